"""schemi_gpu_mixing_width_noalloc (the entry point that allocates nothing:
its reduction's partial sums live in the library's fixed device scratch)
against schemi_gpu_mixing_width, through the C ABI: the same value and the
same gradient, bit for bit, on two calls in a row, the second on a smaller
mesh (so the second call's partial sums are fewer than the first's and any
left over in the shared scratch would show); and the size limit."""
import numpy as np
import pytest

from schemi_amd import Case
from schemi_amd.gpu import GpuSolver, SchemiGpuError

SCRATCH_ERR = 100          # SCHEMI_ERR_SCRATCH
MAX_CELLS = 16777216       # SCHEMI_NOALLOC_MAX_CELLS


def _state(case, steps):
    from schemi_amd import cases

    if case == "rm":
        c, U0 = cases.richtmyer_meshkov_3d(nx=128, ny=64, nz=64, seed=17)
    else:
        c, U0 = cases.rayleigh_taylor_2d(nx=64, ny=128, seed=3)
    s = GpuSolver(c)
    s.set_conserved(U0)
    s.run(steps, 1e-15)
    return c, s


@pytest.mark.gpu
def test_noalloc_mixing_width_matches_allocating_entry(hip_device):
    import torch

    for name, steps in (("rm", 5), ("rt", 3)):  # 524288 cells, then 8192
        case, s = _state(name, steps)
        n = case.ncells
        g1 = torch.full((case.Ns + 5, n), float("nan"), dtype=torch.float64, device="cuda")
        g2 = torch.full((case.Ns + 5, n), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        for sp in range(case.Ns):
            J1 = s.mixing_width(sp, g1.data_ptr())
            J2 = s.mixing_width_noalloc(sp, g2.data_ptr())
            torch.cuda.synchronize()
            assert J1 > 0.0 and J2 == J1, (name, sp, J1, J2)
            assert torch.equal(g1, g2), (name, sp)
            assert s.mixing_width_noalloc(sp) == J1  # without the gradient
        s.close()


@pytest.mark.gpu
def test_noalloc_mixing_width_rejects_meshes_beyond_the_scratch(hip_device):
    case = Case(dim=3, N=(256, 256, 257), L=(1.0, 1.0, 1.0), M=[29.0, 4.0],
                Cv=[2.5, 1 / 0.67])
    assert case.ncells > MAX_CELLS
    s = GpuSolver(case)
    with pytest.raises(SchemiGpuError) as e:
        s.mixing_width_noalloc(0)
    assert e.value.code == SCRATCH_ERR
    s.close()
