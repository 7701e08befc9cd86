"""Host-side pieces of the gradient workload (no GPU): the differentiable
Rayleigh-Taylor initial state matches the numpy generator of the same set-up,
and the new entry points are exported."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def test_rt_modes_state_matches_generator():
    from schemi_amd import adjoint as adj
    from schemi_amd import cases

    a = np.array([0.004, 0.0025, 0.003, 0.0015])
    case, U_np = cases.rayleigh_taylor_2d(64, 128, modes=list(a))
    U_t = adj.rt_modes_state(adj.rt_case(64, 128), torch.as_tensor(a)).numpy()
    assert U_t.shape == U_np.shape
    for q in range(U_np.shape[0]):
        scale = max(np.abs(U_np[q]).max(), 1e-300)
        assert np.abs(U_t[q] - U_np[q]).max() <= 1e-13 * scale, q


def test_rt_modes_state_is_differentiable():
    from schemi_amd import adjoint as adj

    case = adj.rt_case(16, 32)
    a = torch.tensor([0.01, 0.005], dtype=torch.float64, requires_grad=True)
    U = adj.rt_modes_state(case, a)
    (g,) = torch.autograd.grad(U[0].sum(), a)
    h = 1e-5
    for k in range(2):
        e = torch.zeros(2, dtype=torch.float64)
        e[k] = h
        # difference per cell first: the sums are large, the change is small
        fd = (adj.rt_modes_state(case, a.detach() + e)[0] -
              adj.rt_modes_state(case, a.detach() - e)[0]).sum() / (2 * h)
        assert abs(float(g[k]) - float(fd)) <= 1e-4 * abs(float(fd))
