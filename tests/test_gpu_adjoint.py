"""Adjoint of the GPU step (schemi_gpu_adjoint_step), the mixing-width
functional (schemi_gpu_mixing_width) and gradients of the Rayleigh-Taylor
mixing width with respect to the initial interface-mode amplitudes, checked
against central finite differences of the GPU FP64 solver and of the oracle
(which reproduces the reference CPU solver bit for bit)."""
import numpy as np
import pytest

from schemi_amd import errors
from schemi_amd.gpu import PREC_FP64, PREC_MIXED, GpuSolver, SchemiGpuError
from schemi_amd.gpu import conserved_from_primitive
from tests.cases_small import cases, random_primitives

torch = pytest.importorskip("torch")

ADJ_CASES = ["1d_zero", "2d_hll_grav", "2d_periodic_hll", "2d_zero_fixed", "2d_vdw_periodic_hll",
             "3d_vanAlbada", "3d_periodic_x_vanLeer", "3d_hll_partial_tiles", "3d_stiffened_hll",
             # 1-D with limited slopes: limiter branches frozen as well
             "1d_minmod", "1d_vanLeer", "1d_vanAlbada", "1d_linear", "1d_hll", "1d_slip",
             "1d_fixed", "1d_periodic", "1d_vdw_vanLeer", "1d_stiffened_minmod"]


def _dev(x):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda")


@pytest.mark.gpu
@pytest.mark.parametrize("what", [0, 1], ids=["step", "stage"])
@pytest.mark.parametrize("name", ADJ_CASES)
def test_adjoint_dot_product(name, what, hip_device):
    """<lam, (S(U + h d) - S(U - h d)) / 2h> = <S'(U)^T lam, d> for the SSP-RK2
    step (and one stage), lam random everywhere, d random (scaled per field)
    on the first and last cell and 62 random ones (half the cells of a
    smaller mesh).  The support
    keeps the faces whose frozen branches could switch within +-h d (a kink
    of min/max or of the HLLC wave pattern) to a few hundred; the FP64
    central difference with h = 1e-6 is then accurate to ~1e-9."""
    case = cases()[name]
    c, v, p = random_primitives(case, 3)
    U = conserved_from_primitive(case, c, v, p)
    rng = np.random.default_rng(1)
    n = case.ncells
    cells = np.unique(np.concatenate([[0, n - 1], rng.choice(n, min(62, n // 2), replace=False)]))
    d = np.zeros_like(U)
    d[:, cells] = rng.standard_normal((U.shape[0], cells.size))
    d *= np.abs(U).max(axis=1, keepdims=True)
    lam = rng.standard_normal(U.shape)
    dt = 2e-4
    s = GpuSolver(case, PREC_FP64)

    def fwd(X):
        s.set_conserved(X)
        if what == 0:
            s.step(dt)
        else:
            s.stage(dt)
        return s.conserved()

    h = 1e-6
    fd = float(np.sum(lam * (fwd(U + h * d) - fwd(U - h * d)))) / (2 * h)
    s.set_conserved(U)
    lt = _dev(lam)
    out = torch.empty_like(lt)
    torch.cuda.synchronize()
    s.adjoint_step(dt, lt.data_ptr(), out.data_ptr(), what)
    ad = float(np.sum(out.cpu().numpy() * d))
    print(name, what, "adjoint", ad, "fd", fd)
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)
    # the state is left unchanged, and in-place use gives the same bits
    assert np.array_equal(s.conserved(), U)
    out2 = lt.clone()
    torch.cuda.synchronize()
    s.adjoint_step(dt, out2.data_ptr(), out2.data_ptr(), what)
    assert torch.equal(out, out2)
    s.close()


@pytest.mark.gpu
def test_adjoint_unsupported(hip_device):
    """No adjoint for the linear limiter in 2-D/3-D, MUSCL-Hancock or the
    Redlich-Kwong mixture."""
    for name in ("2d_linear", "3d_linear_dyadic", "hancock_2d_vanLeer", "1d_rk_vanLeer"):
        case = cases()[name]
        s = GpuSolver(case)
        c, v, p = random_primitives(case, 3)
        s.set_conserved(conserved_from_primitive(case, c, v, p))
        lt = torch.zeros((case.Ns + 5, case.ncells), dtype=torch.float64, device="cuda")
        with pytest.raises(SchemiGpuError) as e:
            s.adjoint_step(1e-4, lt.data_ptr(), lt.data_ptr(), 0)
        assert e.value.code == errors.initialisationError
        s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["2d_hll_grav", "3d_periodic_x_vanLeer"])
def test_mixing_width_closed_form(name, hip_device):
    """J = sum_i V_i Y (1 - Y) with Y = rho_1 / rho, and dJ/drho_1 =
    V (1 - 2Y) / rho, dJ/drho = -V (1 - 2Y) Y / rho, zero elsewhere; cell
    volumes from the oracle's mesh."""
    from oracle.oracle import Oracle

    case = cases()[name]
    c, v, p = random_primitives(case, 4)
    U = conserved_from_primitive(case, c, v, p)
    _, V = Oracle(case).cell_geometry()
    Ns = case.Ns
    Y = U[0] / U[Ns]
    s = GpuSolver(case)
    s.set_conserved(U)
    g = torch.empty((Ns + 5, case.ncells), dtype=torch.float64, device="cuda")
    J = s.mixing_width(0, g.data_ptr())
    assert np.isclose(J, np.sum(V * Y * (1 - Y)), rtol=1e-13)
    G = g.cpu().numpy()
    ref = np.zeros_like(U)
    ref[0] = V * (1 - 2 * Y) / U[Ns]
    ref[Ns] = -V * (1 - 2 * Y) * Y / U[Ns]
    assert np.allclose(G, ref, rtol=1e-14, atol=0)
    assert s.mixing_width(0) == J  # deterministic
    s.close()


def _oracle_J(case, U0, nsteps, dt):
    from oracle.oracle import Oracle

    Ns = case.Ns
    o = Oracle(case)
    o.set_conserved(U0[:Ns], U0[Ns], U0[Ns + 1:Ns + 4], U0[Ns + 4])
    for _ in range(nsteps):
        assert o.step(dt, 1) == 0
    U = o.conserved()
    _, V = o.cell_geometry()
    Y = U[0] / U[Ns]
    return float(np.sum(V * Y * (1 - Y)))


@pytest.mark.gpu
def test_rt_mixing_width_gradient_small(hip_device):
    """dJ/da_k of the Rayleigh-Taylor mixing width after 40 steps (64 x 128,
    4 modes, checkpoints every 10 steps): adjoint vs central differences
    (h = 1e-7) of the GPU FP64 solver and of the oracle, <= 1e-4 relative."""
    from schemi_amd import adjoint as adj

    case = adj.rt_case(64, 128)
    a = np.array([0.004, 0.0025, 0.003, 0.0015])
    nsteps, dt = 40, 1e-3
    J, g = adj.mixing_width_gradient(case, a, nsteps, dt, every=10)
    h = 1e-7
    for k in range(len(a)):
        e = np.zeros_like(a)
        e[k] = h
        fd = (adj.mixing_width_forward(case, a + e, nsteps, dt) -
              adj.mixing_width_forward(case, a - e, nsteps, dt)) / (2 * h)
        print("mode", k + 1, "adjoint", g[k], "fd(gpu)", fd)
        assert abs(g[k] - fd) <= 1e-4 * abs(fd)
    # the oracle (the reference's arithmetic) on the same initial states
    J0 = adj.mixing_width_forward(case, a, nsteps, dt)
    assert J == J0
    Uc = {}
    for sgn in (1, -1):
        e = np.zeros_like(a)
        e[0] = sgn * h
        Uc[sgn] = adj.rt_modes_state(case, torch.as_tensor(a + e)).numpy()
    fdo = (_oracle_J(case, Uc[1], nsteps, dt) - _oracle_J(case, Uc[-1], nsteps, dt)) / (2 * h)
    print("mode 1 fd(oracle)", fdo)
    assert abs(g[0] - fdo) <= 1e-4 * abs(fdo)


@pytest.mark.gpu
def test_rt_mixing_width_gradient_full(hip_device):
    """The gradient workload at full size: 256 x 1024, 16 modes, 400 steps
    with checkpoints every 20 steps, FP64 adjoint, linearised along the FP64
    and along the mixed-precision forward.  Four modes against central
    differences of the GPU FP64 solver, <= 1e-4 relative.  The differences
    use h = 1e-8: the interface is two cells wide, so J is strongly curved in
    a and h = 1e-6 leaves a truncation error of 1e-4 to 1e-3 (measured: the
    differences converge to the adjoint as h decreases, 7e-5 at 1e-7 and
    7e-7 at 1e-8 for mode 1), while FP64 rounding in J stays near 1e-7
    relative at h = 1e-8.  The same four modes against the committed central
    differences of the oracle (tests/golden/rt_fd_oracle_256x1024.json,
    written by tests/golden/make_rt_fd.py, h = 1e-8), <= 1e-4 relative."""
    import json
    import os

    from schemi_amd import adjoint as adj

    with open(os.path.join(os.path.dirname(__file__), "golden",
                           "rt_fd_oracle_256x1024.json")) as f:
        ofd = json.load(f)

    case = adj.rt_case(256, 1024)
    rng = np.random.default_rng(2024)
    a = rng.uniform(0.0, 0.005, 16)
    nsteps, dt = 400, 2.5e-4
    assert (ofd["seed"], ofd["nsteps"], ofd["dt"]) == (2024, nsteps, dt)
    import time

    t0 = time.perf_counter()
    J64, g64 = adj.mixing_width_gradient(case, a, nsteps, dt, every=20, precision=PREC_FP64)
    t1 = time.perf_counter()
    Jmx, gmx = adj.mixing_width_gradient(case, a, nsteps, dt, every=20, precision=PREC_MIXED)
    t2 = time.perf_counter()
    print("J fp64", J64, "J mixed", Jmx, "gradient wall s: fp64 fwd", t1 - t0, "mixed fwd", t2 - t1)
    h = 1e-8
    for k in (0, 3, 7, 15):
        e = np.zeros_like(a)
        e[k] = h
        fd = (adj.mixing_width_forward(case, a + e, nsteps, dt) -
              adj.mixing_width_forward(case, a - e, nsteps, dt)) / (2 * h)
        r64 = abs(g64[k] - fd) / abs(fd)
        rmx = abs(gmx[k] - fd) / abs(fd)
        print("mode", k + 1, "fd(gpu fp64)", fd, "adjoint(fp64 fwd)", g64[k], "rel", r64,
              "adjoint(mixed fwd)", gmx[k], "rel", rmx)
        assert r64 <= 1e-4
        assert rmx <= 1e-4
        fdo = ofd["fd"][f"{k}:1e-08"]
        o64 = abs(g64[k] - fdo) / abs(fdo)
        omx = abs(gmx[k] - fdo) / abs(fdo)
        print("mode", k + 1, "fd(oracle)", fdo, "rel fp64", o64, "rel mixed", omx)
        assert o64 <= 1e-4
        assert omx <= 1e-4
