"""The lean mixed-precision stage kernel (ideal mixture, SSP-RK2, slopes that
vanish, free/slip/periodic walls, whole 64 x 8 tiles).

It replaces stage0_kernel on the mixed path wherever it applies, so every
case here (a) stays within the mixed path's tolerance of the FP64 path,
(b) agrees with the general mixed kernel (SCHEMI_GPU_LEAN=0) to the same
tolerance while differing from it in the last bits (the lean kernel ran),
and (c) keeps sum_k rho_k = rho.  Tolerance: relative L1 of every conserved
field <= 1e-5 after 5 steps (FP32 closure and Riemann solve, FP64
accumulation and update)."""
import os

import numpy as np
import pytest

from schemi_amd import Case, boundaryConditionType as BCT, errors
from schemi_amd import flowSolver as FS, typeOfTVDLimiter as TL
from schemi_amd.gpu import (PREC_FP64, PREC_MIXED, GpuSolver, SchemiGpuError,
                            conserved_from_primitive, periodic)
from tests.cases_small import random_primitives


def _lean_cases():
    out = {}
    c = Case(dim=3, N=(64, 16, 16), L=(2.0, 1.0, 1.0), limiter=TL.vanAlbada,
             M=[29.0, 146.0], Cv=[2.5, 1 / 0.09])
    periodic(c, 1)
    periodic(c, 2)
    out["3d_periodic_yz"] = c
    c = Case(dim=3, N=(64, 16, 16), L=(2.0, 1.0, 1.0), limiter=TL.minmod,
             M=[29.0, 4.0], Cv=[2.5, 1 / 0.67])
    for pt in range(6):
        c.set_bc(pt, BCT.slip)
    out["3d_slip"] = c
    c = Case(dim=3, N=(128, 8, 8), L=(1.0, 0.5, 0.5), flux=FS.HLL, limiter=TL.vanLeer,
             M=[29.0, 4.0], Cv=[2.5, 1 / 0.67])
    out["3d_hll_free"] = c
    # Lz < Ly: the top face lies below the last cell centre, its normal is -z
    c = Case(dim=3, N=(64, 8, 8), L=(1.0, 0.5, 0.25), limiter=TL.vanAlbada,
             M=[29.0, 4.0], Cv=[2.5, 1 / 0.67])
    out["3d_lz_flip"] = c
    c = Case(dim=2, N=(64, 64, 1), L=(0.25, 1.0, 1.0 / 64), limiter=TL.vanLeer,
             M=[87.0, 29.0], Cv=[2.5, 2.5], gravity=True, g=(0.0, -1.0, 0.0))
    periodic(c, 0)
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    out["2d_rt_periodic_x"] = c
    c = Case(dim=2, N=(128, 32, 1), L=(1.0, 1.0, 0.1), flux=FS.HLL, limiter=TL.minmod,
             M=[44.0, 28.0], Cv=[1 / 0.3, 2.5])
    periodic(c, 0)
    periodic(c, 1)
    out["2d_hll_periodic"] = c
    c = Case(dim=3, N=(64, 8, 16), L=(1.0, 0.5, 0.5), limiter=TL.zero, M=[29.0], Cv=[2.5])
    periodic(c, 2)
    out["3d_one_species"] = c
    c = Case(dim=3, N=(64, 8, 8), L=(1.0, 0.5, 0.5), limiter=TL.vanAlbada,
             M=[29.0, 4.0, 44.0], Cv=[2.5, 1 / 0.67, 1 / 0.3])
    out["3d_three_species"] = c
    return out


def _run(case, U0, prec, nsteps, dt, lean=True):
    old = os.environ.get("SCHEMI_GPU_LEAN")
    if not lean:
        os.environ["SCHEMI_GPU_LEAN"] = "0"
    try:
        s = GpuSolver(case, prec, const_dt=dt)
    finally:
        if old is None:
            os.environ.pop("SCHEMI_GPU_LEAN", None)
        else:
            os.environ["SCHEMI_GPU_LEAN"] = old
    s.set_conserved(U0)
    s.run(nsteps, dt)
    U = s.conserved()
    s.close()
    return U


def _l1(a, b):
    return np.abs(a - b).sum(axis=1) / np.maximum(np.abs(b).sum(axis=1), 1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(_lean_cases()))
def test_lean_kernel_against_fp64_and_general_mixed(name, hip_device):
    case = _lean_cases()[name]
    c, v, p = random_primitives(case, 21)
    U0 = conserved_from_primitive(case, c, v, p)
    dt = 2e-4
    U64 = _run(case, U0, PREC_FP64, 5, dt)
    Ulean = _run(case, U0, PREC_MIXED, 5, dt)
    Ugen = _run(case, U0, PREC_MIXED, 5, dt, lean=False)
    Ns = case.Ns
    # momentum components that are identically zero stay zero: compare the
    # fields with a nonzero reference only
    live = np.abs(U64).sum(axis=1) > 0
    assert np.all(_l1(Ulean, U64)[live] < 1e-5), _l1(Ulean, U64)
    assert np.all(_l1(Ulean, Ugen)[live] < 1e-5), _l1(Ulean, Ugen)
    assert not np.array_equal(Ulean, Ugen), "the lean kernel did not run"
    assert np.max(np.abs(Ulean[:Ns].sum(0) - Ulean[Ns]) / Ulean[Ns]) < 1e-14


@pytest.mark.gpu
def test_lean_kernel_conserves_on_a_periodic_box(hip_device):
    """Each face flux is evaluated once and added to one cell and subtracted
    from the other in FP64, so on a fully periodic box the totals of rho, m
    and E are conserved to FP64 round-off (the species renormalisation moves
    the partial densities by rounding amounts only)."""
    case = Case(dim=3, N=(64, 16, 16), L=(1.0, 0.5, 0.5), limiter=TL.vanAlbada,
                M=[29.0, 4.0], Cv=[2.5, 1 / 0.67])
    for a in range(3):
        periodic(case, a)
    c, v, p = random_primitives(case, 5, amp_v=0.4)
    U0 = conserved_from_primitive(case, c, v, p)
    U = _run(case, U0, PREC_MIXED, 20, 2e-4)
    Ns = case.Ns
    for q in range(Ns, Ns + 5):
        scale = np.abs(U0[q]).sum()
        assert abs(U[q].sum() - U0[q].sum()) <= 1e-12 * scale, q
    for q in range(Ns):
        assert abs(U[q].sum() - U0[q].sum()) <= 1e-9 * np.abs(U0[q]).sum(), q


@pytest.mark.gpu
def test_lean_kernel_reports_riemann_error(hip_device):
    case = Case(dim=3, N=(64, 8, 8), L=(1.0, 0.5, 0.5), limiter=TL.vanAlbada,
                M=[29.0, 4.0], Cv=[2.5, 1 / 0.67])
    c, v, p = random_primitives(case, 3)
    p[100] = -1.0  # negative pressure: the sound speed is NaN
    U0 = conserved_from_primitive(case, c, v, p)
    s = GpuSolver(case, PREC_MIXED)
    s.set_conserved(U0)
    with pytest.raises(SchemiGpuError) as e:
        s.step(1e-4)
    assert e.value.code == errors.RiemannSolverError
