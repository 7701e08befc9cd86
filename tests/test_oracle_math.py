"""Closed-form checks of the oracle's building blocks.

These pin the oracle where the golden vectors leave it open: the libstdc++
3-argument hypot, the limiter functions phi(r), the HLL/HLLC flux properties
(consistency, upwinding, exact contact, mirror antisymmetry), conservation of
the full step with reflecting walls, and the SSP-RK2 update.
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle, hypot3, limiter_phi
from schemi_amd import Case, boundaryConditionType as BCT, flowSolver as FS
from schemi_amd import typeOfTVDLimiter as TL
from schemi_amd.gpu import conserved_from_primitive
from tests.cases_small import random_primitives
from tests.golden_io import per_cell_rel

BASE = dict(M=[29.0, 4.0], Cv=[2.5, 1 / 0.67])


def test_hypot3():
    assert hypot3(3.0, 4.0, 12.0) == 13.0
    assert hypot3(0.0, 0.0, 0.0) == 0.0
    assert hypot3(-2.5, 0.0, 0.0) == 2.5
    r = np.random.default_rng(0)
    for x, y, z in r.normal(size=(200, 3)):
        assert abs(hypot3(x, y, z) - math.sqrt(x * x + y * y + z * z)) <= 4e-16 * math.sqrt(
            x * x + y * y + z * z)


@pytest.mark.parametrize("r", [-1.0, -0.3, 0.0, 0.25, 0.5, 1.0, 1.5, 3.0, 10.0])
def test_limiter_closed_forms(r):
    xi = 2 / (1 + r) if r != -1.0 else math.inf
    minmod = 0.0 if r <= 0 else (r if r <= 1 else min(1.0, xi))
    vanleer = 0.0 if r < 0 else min(2 * r / (1 + r), xi)
    vanalbada = 0.0 if r <= 0 else min((r * r + r) / (1 + r * r), xi)
    if r == -1.0:
        return
    assert limiter_phi(TL.minmod, r) == minmod
    assert limiter_phi(TL.vanLeer, r) == vanleer
    assert limiter_phi(TL.vanAlbada, r) == vanalbada
    # TVD region: 0 <= phi <= min(2r, 2) for r > 0 (Sweby diagram)
    for lim in (TL.minmod, TL.vanLeer, TL.vanAlbada):
        ph = limiter_phi(lim, r)
        assert ph >= 0.0
        if r > 0:
            assert ph <= min(2 * r, 2.0) + 1e-15


def _phys_flux(case, N, c, v, p):
    """F(U) & N = (rho_k u, m u + p N, u (E + p)) from primitives."""
    rho_k = [c[k] * case.M[k] for k in range(case.Ns)]
    rho = sum(rho_k)
    cv = sum((c[k] / sum(c)) * case.Cv[k] for k in range(case.Ns))
    U = p / (case.R / cv)
    E = U + rho * np.dot(v, v) / 2
    u = np.dot(v, N)
    # density fluxes in the reference's order: total density first, then species
    return (np.array([rho * u] + [r * u for r in rho_k]),
            np.array([rho * v[a] * u + p * N[a] for a in range(3)]), u * (E + p))


@pytest.mark.parametrize("flux", [FS.HLLC, FS.HLL])
def test_flux_consistency_and_upwinding(flux):
    case = Case(dim=1, N=(4, 1, 1), L=(1.0, 1.0, 1.0), flux=flux, **BASE)
    o = Oracle(case)
    r = np.random.default_rng(1)
    for _ in range(50):
        N = np.array([1.0, 0.0, 0.0]) * (1 if r.random() < 0.5 else -1)
        c = 0.01 + 0.05 * r.random(2)
        v = r.normal(size=3) * 0.5
        p = 0.2 + r.random()
        err, Fr, Fm, FE = o.face_flux(N, c, v, p, c, v, p)
        assert err == 0
        Pr, Pm, PE = _phys_flux(case, N, c, v, p)
        np.testing.assert_allclose(Fr, Pr, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(Fm, Pm, rtol=1e-12, atol=1e-14)
        assert math.isclose(FE, PE, rel_tol=1e-12, abs_tol=1e-14)
    # supersonic to +N: the flux is the left physical flux
    N = np.array([1.0, 0.0, 0.0])
    cL, vL, pL = np.array([0.03, 0.01]), np.array([5.0, 0.1, 0.0]), 1.0
    cR, vR, pR = np.array([0.01, 0.02]), np.array([4.0, -0.2, 0.0]), 0.5
    err, Fr, Fm, FE = o.face_flux(N, cL, vL, pL, cR, vR, pR)
    Pr, Pm, PE = _phys_flux(case, N, cL, vL, pL)
    np.testing.assert_allclose(Fr, Pr, rtol=1e-13)
    np.testing.assert_allclose(Fm, Pm, rtol=1e-13)
    assert math.isclose(FE, PE, rel_tol=1e-13)


def test_hllc_isolated_contact():
    """u_L = u_R, p_L = p_R: HLLC transports the upwind densities exactly."""
    case = Case(dim=1, N=(4, 1, 1), L=(1.0, 1.0, 1.0), flux=FS.HLLC, **BASE)
    o = Oracle(case)
    N = np.array([1.0, 0.0, 0.0])
    v = np.array([0.3, 0.0, 0.0])
    cL, cR = np.array([0.04, 0.0]), np.array([0.0, 0.2])
    err, Fr, Fm, FE = o.face_flux(N, cL, v, 1.0, cR, v, 1.0)
    assert err == 0
    np.testing.assert_allclose(Fr, [0.04 * 29 * 0.3, 0.04 * 29 * 0.3, 0.0], rtol=1e-12,
                               atol=1e-15)
    np.testing.assert_allclose(Fm[0], 0.04 * 29 * 0.3 * 0.3 + 1.0, rtol=1e-12)


@pytest.mark.parametrize("flux", [FS.HLLC, FS.HLL])
def test_flux_mirror_antisymmetry(flux):
    """F(U_L, U_R, N) = -F(U_R', U_L', -N) with normal velocities mirrored."""
    case = Case(dim=1, N=(4, 1, 1), L=(1.0, 1.0, 1.0), flux=flux, **BASE)
    o = Oracle(case)
    r = np.random.default_rng(2)
    N = np.array([1.0, 0.0, 0.0])
    for _ in range(30):
        cL, cR = 0.01 + 0.05 * r.random(2), 0.01 + 0.05 * r.random(2)
        vL, vR = r.normal(size=3) * 0.3, r.normal(size=3) * 0.3
        pL, pR = 0.2 + r.random(), 0.2 + r.random()
        _, F1r, F1m, F1E = o.face_flux(N, cL, vL, pL, cR, vR, pR)
        _, F2r, F2m, F2E = o.face_flux(-N, cR, vR, pR, cL, vL, pL)
        np.testing.assert_allclose(F1r, -F2r, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(F1m, -F2m, rtol=1e-12, atol=1e-14)
        assert math.isclose(F1E, -F2E, rel_tol=1e-12, abs_tol=1e-14)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_step_conserves_with_reflecting_walls(dim):
    N = {1: (50, 1, 1), 2: (20, 16, 1), 3: (12, 8, 6)}[dim]
    case = Case(dim=dim, N=N, L=(1.0, 1.0, 1.0), limiter=TL.vanLeer, **BASE)
    for p in range(6):
        case.set_bc(p, BCT.slip)
    c, v, p = random_primitives(case, 3)
    U0 = conserved_from_primitive(case, c, v, p)
    o = Oracle(case)
    o.set_conserved(U0[:2], U0[2], U0[3:6], U0[6])
    for _ in range(5):
        assert o.step(1e-4, 1) == 0
    U1 = o.conserved()
    for q in (0, 1, 2, 6):  # partial densities, density, energy
        assert math.isclose(U1[q].sum(), U0[q].sum(), rel_tol=1e-13)


def test_rk2_linear_amplification():
    """Uniform flow carrying a small species perturbation: the step is the
    SSP-RK2 combination of two Euler stages (U1 = L(Un), Un+1 = (Un + L(U1))/2)."""
    case = Case(dim=1, N=(32, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.zero, **BASE)
    c, v, p = random_primitives(case, 4)
    U0 = conserved_from_primitive(case, c, v, p)
    a = Oracle(case)
    a.set_conserved(U0[:2], U0[2], U0[3:6], U0[6])
    a.step(1e-3, 1)
    b = Oracle(case)
    b.set_conserved(U0[:2], U0[2], U0[3:6], U0[6])
    b.stage(1e-3)
    b.stage(1e-3)
    U2 = b.conserved()
    expect = 0.5 * U2 + 0.5 * U0
    np.testing.assert_allclose(a.conserved(), expect, rtol=1e-15, atol=0)


def test_vdw_tends_to_ideal_mixture():
    """With the critical pressure raised 1e8-fold at fixed Tc, the van der
    Waals a_kl ~ 1/Pc and b_kl ~ 1/Pc vanish: one step of the van der Waals
    oracle approaches the ideal-mixture step (relative 1e-6), while at the
    test's Pc = 1 the two differ by far more (the model is really engaged)."""
    from schemi_amd import gasModel

    def run(**eos):
        case = Case(dim=1, N=(48, 1, 1), L=(1.0, 1.0, 1.0), M=[29.0, 4.0], Cv=[2.5, 1 / 0.67],
                    **eos)
        o = Oracle(case)
        r = np.random.default_rng(3)
        c = np.stack([0.02 + 0.02 * r.random(48), 0.1 + 0.1 * r.random(48)])
        v = np.zeros((3, 48))
        v[0] = 0.3 * r.standard_normal(48)
        p = 0.8 + 0.4 * r.random(48)
        o.init_primitive(c, v, p)
        assert o.step(2e-3, 0) == 0
        return o.state()

    ideal = run()
    near = run(gas_model=gasModel.vanDerWaals, Tc=[24.0, 3.2], Pc=[1e8, 1e8])
    far = run(gas_model=gasModel.vanDerWaals, Tc=[24.0, 3.2], Pc=[1.0, 1.0])
    for key in ("p", "E", "rho0"):
        scale = np.abs(ideal[key]).max()
        assert np.abs(near[key] - ideal[key]).max() <= 1e-6 * scale, key
    # at equal initial p the van der Waals internal energy differs by ~a c0^2
    assert np.abs(far["E"] - ideal["E"]).max() > 1e-2 * np.abs(ideal["E"]).max()


def test_rk_initial_state_matches_oracle():
    """The host's Redlich-Kwong UvFromp (FP64 Newton root of the cubic in
    sqrt(T)) gives the oracle's conserved state (long-double Cardano, the
    reference's solver) to ~1e-15, and the mixing matrices agree exactly."""
    from schemi_amd import Case, gasModel as GM
    from schemi_amd.gpu import conserved_from_primitive, rk_matrices
    from oracle.oracle import Oracle

    case = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), M=[29.0, 4.0], Cv=[2.5, 1 / 0.67],
                gas_model=GM.RedlichKwong, Tc=[24.0, 3.2], Pc=[1.0, 1.0])
    c, v, p = random_primitives(case, 9)
    U = conserved_from_primitive(case, c, v, p)
    o = Oracle(case)
    o.init_primitive(c, v, p)
    ref = o.conserved()
    for q, err in enumerate(per_cell_rel(U, ref, case.Ns)):
        assert err <= 2e-15, (q, err)
    a, b = rk_matrices(case)
    assert a[0, 0] > 1000 and 2.0 < b[0, 0] < 2.2  # Tc 24, Pc 1: strongly non-ideal


def test_stiffened_initial_state_matches_oracle():
    """The host's stiffened UvFromp gives the oracle's conserved state bit for
    bit (closed form, same operation order)."""
    from schemi_amd import Case, gasModel as GM
    from schemi_amd.gpu import conserved_from_primitive
    from oracle.oracle import Oracle

    case = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), M=[29.0, 4.0], Cv=[2.5, 1 / 0.67],
                gas_model=GM.stiffened, p0=[0.3, 0.05], gammas=[2.8, 1.6])
    c, v, p = random_primitives(case, 9)
    U = conserved_from_primitive(case, c, v, p)
    o = Oracle(case)
    o.init_primitive(c, v, p)
    ref = o.conserved()
    for q, err in enumerate(per_cell_rel(U, ref, case.Ns)):
        assert err <= 1e-15, (q, err)


KATAOKA = dict(Tc=[24.0, 3.2], Pc=[1.0, 1.0], epsLJ=[10.0, 5.0], sigmaLJ=[1.52e-8, 1.2e-8])


@pytest.mark.parametrize("kb", [("critical", 1.0), ("Kataoka", 2.0)])
def test_kataoka_initial_state_matches_oracle(kb):
    """The host's Kataoka-van der Waals UvFromp (numpy Newton on T with the
    reference's tolerance) gives the oracle's conserved state to the Newton
    tolerance, and the host's mixing matrices are the oracle's."""
    from schemi_amd import Case, gasModel as GM
    from schemi_amd.gpu import conserved_from_primitive, kataoka_matrices
    from oracle.oracle import Oracle

    case = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), M=[29.0, 4.0], Cv=[2.5, 1 / 0.67],
                gas_model=GM.KataokaVanDerWaals, kataoka_b=kb, **KATAOKA)
    c, v, p = random_primitives(case, 9)
    U = conserved_from_primitive(case, c, v, p)
    o = Oracle(case)
    o.init_primitive(c, v, p)
    ref = o.conserved()
    for q, err in enumerate(per_cell_rel(U, ref, case.Ns)):
        assert err <= 1e-12, (q, err)
    V0, eps, b = kataoka_matrices(case)
    # V0_kk = sigma^3 N_A / sqrt(2); eps geometric mean; b from Tc/Pc or sigma
    assert np.isclose(V0[0, 0], 1.52e-8 ** 3 * 6.02214076e23 / np.sqrt(2), rtol=1e-15)
    assert np.isclose(eps[0, 1], np.sqrt(50.0), rtol=1e-15)
    if kb[0] == "critical":
        assert np.isclose(b[1, 1], case.R * 3.2 / 8.0, rtol=1e-15)
    else:
        assert np.isclose(b[1, 1], 2.0 * 1.2e-8 ** 3 * 6.02214076e23, rtol=1e-15)
    # mixing rule across species: ((cbrt b_00 + cbrt b_11) / 2)^3
    assert np.isclose(b[0, 1], (0.5 * (np.cbrt(b[0, 0]) + np.cbrt(b[1, 1]))) ** 3, rtol=1e-14)


@pytest.mark.parametrize("model", ["ideal", "vdw", "rk", "stiffened", "kataoka"])
def test_eos_round_trip(model):
    """p -> U (UvFromp at initialisation) -> p (pFromUv of the conserved
    state) returns the pressure to a few ulps for every mixture model, and the
    temperature of the conserved state equals the initial one."""
    from schemi_amd import Case, gasModel as GM
    from oracle.oracle import Oracle

    eos = {"ideal": {}, "vdw": dict(gas_model=GM.vanDerWaals, Tc=[24.0, 3.2], Pc=[1.0, 1.0]),
           "rk": dict(gas_model=GM.RedlichKwong, Tc=[24.0, 3.2], Pc=[1.0, 1.0]),
           "stiffened": dict(gas_model=GM.stiffened, p0=[0.3, 0.05], gammas=[2.8, 1.6]),
           "kataoka": dict(gas_model=GM.KataokaVanDerWaals, **KATAOKA)}[model]
    case = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), M=[29.0, 4.0], Cv=[2.5, 1 / 0.67], **eos)
    c, v, p = random_primitives(case, 13)
    o = Oracle(case)
    o.init_primitive(c, v, p)
    st0 = o.state()
    U = o.conserved()
    Ns = case.Ns
    o2 = Oracle(case)
    o2.set_conserved(U[:Ns], U[Ns], U[Ns + 1:Ns + 4], U[Ns + 4])
    st = o2.state()
    # Kataoka-van der Waals: both directions stop their Newton iteration at
    # |T_{k+1} - T_k| <= 1e-12 (absolute), so the round trip holds to that
    tol = 1e-11 if model == "kataoka" else 1e-13
    assert np.max(np.abs(st["p"] - p) / np.abs(p)) < tol
    assert np.max(np.abs(st["T"] - st0["T"]) / np.abs(st0["T"])) < tol
