"""Periodic boundaries of the oracle (no reference counterpart).

Pinned by properties: a periodic step conserves every conserved sum exactly
up to round-off, it commutes with a cyclic shift of the state along the
periodic axis, and a state mirror-symmetric about the domain boundary gives
the same step as the reference's slip walls (the mirror surrogate).
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from schemi_amd import Case, boundaryConditionType as BCT, typeOfTVDLimiter as TL
from schemi_amd.gpu import conserved_from_primitive, periodic
from tests.cases_small import random_primitives

BASE = dict(M=[29.0, 4.0], Cv=[2.5, 1 / 0.67])


def run(case, U0, steps=3, dt=2e-4):
    o = Oracle(case)
    o.set_conserved(U0[:2], U0[2], U0[3:6], U0[6])
    for _ in range(steps):
        assert o.step(dt, 1) == 0
    return o.conserved()


@pytest.mark.parametrize("lim", [TL.minmod, TL.vanLeer, TL.linear])
def test_periodic_conservation_and_shift(lim):
    case = periodic(Case(dim=1, N=(48, 1, 1), L=(1.0, 1.0, 1.0), limiter=lim, **BASE), 0)
    c, v, p = random_primitives(case, 21)
    U0 = conserved_from_primitive(case, c, v, p)
    U1 = run(case, U0)
    for q in range(U0.shape[0]):
        assert abs(U1[q].sum() - U0[q].sum()) <= 1e-13 * np.abs(U0[q]).sum()
    # shift invariance (uniform dyadic mesh: every cell sees the same geometry)
    case2 = periodic(Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=lim, **BASE), 0)
    c, v, p = random_primitives(case2, 22)
    U0 = conserved_from_primitive(case2, c, v, p)
    a = run(case2, U0)
    b = run(case2, np.roll(U0, 5, axis=1))
    assert np.array_equal(np.roll(a, 5, axis=1), b)


@pytest.mark.parametrize("lim", [TL.zero, TL.linear])
def test_periodic_matches_mirror_surrogate(lim):
    """Mirror-symmetric data about x = 0 and x = L: periodic == slip walls on
    the unmirrored half.  Holds for limiters whose slopes are
    reflection-symmetric (zero, linear; every TVD limiter in multi-D, where
    phi = 0).  The reference's 1-D TVD ratio min/(max + eps) is not symmetric
    under reflection, so there the two set-ups differ from the second stage
    on (they agree bit for bit after one stage)."""
    n = 32
    base = Case(dim=1, N=(n, 1, 1), L=(1.0, 1.0, 1.0), limiter=lim, **BASE)
    c, v, p = random_primitives(base, 23)
    cp = np.concatenate([c, c[:, ::-1]], axis=1)
    vp = np.concatenate([v, -v[:, ::-1]], axis=1)
    pp = np.concatenate([p, p[::-1]])
    per = periodic(Case(dim=1, N=(2 * n, 1, 1), L=(2.0, 1.0, 1.0), limiter=lim, **BASE), 0)
    Up = conserved_from_primitive(per, cp, vp, pp)
    slip = Case(dim=1, N=(n, 1, 1), L=(1.0, 1.0, 1.0), limiter=lim, **BASE)
    slip.set_bc(0, BCT.slip)
    slip.set_bc(1, BCT.slip)
    Us = conserved_from_primitive(slip, c, v, p)
    a = run(per, Up, steps=3)
    b = run(slip, Us, steps=3)
    np.testing.assert_allclose(a[:, :n], b, rtol=1e-12, atol=1e-14)


def test_periodic_matches_mirror_surrogate_2d():
    """2-D van Leer (phi = 0 off the 1-D case): x-periodic mirror vs slip."""
    nx, ny = 16, 8
    base = Case(dim=2, N=(nx, ny, 1), L=(1.0, 0.5, 0.1), limiter=TL.vanLeer, **BASE)
    for pt in (3, 4):
        base.set_bc(pt, BCT.slip)
    c, v, p = random_primitives(base, 24)
    r = lambda a: a.reshape(a.shape[:-1] + (ny, nx))
    mir = lambda a: np.concatenate([r(a), r(a)[..., ::-1]], axis=-1).reshape(a.shape[:-1] + (-1,))
    vm = np.concatenate([r(v), r(v)[..., ::-1] * np.array([-1, 1, 1])[:, None, None]],
                        axis=-1).reshape(3, -1)
    per = Case(dim=2, N=(2 * nx, ny, 1), L=(2.0, 0.5, 0.1), limiter=TL.vanLeer, **BASE)
    periodic(per, 0)
    for pt in (3, 4):
        per.set_bc(pt, BCT.slip)
    Up = conserved_from_primitive(per, mir(c), vm, mir(p[None])[0])
    for pt in (0, 1):
        base.set_bc(pt, BCT.slip)
    Us = conserved_from_primitive(base, c, v, p)
    a = run(per, Up, steps=3).reshape(7, ny, 2 * nx)[:, :, :nx].reshape(7, -1)
    b = run(base, Us, steps=3)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("solver", ["explicit", "Jacobi"])
def test_periodic_diffusion_conservation_and_shift(solver):
    """The diffusion stage on a periodic box: the wrapped face is an inner
    face, so the stage's face fluxes cancel in pairs and the step conserves
    the species masses, the total mass and the momentum to round-off, and
    commutes with a cyclic shift of the state (bit for bit with the explicit
    solver; the Jacobi rows sum their triangles by cell index, which a shift
    reorders, so there to 1e-13)."""
    kw = dict(diffusion=True, transport=(0.02, 0.1, 0.01), sle_solver=solver, **BASE)
    case = Case(dim=2, N=(32, 16, 1), L=(1.0, 0.5, 0.1), limiter=TL.vanLeer, **kw)
    periodic(case, 0)
    periodic(case, 1)
    c, v, p = random_primitives(case, 24)
    U0 = conserved_from_primitive(case, c, v, p)
    U1 = run(case, U0)
    for q in range(6):  # rho_k, rho, momentum
        assert abs(U1[q].sum() - U0[q].sum()) <= 2e-13 * np.abs(U0[q]).sum(), q
    # the stage acts: it differs from the step without diffusion
    plain = Case(dim=2, N=(32, 16, 1), L=(1.0, 0.5, 0.1), limiter=TL.vanLeer, **BASE)
    periodic(plain, 0)
    periodic(plain, 1)
    assert np.max(np.abs(run(plain, U0) - U1)) > 1e-6
    # cyclic shift along x and y
    sh = lambda U: np.roll(U.reshape(U.shape[0], 16, 32), (3, 5), axis=(1, 2)).reshape(U.shape)
    a = run(case, U0)
    b = run(case, sh(U0))
    if solver == "explicit":
        assert np.array_equal(sh(a), b)
    else:
        np.testing.assert_allclose(sh(a), b, rtol=1e-13, atol=0)
