"""Small parity cases shared by the CPU and GPU tests (seeded, deterministic)."""
import copy

import numpy as np

from schemi_amd import Case, typeOfTVDLimiter as TL, flowSolver as FS, gasModel as GM
from schemi_amd import boundaryConditionType as BCT
from schemi_amd.gpu import periodic

BASE = dict(M=[29.0, 4.0], Cv=[2.5, 1 / 0.67])


def random_primitives(case, seed, amp_v=0.2):
    r = np.random.default_rng(seed)
    n = case.ncells
    c = np.stack([0.02 + 0.03 * r.random(n) for _ in range(case.Ns)])
    v = amp_v * (r.random((3, n)) - 0.5)
    if case.dim < 3:
        v[2] = 0
    if case.dim < 2:
        v[1] = 0
    p = 0.5 + r.random(n)
    return c, v, p


def cases():
    out = {}
    for lim in (TL.minmod, TL.vanLeer, TL.vanAlbada, TL.linear, TL.zero):
        out[f"1d_{lim.name}"] = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0), limiter=lim, **BASE)
    out["1d_hll"] = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0), flux=FS.HLL, **BASE)
    c = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0), **BASE)
    c.set_bc(0, BCT.slip)
    c.set_bc(1, BCT.slip)
    out["1d_slip"] = c
    c = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0), **BASE)
    c.set_bc(0, BCT.fixedValueCell, "p", (1.2, 0, 0))
    c.set_bc(0, BCT.fixedValueCell, "v", (0.1, 0.0, 0.0))
    out["1d_fixed"] = c
    for lim in (TL.vanLeer, TL.linear):
        c = Case(dim=2, N=(40, 32, 1), L=(0.25, 1.0, 0.1), limiter=lim, **BASE)
        c.set_bc(3, BCT.slip)
        c.set_bc(4, BCT.slip)
        out[f"2d_{lim.name}"] = c
    c = Case(dim=2, N=(36, 20, 1), L=(0.3, 0.7, 0.1), limiter=TL.minmod, flux=FS.HLL,
             gravity=True, g=(0.0, -1.0, 0.0), **BASE)
    out["2d_hll_grav"] = c
    for lim in (TL.vanAlbada, TL.linear):
        c = Case(dim=3, N=(40, 16, 16), L=(2.0, 1.0, 1.0), limiter=lim, **BASE)
        for pt in (2, 3, 4, 5):
            c.set_bc(pt, BCT.slip)
        out[f"3d_{lim.name}"] = c
    c = Case(dim=3, N=(33, 9, 18), L=(1.0, 0.5, 0.5), limiter=TL.minmod, flux=FS.HLL, **BASE)
    out["3d_hll_partial_tiles"] = c
    # periodic boundaries (no reference counterpart; the oracle pairs the faces)
    c = Case(dim=1, N=(256, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer, **BASE)
    periodic(c, 0)
    out["1d_periodic"] = c
    c = Case(dim=2, N=(64, 32, 1), L=(1.0, 1.0, 0.1), limiter=TL.minmod, flux=FS.HLL, **BASE)
    periodic(c, 0)
    periodic(c, 1)
    out["2d_periodic_hll"] = c
    c = Case(dim=3, N=(24, 16, 20), L=(1.5, 1.0, 1.0), limiter=TL.linear, **BASE)
    periodic(c, 1)
    periodic(c, 2)
    out["3d_periodic_yz_linear"] = c
    # the linear limiter on separable (dyadic) meshes uses the fused slope
    # kernel; the non-dyadic ones above use the general-geometry kernels
    c = Case(dim=2, N=(32, 32, 1), L=(0.25, 1.0, 0.125), limiter=TL.linear, flux=FS.HLL, **BASE)
    c.set_bc(3, BCT.slip)
    out["2d_linear_dyadic"] = c
    c = Case(dim=3, N=(32, 16, 16), L=(2.0, 1.0, 1.0), limiter=TL.linear, **BASE)
    for pt in (2, 3, 4, 5):
        c.set_bc(pt, BCT.slip)
    out["3d_linear_dyadic"] = c
    c = Case(dim=2, N=(24, 16, 1), L=(0.7, 0.5, 0.1), limiter=TL.linear, **BASE)
    periodic(c, 1)
    out["2d_periodic_y_linear"] = c
    c = Case(dim=2, N=(40, 24, 1), L=(0.3, 0.7, 0.1), limiter=TL.zero, **BASE)
    c.set_bc(0, BCT.fixedValueCell, "p", (1.2, 0, 0))
    out["2d_zero_fixed"] = c
    c = Case(dim=3, N=(64, 16, 8), L=(1.0, 0.5, 0.25), limiter=TL.vanLeer, flux=FS.HLL, **BASE)
    periodic(c, 0)
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    out["3d_periodic_x_vanLeer"] = c
    # Lz != Ly in 3-D (z spaced by Ly / Nz): reversed top face normal (Lz < Ly)
    # for the fused linear kernel, and a top face above the last cell (Lz > Ly)
    # for the general-geometry kernels
    c = Case(dim=3, N=(32, 8, 8), L=(1.0, 0.25, 0.125), limiter=TL.linear, **BASE)
    c.set_bc(5, BCT.slip)
    out["3d_lz_short_linear_dyadic"] = c
    c = Case(dim=3, N=(10, 6, 6), L=(1.0, 0.6, 0.65), limiter=TL.linear, flux=FS.HLL, **BASE)
    out["3d_lz_long_linear"] = c
    # van der Waals mixture (c b ~ 0.1, a c^2 ~ p / 4 for these states)
    vdw = dict(gas_model=GM.vanDerWaals, Tc=[24.0, 3.2], Pc=[1.0, 1.0])
    out["1d_vdw_vanLeer"] = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer,
                                 **BASE, **vdw)
    c = Case(dim=2, N=(64, 32, 1), L=(1.0, 0.5, 0.1), limiter=TL.minmod, flux=FS.HLL, **BASE,
             **vdw)
    periodic(c, 0)
    periodic(c, 1)
    out["2d_vdw_periodic_hll"] = c
    c = Case(dim=3, N=(24, 10, 12), L=(1.0, 0.5, 0.5), limiter=TL.vanAlbada, **BASE, **vdw)
    out["3d_vdw_vanAlbada"] = c
    c = Case(dim=2, N=(24, 16, 1), L=(0.7, 0.5, 0.1), limiter=TL.linear, **BASE, **vdw)
    out["2d_vdw_linear_general"] = c
    # Redlich-Kwong mixture (cubic in sqrt(T)): compared with a tolerance
    rk = dict(gas_model=GM.RedlichKwong, Tc=[24.0, 3.2], Pc=[1.0, 1.0])
    out["1d_rk_vanLeer"] = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer,
                                **BASE, **rk)
    c = Case(dim=2, N=(64, 32, 1), L=(1.0, 0.5, 0.1), limiter=TL.minmod, flux=FS.HLL, **BASE,
             **rk)
    periodic(c, 0)
    c.set_bc(3, BCT.slip)
    out["2d_rk_periodic_hll"] = c
    out["3d_rk_vanAlbada"] = Case(dim=3, N=(24, 10, 12), L=(1.0, 0.5, 0.5),
                                  limiter=TL.vanAlbada, **BASE, **rk)
    out["2d_rk_linear_general"] = Case(dim=2, N=(24, 16, 1), L=(0.7, 0.5, 0.1),
                                       limiter=TL.linear, **BASE, **rk)
    # stiffened mixture
    st = dict(gas_model=GM.stiffened, p0=[0.3, 0.05], gammas=[2.8, 1.6])
    out["1d_stiffened_minmod"] = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0),
                                      limiter=TL.minmod, **BASE, **st)
    c = Case(dim=3, N=(24, 10, 12), L=(1.0, 0.5, 0.5), limiter=TL.vanAlbada, flux=FS.HLL,
             **BASE, **st)
    c.set_bc(3, BCT.slip)
    out["3d_stiffened_hll"] = c
    # Kataoka-van der Waals mixture (Newton for T): compared with a tolerance
    kat = dict(gas_model=GM.KataokaVanDerWaals, Tc=[24.0, 3.2], Pc=[1.0, 1.0],
               epsLJ=[10.0, 5.0], sigmaLJ=[1.52e-8, 1.2e-8], kataoka_b=("critical", 1.0))
    out["1d_kataoka_vanLeer"] = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0),
                                     limiter=TL.vanLeer, **BASE, **kat)
    c = Case(dim=2, N=(64, 32, 1), L=(1.0, 0.5, 0.1), limiter=TL.minmod, flux=FS.HLL, **BASE,
             **dict(kat, kataoka_b=("Kataoka", 2.0)))
    periodic(c, 0)
    out["2d_kataoka_periodic_hll"] = c
    out["3d_kataoka_vanAlbada"] = Case(dim=3, N=(24, 10, 12), L=(1.0, 0.5, 0.5),
                                       limiter=TL.vanAlbada, **BASE, **kat)
    # explicit molecular diffusion stage after each step (constant transport
    # coefficients; every boundary kind, non-separable 3-D geometry with a
    # reversed top face)
    dif = dict(diffusion=True, transport=(0.004, 0.02, 0.002))
    out["1d_diffusion_vanLeer"] = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0),
                                       limiter=TL.vanLeer, **BASE, **dif)
    c = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.minmod, flux=FS.HLL, **BASE,
             **dif)
    c.set_bc(0, BCT.slip)
    c.set_bc(1, BCT.fixedValueCell, "p", (1.2, 0, 0))
    c.set_bc(1, BCT.fixedValueCell, "v", (0.05, -0.02, 0.01))
    c.set_bc(1, BCT.fixedValueCell, "c", (0.03, 0, 0))
    out["1d_diffusion_slip_fixed"] = c
    c = Case(dim=2, N=(64, 32, 1), L=(1.0, 0.5, 0.1), limiter=TL.minmod, flux=FS.HLL, **BASE,
             **dif)
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    out["2d_diffusion_hll"] = c
    out["2d_diffusion_linear"] = Case(dim=2, N=(24, 16, 1), L=(0.7, 0.5, 0.1),
                                      limiter=TL.linear, **BASE, **dif)
    c = Case(dim=3, N=(24, 10, 12), L=(1.0, 0.5, 0.35), limiter=TL.vanAlbada, **BASE, **dif)
    c.set_bc(2, BCT.slip)
    c.set_bc(5, BCT.slip)
    out["3d_diffusion_lz_slip"] = c
    out["3d_diffusion_linear_lz"] = Case(dim=3, N=(10, 6, 6), L=(1.0, 0.6, 0.65),
                                         limiter=TL.linear, flux=FS.HLL, **BASE, **dif)
    # the Jacobi matrix solver (implicit Laplacians), converged and capped
    jac = dict(diffusion=True, transport=(0.02, 0.1, 0.01), sle_solver="Jacobi")
    out["1d_jacobi_vanLeer"] = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0),
                                    limiter=TL.vanLeer, **BASE, **jac)
    c = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.minmod, flux=FS.HLL, **BASE,
             **jac)
    c.set_bc(0, BCT.slip)
    c.set_bc(1, BCT.fixedValueCell, "p", (1.2, 0, 0))
    c.set_bc(1, BCT.fixedValueCell, "v", (0.05, -0.02, 0.01))
    c.set_bc(1, BCT.fixedValueCell, "c", (0.03, 0, 0))
    out["1d_jacobi_slip_fixed"] = c
    c = Case(dim=2, N=(64, 32, 1), L=(1.0, 0.5, 0.1), limiter=TL.minmod, flux=FS.HLL, **BASE,
             **jac)
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    out["2d_jacobi_hll"] = c
    c = Case(dim=3, N=(24, 10, 12), L=(1.0, 0.5, 0.35), limiter=TL.vanAlbada, **BASE, **jac)
    c.set_bc(2, BCT.slip)
    c.set_bc(5, BCT.slip)
    out["3d_jacobi_lz_slip"] = c
    out["1d_jacobi_capped"] = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0),
                                   limiter=TL.vanLeer, sle_iters=3, **BASE, **jac)
    # periodic axes (the wrapped face is an inner face)
    for nm, extra in (("diffusion", dif), ("jacobi", jac)):
        c = Case(dim=2, N=(64, 32, 1), L=(1.0, 0.5, 0.1), limiter=TL.vanLeer, **BASE, **extra)
        periodic(c, 0)
        periodic(c, 1)
        out[f"2d_{nm}_periodic_box"] = c
    c = Case(dim=3, N=(64, 10, 12), L=(1.0, 0.5, 0.5), limiter=TL.minmod, flux=FS.HLL, **BASE,
             **jac)
    periodic(c, 0)
    c.set_bc(3, BCT.slip)
    out["3d_jacobi_periodic_x"] = c
    # molar-mass correction of the species fluxes and the explicit enthalpy
    # flow (small D: the seeded mass fractions are rough)
    sp = dict(diffusion=True, transport=(0.004, 2e-4, 0.002), molar_mass_correction=True,
              enthalpy_flow=True)
    out["1d_molmass_enthalpy"] = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0),
                                      limiter=TL.vanLeer, **BASE, **sp)
    c = Case(dim=3, N=(24, 10, 12), L=(1.0, 0.5, 0.35), limiter=TL.vanAlbada, **BASE, **sp)
    c.set_bc(2, BCT.slip)
    c.set_bc(5, BCT.slip)
    out["3d_molmass_enthalpy_lz"] = c
    c = Case(dim=2, N=(64, 32, 1), L=(1.0, 0.5, 0.1), limiter=TL.minmod, flux=FS.HLL, **BASE,
             **dict(jac, molar_mass_correction=True))
    periodic(c, 0)
    c.set_bc(3, BCT.slip)
    out["2d_molmass_jacobi_periodic_x"] = c
    out["1d_molmass_enthalpy_3sp"] = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0),
                                          limiter=TL.minmod, M=[29.0, 4.0, 44.0],
                                          Cv=[2.5, 1 / 0.67, 3.4], **sp)
    # van der Waals mixture in the stage (non-ideality in the energy balance)
    vd = dict(gas_model=GM.vanDerWaals, Tc=[24.0, 3.2], Pc=[1.0, 1.0])
    out["1d_vdw_diffusion_molmass"] = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0),
                                           limiter=TL.vanLeer, **BASE, **vd,
                                           **dict(dif, transport=(0.004, 0.002, 0.002)),
                                           molar_mass_correction=True)
    c = Case(dim=2, N=(64, 32, 1), L=(1.0, 0.5, 0.1), limiter=TL.minmod, flux=FS.HLL, **BASE,
             **vd, **jac)
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    out["2d_vdw_jacobi"] = c
    c = Case(dim=3, N=(24, 10, 12), L=(1.0, 0.5, 0.35), limiter=TL.vanAlbada, **BASE, **vd,
             **dif)
    c.set_bc(2, BCT.slip)
    out["3d_vdw_diffusion_lz"] = c
    # Gauss-Seidel decks: compared with the oracle's sequential sweeps within
    # the solver tolerance
    gs = dict(jac, sle_solver="GaussSeidel")
    out["1d_gs_vanLeer"] = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer,
                                **BASE, **gs)
    c = Case(dim=3, N=(24, 10, 12), L=(1.0, 0.5, 0.35), limiter=TL.vanAlbada, **BASE, **gs)
    c.set_bc(2, BCT.slip)
    c.set_bc(5, BCT.slip)
    out["3d_gs_lz_slip"] = c
    # MUSCL-Hancock (SecondOrder) on every mesh kind it runs on: 1-D with each
    # limiter and boundary kind, and 2-D/3-D where the slopes vanish
    # (including the reversed top face normal, 3d_periodic_x_vanLeer)
    for name in ("1d_minmod", "1d_vanLeer", "1d_vanAlbada", "1d_linear", "1d_zero", "1d_hll",
                 "1d_slip", "1d_fixed", "1d_periodic", "1d_vdw_vanLeer", "1d_rk_vanLeer",
                 "1d_kataoka_vanLeer", "1d_diffusion_vanLeer", "2d_diffusion_hll",
                 "1d_jacobi_vanLeer", "2d_vanLeer",
                 "2d_hll_grav", "2d_periodic_hll", "2d_zero_fixed", "2d_vdw_periodic_hll",
                 "3d_vanAlbada", "3d_hll_partial_tiles", "3d_periodic_x_vanLeer",
                 "3d_vdw_vanAlbada"):
        c = copy.deepcopy(out[name])
        c.step_solver = 0
        out["hancock_" + name] = c
    # the linear limiter with MUSCL-Hancock in 2-D/3-D (general-geometry
    # kernels): non-dyadic, dyadic, periodic, reversed top face
    for name in ("2d_linear", "2d_linear_dyadic", "2d_periodic_y_linear", "3d_linear",
                 "3d_lz_short_linear_dyadic", "3d_lz_long_linear", "2d_vdw_linear_general",
                 "2d_rk_linear_general"):
        c = copy.deepcopy(out[name])
        c.step_solver = 0
        out["hancock_" + name] = c
    # reversed top face normal with a reflecting top
    c = Case(dim=3, N=(16, 8, 8), L=(1.0, 0.5, 0.25), limiter=TL.minmod, step_solver=0, **BASE)
    c.set_bc(5, BCT.slip)
    out["hancock_3d_lz_short_slip"] = c
    return out
