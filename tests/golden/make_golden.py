"""Writes the golden fixtures of tests/golden/ from the compiled reference.

Run in the development container only (it needs oracle/_ref/schemi, built by
``make -C oracle ref`` from the reference sources):

    python tests/golden/make_golden.py [group]

Each fixture holds the case description, the time steps, the reference's
initial state and its full-precision state after selected steps, parsed from
the reference's output tables (20 significant digits round-trip doubles).
Initial fields other than the 1-D shock tube are injected through the
reference's restart reader.
"""
from __future__ import annotations

import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import decks  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402
from schemi_amd import Case, typeOfTVDLimiter as TL, flowSolver as FS, gasModel  # noqa: E402
from schemi_amd import boundaryConditionType as BCT  # noqa: E402
from schemi_amd import BC  # noqa: E402
from schemi_amd.cases import cv_of  # noqa: E402

KEYS = ("c0", "c", "rho0", "rho", "v", "p", "U", "T", "m", "E")


def case_json(case: Case) -> str:
    eos = {} if int(case.gas_model) == 0 else {
        "gas_model": int(case.gas_model), "Tc": list(case.Tc), "Pc": list(case.Pc)}
    if int(case.gas_model) == 3:
        eos.update(p0=list(case.p0), gammas=list(case.gammas))
    if int(case.gas_model) == 4:
        eos.update(epsLJ=list(case.epsLJ), sigmaLJ=list(case.sigmaLJ),
                   kataoka_b=list(case.kataoka_b))
    if int(case.step_solver) != 1:
        eos["step_solver"] = int(case.step_solver)
    if case.diffusion:
        eos.update(diffusion=True, transport=list(case.transport))
        if case.sle_solver != "explicit":
            eos.update(sle_solver=case.sle_solver, sle_iters=int(case.sle_iters))
        if case.enthalpy_flow:
            eos["enthalpy_flow"] = True
        if case.molar_mass_correction:
            eos["molar_mass_correction"] = True
        if case.reconstruction != "linear":
            eos["reconstruction"] = case.reconstruction
    if case.timestep_source:
        eos.update(timestep_source=True, source_coeffs=list(case.source_coeffs))
    t = case.turbulence
    if t is not None:
        eos["turbulence"] = {
            "model": int(t.model), "params": list(t.params()), "C": list(t.C),
            "bc_k": [[int(b.type), list(b.value)] for b in t.bc_k],
            "bc_eps": [[int(b.type), list(b.value)] for b in t.bc_eps]}
        if t.a_field:
            eos["turbulence"].update(sigmaa=t.sigmaa,
                                     bc_a=[[int(b.type), list(b.value)] for b in t.bc_a])
        if t.b_field:
            eos["turbulence"].update(sigmab=t.sigmab, minb=t.minb,
                                     bc_b=[[int(b.type), list(b.value)] for b in t.bc_b])
    return json.dumps({**eos,
        "dim": case.dim, "N": list(case.N), "L": list(case.L), "M": list(case.M),
        "Cv": list(case.Cv), "R": case.R, "flux": int(case.flux),
        "limiter": int(case.limiter), "gravity": bool(case.gravity), "g": list(case.g),
        "courant": case.courant, "min_time": case.min_time,
        "bc_c": [[[int(b.type), list(b.value)] for b in row] for row in case.bc_c],
        "bc_v": [[int(b.type), list(b.value)] for b in case.bc_v],
        "bc_p": [[int(b.type), list(b.value)] for b in case.bc_p],
    })


def seeded_primitives(case: Case, seed: int):
    r = np.random.default_rng(seed)
    n = case.ncells
    c = np.stack([0.02 + 0.03 * r.random(n) for _ in range(case.Ns)])
    v = 0.2 * (r.random((3, n)) - 0.5)
    if case.dim < 3:
        v[2] = 0
    if case.dim < 2:
        v[1] = 0
    p = 0.5 + r.random(n)
    return c, v, p


def seeded_turbulence(case: Case, seed: int):
    """Seeded k and epsilon: eps / k between 2 and 60, turbulent viscosity
    Cmu k^2 / eps up to ~1e-4."""
    r = np.random.default_rng(seed)
    n = case.ncells
    return 1e-3 + 5e-3 * r.random(n), 1e-2 + 5e-2 * r.random(n)


def seeded_a(case: Case, seed: int):
    """Seeded turbulent mass flux a (3, n): components up to 0.1 in the
    mesh's directions, so that a.a / 2k straddles b and CMSA."""
    r = np.random.default_rng(seed + 1000)
    a = 0.1 * (r.random((3, case.ncells)) - 0.5)
    a[case.dim:] = 0
    return a


def seeded_b(case: Case, seed: int):
    """Seeded b of the BHR model, 0.05 .. 0.5."""
    return 0.05 + 0.45 * np.random.default_rng(seed + 2000).random(case.ncells)


def save(name, case, dts, snaps, times=None, restart=None, restart_turb=None, restart_a=None,
         restart_b=None):
    out = {"case": np.array(case_json(case)), "dts": np.asarray(dts, dtype=np.float64),
           "steps": np.array(sorted(snaps), dtype=np.int64)}
    if times is not None:
        out["times"] = np.asarray(times, dtype=np.float64)
    if restart is not None:
        # primitive fields given to the reference's restart reader
        out["init_c"], out["init_v"], out["init_p"] = (np.asarray(x) for x in restart)
    if restart_turb is not None:
        out["init_k"], out["init_eps"] = (np.asarray(x) for x in restart_turb)
    if restart_a is not None:
        out["init_a"] = np.asarray(restart_a)
    if restart_b is not None:
        out["init_b"] = np.asarray(restart_b)
    keys = KEYS + (("k", "eps") if case.turbulence is not None else ())
    if case.turbulence is not None and case.turbulence.a_field:
        keys += ("a", "b")
    for s, st in snaps.items():
        for k in keys:
            out[f"s{s}_{k}"] = np.asarray(st[k], dtype=np.float64)
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path) // 1024} KiB)")


def run_const(name, case, nsteps, dt, keep, zones=None, restart=None, restart_turb=None,
              restart_a=None, restart_b=None):
    t = 0.0
    for _ in range(nsteps):
        t = t + dt
    with tempfile.TemporaryDirectory(dir=os.path.join(ROOT, "oracle", "_ref")) as root:
        decks.write_deck(root, case, t_end=t, zones=zones, restart=restart, const_dt=dt,
                         restart_turb=restart_turb, restart_a=restart_a, restart_b=restart_b)
        r = decks.run(root)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: reference failed\n{r.stdout[-2000:]}\n{r.stderr}")
        off = 1 if restart is not None else 0
        snaps = {}
        if restart is None:
            snaps[0] = decks.read_output(root, 0, case.Ns)
        for s in keep:
            snaps[s] = decks.read_output(root, s + off, case.Ns)
    save(name, case, [dt] * nsteps, snaps, restart=restart, restart_turb=restart_turb,
         restart_a=restart_a, restart_b=restart_b)


def run_cfl_shock_tube(nsteps=200, keep=(1, 2, 10, 200)):
    """Config-0 shock tube with the reference's Courant step (first step =
    #Minimal_timestep).  The oracle finds the time after nsteps; the
    reference is run to exactly that time, so both take the same steps."""
    case = Case(dim=1, N=(1000, 1, 1), L=(1.0, 1.0, 1.0), M=[29.0, 4.0],
                Cv=[cv_of(1.4), cv_of(1.67)], flux=FS.HLLC, limiter=TL.minmod,
                courant=0.4, min_time=1e-15)
    zones = (([1 / 29, 0.0], (0., 0., 0.), 1.0), ([0.0, 0.125 / 4], (0., 0., 0.), 0.1), 0.5)
    with tempfile.TemporaryDirectory(dir=os.path.join(ROOT, "oracle", "_ref")) as root:
        # initial state from an initialisation-only reference run
        decks.write_deck(root, case, t_end=1.0, zones=zones)
        main = os.path.join(root, "set", "main.txt")
        txt = open(main).read().replace("#Start_From_Time-Step\nno",
                                        "#Start_From_Time-Step\ninitialisation")
        open(main, "w").write(txt)
        decks.run(root)
        init = decks.read_output(root, 0, case.Ns)
    o = Oracle(case)
    o.set_full(init)
    t, dt, dts, times = 0.0, case.min_time, [], []
    for _ in range(nsteps):
        assert o.step(dt, 0) == 0
        dts.append(dt)
        t = t + dt
        times.append(t)
        dt = o.courant_dt(case.courant, case.min_time)
    t_end = t
    with tempfile.TemporaryDirectory(dir=os.path.join(ROOT, "oracle", "_ref")) as root:
        decks.write_deck(root, case, t_end=t_end, zones=zones)
        r = decks.run(root)
        if r.returncode != 0:
            raise RuntimeError(r.stderr)
        tt = decks.read_times(root)
        if tt.shape[0] != nsteps + 1:
            raise RuntimeError(f"reference took {tt.shape[0] - 1} steps, expected {nsteps}")
        snaps = {0: decks.read_output(root, 0, case.Ns)}
        for s in keep:
            snaps[s] = decks.read_output(root, s, case.Ns)
        ref_times = tt[1:, 1]
    save("shock_tube_1d_cfl", case, dts, snaps, times=ref_times)


def main():
    os.makedirs(os.path.join(ROOT, "oracle", "_ref"), exist_ok=True)
    base = dict(M=[29.0, 4.0], Cv=[2.5, 1 / 0.67])
    # 1-D shock tube (config 0), constant step
    c0 = Case(dim=1, N=(1000, 1, 1), L=(1.0, 1.0, 1.0), M=[29.0, 4.0],
              Cv=[cv_of(1.4), cv_of(1.67)], flux=FS.HLLC, limiter=TL.minmod)
    zones = (([1 / 29, 0.0], (0., 0., 0.), 1.0), ([0.0, 0.125 / 4], (0., 0., 0.), 0.1), 0.5)
    run_const("shock_tube_1d_const", c0, 10, 1e-4, (1, 2, 10), zones=zones)
    run_cfl_shock_tube()
    # 1-D, every limiter and both fluxes, seeded fields, slip/fixed boundaries
    for lim in (TL.vanLeer, TL.vanAlbada, TL.linear, TL.zero):
        c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=lim, **base)
        run_const(f"seeded_1d_{lim.name}", c, 5, 2e-4, (1, 5),
                  restart=seeded_primitives(c, 11))
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), flux=FS.HLL, **base)
    c.set_bc(0, BCT.slip)
    c.set_bc(1, BCT.fixedValueCell, "p", (1.2, 0, 0))
    c.set_bc(1, BCT.fixedValueCell, "v", (0.1, 0.0, 0.0))
    run_const("seeded_1d_hll_slip_fixed", c, 5, 2e-4, (1, 5), restart=seeded_primitives(c, 12))
    # 2-D Rayleigh-Taylor-like: gravity, van Leer, slip walls
    c = Case(dim=2, N=(16, 32, 1), L=(0.25, 1.0, 1.0 / 32), M=[87.0, 29.0],
             Cv=[cv_of(1.4), cv_of(1.4)], limiter=TL.vanLeer, gravity=True, g=(0.0, -1.0, 0.0))
    for pt in (0, 1, 3, 4):
        c.set_bc(pt, BCT.slip)
    run_const("seeded_2d_rt_vanleer", c, 5, 2e-4, (1, 5), restart=seeded_primitives(c, 13))
    # 2-D linear limiter + HLL on a non-dyadic mesh
    c = Case(dim=2, N=(12, 10, 1), L=(0.3, 0.7, 0.1), limiter=TL.linear, flux=FS.HLL, **base)
    run_const("seeded_2d_linear_hll", c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, 14))
    # 3-D Richtmyer-Meshkov-like: van Albada, free x, slip y/z
    c = Case(dim=3, N=(16, 8, 8), L=(2.0, 1.0, 1.0), M=[29.0, 146.0],
             Cv=[cv_of(1.4), cv_of(1.09)], limiter=TL.vanAlbada)
    for pt in (2, 3, 4, 5):
        c.set_bc(pt, BCT.slip)
    run_const("seeded_3d_rm_vanalbada", c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, 15))
    c = Case(dim=3, N=(8, 6, 6), L=(1.0, 0.5, 0.5), limiter=TL.linear, **base)
    run_const("seeded_3d_linear", c, 3, 2e-4, (1, 3), restart=seeded_primitives(c, 16))
    lz_cases(base)
    vdw_cases(base)
    hancock_cases(base)


def lz_cases(base):
    """3-D meshes with Lz != Ly: the reference spaces z by Ly / Nz, so with
    Lz < Ly the top face plane lies below the last cell centre (its normal
    is reversed) and with Lz > Ly above the last cell."""
    c = Case(dim=3, N=(8, 6, 6), L=(1.0, 0.5, 0.25), limiter=TL.vanLeer, **base)
    c.set_bc(5, BCT.slip)
    run_const("seeded_3d_lz_short_vanleer", c, 3, 2e-4, (1, 3), restart=seeded_primitives(c, 17))
    c = Case(dim=3, N=(8, 4, 4), L=(1.0, 0.5, 0.25), limiter=TL.linear, flux=FS.HLL, **base)
    run_const("seeded_3d_lz_short_linear", c, 3, 2e-4, (1, 3), restart=seeded_primitives(c, 18))
    c = Case(dim=3, N=(8, 4, 4), L=(1.0, 0.5, 0.55), limiter=TL.linear, **base)
    run_const("seeded_3d_lz_long_linear", c, 3, 2e-4, (1, 3), restart=seeded_primitives(c, 19))


def vdw_cases(base):
    """van der Waals mixture (gasModel::vanDerWaals).  With Tc = (24, 3.2),
    Pc = (1, 1) and R = 1, c b is about 0.1 and a c^2 about a quarter of p for
    the seeded states: a clearly non-ideal closure."""
    eos = dict(gas_model=gasModel.vanDerWaals, Tc=[24.0, 3.2], Pc=[1.0, 1.0])
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.minmod, **base, **eos)
    run_const("seeded_1d_vdw_minmod", c, 5, 2e-4, (1, 5), restart=seeded_primitives(c, 21))
    c = Case(dim=2, N=(16, 12, 1), L=(0.5, 0.375, 0.1), limiter=TL.minmod, flux=FS.HLL,
             **base, **eos)
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    run_const("seeded_2d_vdw_hll", c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, 22))
    c = Case(dim=3, N=(8, 6, 6), L=(1.0, 0.5, 0.5), limiter=TL.vanAlbada, **base, **eos)
    run_const("seeded_3d_vdw_vanalbada", c, 3, 2e-4, (1, 3), restart=seeded_primitives(c, 23))


def rk_cases(base):
    """Redlich-Kwong mixture (gasModel::RedlichKwong): with Tc = (24, 3.2),
    Pc = (1, 1), R = 1, c b is about 0.05-0.1 and a c^2 / sqrt(T) about a
    fifth of p for the seeded states."""
    eos = dict(gas_model=gasModel.RedlichKwong, Tc=[24.0, 3.2], Pc=[1.0, 1.0])
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer, **base, **eos)
    run_const("seeded_1d_rk_vanleer", c, 5, 2e-4, (1, 5), restart=seeded_primitives(c, 41))
    c = Case(dim=2, N=(16, 12, 1), L=(0.5, 0.375, 0.1), limiter=TL.minmod, flux=FS.HLL,
             **base, **eos)
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    run_const("seeded_2d_rk_hll", c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, 42))
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.minmod, step_solver=0,
             **base, **eos)
    run_const("seeded_1d_rk_hancock", c, 5, 2e-4, (1, 5), restart=seeded_primitives(c, 43))


def stiffened_cases(base):
    """Stiffened mixture (gasModel::stiffened): p = (gamma - 1) U - gamma p0
    with p0 = (0.3, 0.05) and gamma = (2.8, 1.6) per species, mixed by
    geometric means."""
    eos = dict(gas_model=gasModel.stiffened, p0=[0.3, 0.05], gammas=[2.8, 1.6])
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer, **base, **eos)
    run_const("seeded_1d_stiffened_vanleer", c, 5, 2e-4, (1, 5),
              restart=seeded_primitives(c, 51))
    c = Case(dim=2, N=(16, 12, 1), L=(0.5, 0.375, 0.1), limiter=TL.minmod, flux=FS.HLL,
             **base, **eos)
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    run_const("seeded_2d_stiffened_hll", c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, 52))


KATAOKA = dict(gas_model=gasModel.KataokaVanDerWaals, Tc=[24.0, 3.2], Pc=[1.0, 1.0],
               epsLJ=[10.0, 5.0], sigmaLJ=[1.52e-8, 1.2e-8], kataoka_b=("critical", 1.0))


def kataoka_cases(base):
    """Kataoka-van der Waals mixture (gasModel::KataokaVanDerWaals): V0 c
    about 0.05 and eps / (2 R T) about 0.15 for the seeded states, so the
    association terms change p by about a fifth."""
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer, **base, **KATAOKA)
    run_const("seeded_1d_kataoka_vanleer", c, 5, 2e-4, (1, 5), restart=seeded_primitives(c, 61))
    c = Case(dim=2, N=(16, 12, 1), L=(0.5, 0.375, 0.1), limiter=TL.minmod, flux=FS.HLL,
             **base, **dict(KATAOKA, kataoka_b=("Kataoka", 2.0)))
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    run_const("seeded_2d_kataoka_hll", c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, 62))


def hancock_cases(base):
    """MUSCL-Hancock (#Approximation_order SecondOrder): one Advection with
    the half-step face predictor per step."""
    h = dict(step_solver=0)
    for lim in (TL.vanLeer, TL.minmod):
        c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=lim, **base, **h)
        run_const(f"seeded_1d_hancock_{lim.name}", c, 5, 2e-4, (1, 5),
                  restart=seeded_primitives(c, 31))
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), flux=FS.HLL, limiter=TL.vanAlbada,
             **base, **h)
    c.set_bc(0, BCT.slip)
    c.set_bc(1, BCT.fixedValueCell, "p", (1.2, 0, 0))
    run_const("seeded_1d_hancock_hll_slip_fixed", c, 5, 2e-4, (1, 5),
              restart=seeded_primitives(c, 32))
    c = Case(dim=2, N=(16, 12, 1), L=(0.5, 0.375, 0.1), limiter=TL.minmod, flux=FS.HLL,
             **base, **h)
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    run_const("seeded_2d_hancock_minmod_hll", c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, 33))
    c = Case(dim=3, N=(8, 6, 6), L=(1.0, 0.5, 0.5), limiter=TL.vanAlbada, **base, **h)
    for pt in (2, 3, 4, 5):
        c.set_bc(pt, BCT.slip)
    run_const("seeded_3d_hancock_vanalbada", c, 3, 2e-4, (1, 3), restart=seeded_primitives(c, 34))
    c = Case(dim=2, N=(16, 12, 1), L=(0.5, 0.375, 0.1), limiter=TL.linear, **base, **h)
    c.set_bc(3, BCT.slip)
    run_const("seeded_2d_hancock_linear", c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, 35))
    c = Case(dim=3, N=(8, 6, 6), L=(1.0, 0.5, 0.35), limiter=TL.linear, flux=FS.HLL, **base,
             **h)
    run_const("seeded_3d_hancock_linear_lz", c, 3, 2e-4, (1, 3), restart=seeded_primitives(c, 36))
    c = Case(dim=2, N=(16, 12, 1), L=(0.5, 0.375, 0.1), limiter=TL.vanLeer, **base, **h,
             gas_model=gasModel.vanDerWaals, Tc=[24.0, 3.2], Pc=[1.0, 1.0])
    run_const("seeded_2d_hancock_vdw", c, 3, 2e-4, (1, 3), restart=seeded_primitives(c, 35))


def diffusion_cases(base):
    """Molecular diffusion stage (#Diffusion_flag on, explicit matrix
    solver, linear reconstruction, constant transport coefficients) after
    each step: viscosity mu, diffusion coefficient D and heat conduction
    kappa chosen so that D dt / dx^2 and the thermal and viscous analogues
    are 0.01-0.1 for the seeded states."""
    tr = dict(diffusion=True, transport=(0.004, 0.02, 0.002))
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer, **base, **tr)
    run_const("seeded_1d_diffusion_vanleer", c, 5, 2e-4, (1, 5), restart=seeded_primitives(c, 71))
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), flux=FS.HLL, limiter=TL.minmod,
             **base, **tr)
    c.set_bc(0, BCT.slip)
    c.set_bc(1, BCT.fixedValueCell, "p", (1.2, 0, 0))
    c.set_bc(1, BCT.fixedValueCell, "v", (0.05, 0, 0))
    c.set_bc(1, BCT.fixedValueCell, "c", (0.03, 0, 0))
    run_const("seeded_1d_diffusion_slip_fixed", c, 5, 2e-4, (1, 5),
              restart=seeded_primitives(c, 72))
    c = Case(dim=2, N=(16, 12, 1), L=(0.5, 0.375, 0.1), limiter=TL.minmod, flux=FS.HLL,
             **base, **tr)
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    run_const("seeded_2d_diffusion_hll", c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, 73))
    c = Case(dim=2, N=(16, 12, 1), L=(0.5, 0.375, 0.1), limiter=TL.linear, **base, **tr)
    c.set_bc(3, BCT.slip)
    run_const("seeded_2d_diffusion_linear", c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, 74))
    c = Case(dim=3, N=(8, 6, 6), L=(1.0, 0.5, 0.35), limiter=TL.vanAlbada, **base, **tr)
    for pt in (2, 5):
        c.set_bc(pt, BCT.slip)
    run_const("seeded_3d_diffusion_lz", c, 3, 2e-4, (1, 3), restart=seeded_primitives(c, 75))
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer, step_solver=0,
             **base, **tr)
    run_const("seeded_1d_diffusion_hancock", c, 5, 2e-4, (1, 5),
              restart=seeded_primitives(c, 76))


def jacobi_cases(base):
    """The diffusion stage with the Jacobi matrix solver
    (#Type_of_SLE_solver_flag Jacobi): implicit Laplacians solved to the
    reference's tolerance.  Larger coefficients than the explicit cases, so
    D dt / dx^2 is about 0.1-1 and the solves take several iterations."""
    tr = dict(diffusion=True, transport=(0.02, 0.1, 0.01), sle_solver="Jacobi")
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer, **base, **tr)
    run_const("seeded_1d_jacobi_vanleer", c, 5, 2e-4, (1, 5), restart=seeded_primitives(c, 81))
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), flux=FS.HLL, limiter=TL.minmod,
             **base, **tr)
    c.set_bc(0, BCT.slip)
    c.set_bc(1, BCT.fixedValueCell, "p", (1.2, 0, 0))
    c.set_bc(1, BCT.fixedValueCell, "v", (0.05, 0, 0))
    c.set_bc(1, BCT.fixedValueCell, "c", (0.03, 0, 0))
    run_const("seeded_1d_jacobi_slip_fixed", c, 5, 2e-4, (1, 5),
              restart=seeded_primitives(c, 82))
    c = Case(dim=2, N=(16, 12, 1), L=(0.5, 0.375, 0.1), limiter=TL.minmod, flux=FS.HLL,
             **base, **tr)
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    run_const("seeded_2d_jacobi_hll", c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, 83))
    c = Case(dim=3, N=(8, 6, 6), L=(1.0, 0.5, 0.35), limiter=TL.vanAlbada, **base, **tr)
    for pt in (2, 5):
        c.set_bc(pt, BCT.slip)
    run_const("seeded_3d_jacobi_lz", c, 3, 2e-4, (1, 3), restart=seeded_primitives(c, 84))
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer, step_solver=0,
             sle_iters=3, **base, **tr)
    run_const("seeded_1d_jacobi_hancock_capped", c, 5, 2e-4, (1, 5),
              restart=seeded_primitives(c, 85))


def gauss_seidel_cases(base):
    """The diffusion stage with the Gauss-Seidel matrix solver (the reference
    decks' default): the implicit Laplacians of the Jacobi cases solved by
    forward and backward sweeps."""
    tr = dict(diffusion=True, transport=(0.02, 0.1, 0.01), sle_solver="GaussSeidel")
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer, **base, **tr)
    run_const("seeded_1d_gs_vanleer", c, 5, 2e-4, (1, 5), restart=seeded_primitives(c, 91))
    c = Case(dim=2, N=(16, 12, 1), L=(0.5, 0.375, 0.1), limiter=TL.minmod, flux=FS.HLL,
             **base, **tr)
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    run_const("seeded_2d_gs_hll", c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, 92))
    c = Case(dim=3, N=(8, 6, 6), L=(1.0, 0.5, 0.35), limiter=TL.vanAlbada, **base, **tr)
    for pt in (2, 5):
        c.set_bc(pt, BCT.slip)
    run_const("seeded_3d_gs_lz", c, 3, 2e-4, (1, 3), restart=seeded_primitives(c, 93))


def species_flux_cases(base):
    """The diffusion stage with the molar-mass correction of the species
    fluxes (Gauss elimination per face) and the explicit enthalpy flow, as
    the reference's own diffusion test deck enables them."""
    tr = dict(diffusion=True, transport=(0.004, 0.02, 0.002), molar_mass_correction=True)
    # the enthalpy flow of the seeded (cell-to-cell random) mass fractions is
    # strong: a small D keeps the explicit stage stable (the reference fails
    # with a Riemann error at step 6 with D = 0.02, as does the oracle)
    en = dict(tr, transport=(0.004, 2e-4, 0.002), enthalpy_flow=True)
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer, **base, **en)
    run_const("seeded_1d_molmass_enthalpy", c, 5, 2e-4, (1, 5), restart=seeded_primitives(c, 101))
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), flux=FS.HLL, limiter=TL.minmod, **base,
             **tr)
    c.set_bc(0, BCT.slip)
    c.set_bc(1, BCT.fixedValueCell, "p", (1.2, 0, 0))
    c.set_bc(1, BCT.fixedValueCell, "c", (0.03, 0, 0))
    run_const("seeded_1d_molmass_slip_fixed", c, 5, 2e-4, (1, 5),
              restart=seeded_primitives(c, 102))
    c = Case(dim=2, N=(16, 12, 1), L=(0.5, 0.375, 0.1), limiter=TL.minmod, flux=FS.HLL,
             **base, **en)
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    run_const("seeded_2d_molmass_enthalpy", c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, 103))
    c = Case(dim=3, N=(8, 6, 6), L=(1.0, 0.5, 0.35), limiter=TL.vanAlbada,
             **base, **dict(tr, transport=(0.02, 0.1, 0.01)), sle_solver="Jacobi")
    run_const("seeded_3d_molmass_jacobi", c, 3, 2e-4, (1, 3), restart=seeded_primitives(c, 104))
    b3 = dict(M=[29.0, 4.0, 44.0], Cv=[2.5, 1 / 0.67, 3.4])
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer, **b3, **en)
    run_const("seeded_1d_molmass_3species", c, 5, 2e-4, (1, 5), restart=seeded_primitives(c, 105))


def vdw_diffusion_cases(base):
    """The diffusion stage with the van der Waals mixture: the non-ideality
    -a c^2 in the energy balance and U = (Cv T - a c) c in the recalculated
    cells (Diffusion.cpp:484-485, 991-993)."""
    eos = dict(gas_model=gasModel.vanDerWaals, Tc=[24.0, 3.2], Pc=[1.0, 1.0])
    tr = dict(diffusion=True, transport=(0.004, 0.02, 0.002))
    # with the molar-mass correction D = 0.02 ends in a Riemann error at
    # step 5 in the reference and the oracle alike: D = 0.002
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer, **base, **eos,
             **dict(tr, transport=(0.004, 0.002, 0.002)), molar_mass_correction=True)
    run_const("seeded_1d_vdw_diffusion", c, 5, 2e-4, (1, 5), restart=seeded_primitives(c, 111))
    c = Case(dim=2, N=(16, 12, 1), L=(0.5, 0.375, 0.1), limiter=TL.minmod, flux=FS.HLL,
             **base, **eos, **dict(tr, transport=(0.02, 0.1, 0.01)), sle_solver="Jacobi")
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    run_const("seeded_2d_vdw_jacobi", c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, 112))
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.minmod, step_solver=0,
             **base, **eos, **dict(tr, transport=(0.02, 0.1, 0.01)), sle_solver="GaussSeidel")
    run_const("seeded_1d_vdw_gs_hancock", c, 5, 2e-4, (1, 5), restart=seeded_primitives(c, 113))


def courant_source_cases(base, nsteps=5, keep=(1, 3, 5)):
    """The "Courant-Source" time step: the explicit enthalpy flow of the
    diffusion stage sets the source step (Diffusion.cpp:696-700), which caps
    the next Courant step (Schemi.cpp:828-833).  With D = 0.02 on the seeded
    mass fractions the source step binds from step 2 on and falls below
    #Minimal_timestep at step 5.  The oracle finds the step sequence; the
    reference is run to the time after nsteps and must take the same steps."""
    tr = dict(diffusion=True, transport=(0.004, 0.02, 0.002), molar_mass_correction=True,
              enthalpy_flow=True, timestep_source=True)
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer, courant=0.3,
             min_time=1e-6, **base, **tr)
    init = seeded_primitives(c, 101)
    o = Oracle(c)
    o.init_primitive(*init)
    t, dt, dts, times = 0.0, c.min_time, [], []
    for _ in range(nsteps):
        assert o.step(dt, 0) == 0
        dts.append(dt)
        t = t + dt
        times.append(t)
        dt = o.courant_dt(c.courant, c.min_time)
    with tempfile.TemporaryDirectory(dir=os.path.join(ROOT, "oracle", "_ref")) as root:
        decks.write_deck(root, c, t_end=t, restart=init)
        r = decks.run(root)
        if r.returncode != 0:
            raise RuntimeError(f"reference failed\n{r.stdout[-2000:]}\n{r.stderr}")
        tt = decks.read_times(root)
        if tt.shape[0] != nsteps + 2:
            raise RuntimeError(f"reference took {tt.shape[0] - 2} steps, expected {nsteps}")
        snaps = {s: decks.read_output(root, s + 1, c.Ns) for s in keep}
        ref_times = tt[2:, 1]
    print("dts", dts, "reference times", ref_times)
    save("seeded_1d_courant_source", c, dts, snaps, times=ref_times, restart=init)


def turbulence_cases(base, nsteps=4, keep=(1, 4)):
    """k-epsilon turbulence (#Turbulence_flag on): rho k and rho epsilon
    through the HLLC/HLL flux with the turbulent pressure (2/3) rho k and the
    sound-speed term (HLLCSolver.cpp:26-64), and, with the diffusion stage,
    the turbulent transport coefficients, the k and epsilon Laplacians and the
    shear model's sources (Diffusion.cpp:129-933, shearModel.cpp:77-153)."""
    from schemi_amd.case import Turbulence
    tr = dict(diffusion=True, transport=(2e-3, 2e-3, 2e-3))
    vdw = dict(gas_model=gasModel.vanDerWaals, Tc=[24.0, 3.2], Pc=[1.0, 1.0])

    def slip_y(c):
        for p in (3, 4):
            c.set_bc(p, BCT.slip)
        return c

    def go(name, c, seed, n=nsteps, kp=keep):
        run_const(name, c, n, 2e-4 if c.dim > 1 else 1e-3, kp,
                  restart=seeded_primitives(c, seed), restart_turb=seeded_turbulence(c, seed))

    # advection only (turbulence on, diffusion off)
    go("seeded_2d_turb_advection_hllc",
       slip_y(Case(dim=2, N=(12, 10, 1), L=(1.0, 0.8, 1.0), limiter=TL.minmod,
                   turbulence=Turbulence(), **base)), 201)
    # shear model with the diffusion stage
    go("seeded_2d_turb_shear_hllc",
       slip_y(Case(dim=2, N=(12, 10, 1), L=(1.0, 0.8, 1.0), limiter=TL.minmod,
                   turbulence=Turbulence(), **base, **tr)), 202)
    # the configuration of the Kelvin-Helmholtz workload: HLL, minmod, van der
    # Waals mixture, shear model
    go("seeded_2d_turb_shear_vdw_hll",
       slip_y(Case(dim=2, N=(12, 10, 1), L=(1.0, 0.8, 1.0), flux=FS.HLL, limiter=TL.minmod,
                   turbulence=Turbulence(sigmaE=0.9), **base, **tr, **vdw)), 203)
    # 1-D van Leer (limited slopes of k and epsilon), fixed-value and slip ends
    c = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer,
             turbulence=Turbulence(), **base, **tr)
    c.set_bc(0, BCT.fixedValueCell, "all", (0.03, 0.0, 0.0))
    c.set_bc(1, BCT.slip)
    c.turbulence.bc_k[0] = BC(BCT.fixedValueCell, (2e-3, 0.0, 0.0))
    c.turbulence.bc_eps[0] = BC(BCT.fixedValueCell, (3e-2, 0.0, 0.0))
    c.turbulence.bc_eps[1] = BC(BCT.slip)
    go("seeded_1d_turb_shear_vanleer", c, 204)
    # 3-D van Albada with Lz != Ly (reversed top faces)
    go("seeded_3d_turb_shear_vanalbada",
       Case(dim=3, N=(8, 6, 5), L=(1.0, 0.8, 0.6), limiter=TL.vanAlbada,
            turbulence=Turbulence(), **base, **tr), 205, n=3, kp=(1, 3))
    # zero model: transport and diffusion of k and epsilon, no sources
    from schemi_amd.enums import turbulenceModel
    go("seeded_2d_turb_zero_hllc",
       slip_y(Case(dim=2, N=(12, 10, 1), L=(1.0, 0.8, 1.0), limiter=TL.minmod,
                   turbulence=Turbulence(model=turbulenceModel.zeroModel), **base, **tr)), 206)


def kepsa_cases(base, nsteps=4, keep=(1, 4)):
    """kEpsA (kEpsAModel.cpp): rho a through the HLLC/HLL flux like rho k,
    vector TVD slopes of a in 1-D, and in the diffusion stage the algebraic b
    (calculate_b), grad a, grad p, grad rho, the molecular viscous stress
    divergence, the a Laplacian with mu + rho nu_t / sigma_a, the sources of
    k, epsilon and a with the thetaA limiter, and the a . (grad p - div
    tau) energy sink (Diffusion.cpp:129-1029)."""
    from schemi_amd.case import Turbulence
    tr = dict(diffusion=True, transport=(2e-3, 2e-3, 2e-3))
    vdw = dict(gas_model=gasModel.vanDerWaals, Tc=[24.0, 3.2], Pc=[1.0, 1.0])

    def slip_y(c):
        for p in (3, 4):
            c.set_bc(p, BCT.slip)
            c.turbulence.bc_a[p] = BC(BCT.slip)
        return c

    def go(name, c, seed, n=nsteps, kp=keep):
        run_const(name, c, n, 2e-4 if c.dim > 1 else 1e-3, kp,
                  restart=seeded_primitives(c, seed), restart_turb=seeded_turbulence(c, seed),
                  restart_a=seeded_a(c, seed))

    # advection only: rho a transported with the flux of rho k
    go("seeded_2d_kepsa_advection_hllc",
       slip_y(Case(dim=2, N=(12, 10, 1), L=(1.0, 0.8, 1.0), limiter=TL.minmod,
                   turbulence=Turbulence.kEpsA(), **base)), 211)
    # with the diffusion stage, HLLC, ideal mixture
    go("seeded_2d_kepsa_hllc",
       slip_y(Case(dim=2, N=(12, 10, 1), L=(1.0, 0.8, 1.0), limiter=TL.minmod,
                   turbulence=Turbulence.kEpsA(), **base, **tr)), 212)
    # HLL, van der Waals mixture, thetaA off (CMSA >= 1e15), chi = 0
    go("seeded_2d_kepsa_vdw_hll",
       slip_y(Case(dim=2, N=(12, 10, 1), L=(1.0, 0.8, 1.0), flux=FS.HLL, limiter=TL.minmod,
                   turbulence=Turbulence.kEpsA(CMSA=1e16, chi=0.0, sigmaE=0.9),
                   **base, **tr, **vdw)), 213)
    # 1-D van Leer (vector slopes of a), fixed-value and slip ends
    c = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer,
             turbulence=Turbulence.kEpsA(), **base, **tr)
    c.set_bc(0, BCT.fixedValueCell, "all", (0.03, 0.0, 0.0))
    c.set_bc(1, BCT.slip)
    c.turbulence.bc_k[0] = BC(BCT.fixedValueCell, (2e-3, 0.0, 0.0))
    c.turbulence.bc_a[0] = BC(BCT.fixedValueCell, (0.02, 0.0, 0.0))
    c.turbulence.bc_a[1] = BC(BCT.slip)
    go("seeded_1d_kepsa_vanleer", c, 214)
    # 3-D van Albada with Lz != Ly
    go("seeded_3d_kepsa_vanalbada",
       Case(dim=3, N=(8, 6, 5), L=(1.0, 0.8, 0.6), limiter=TL.vanAlbada,
            turbulence=Turbulence.kEpsA(), **base, **tr), 215, n=3, kp=(1, 3))


def bhr_cases(base, nsteps=4, keep=(1, 4)):
    """BHR (BHRModel.cpp, the Goncharov initialisation model off as in every
    reference deck): rho a and rho b through the HLLC/HLL flux and the TVD
    slopes, and in the diffusion stage the a and b Laplacians, the sources
    with the turbulent Mach number, the pressure-dilatation term Pi_K, the
    redistribution rho div(a a), the energy sink, rho eps (1 + CMS_M M_t^2)
    in the temperature equation and the source step with |U| / G
    (Diffusion.cpp:129-1029, BHRModel.cpp:139-272)."""
    from schemi_amd.case import Turbulence
    tr = dict(diffusion=True, transport=(2e-3, 2e-3, 2e-3))
    vdw = dict(gas_model=gasModel.vanDerWaals, Tc=[24.0, 3.2], Pc=[1.0, 1.0])

    def slip_y(c):
        for p in (3, 4):
            c.set_bc(p, BCT.slip)
            c.turbulence.bc_a[p] = BC(BCT.slip)
        return c

    def go(name, c, seed, n=nsteps, kp=keep):
        run_const(name, c, n, 2e-4 if c.dim > 1 else 1e-3, kp,
                  restart=seeded_primitives(c, seed), restart_turb=seeded_turbulence(c, seed),
                  restart_a=seeded_a(c, seed), restart_b=seeded_b(c, seed))

    go("seeded_2d_bhr_advection_hllc",
       slip_y(Case(dim=2, N=(12, 10, 1), L=(1.0, 0.8, 1.0), limiter=TL.minmod,
                   turbulence=Turbulence.BHR(), **base)), 251)
    go("seeded_2d_bhr_hllc",
       slip_y(Case(dim=2, N=(12, 10, 1), L=(1.0, 0.8, 1.0), limiter=TL.minmod,
                   turbulence=Turbulence.BHR(), **base, **tr)), 252)
    go("seeded_2d_bhr_vdw_hll",
       slip_y(Case(dim=2, N=(12, 10, 1), L=(1.0, 0.8, 1.0), flux=FS.HLL, limiter=TL.minmod,
                   turbulence=Turbulence.BHR(sigmaE=0.9), **base, **tr, **vdw)), 253)
    c = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer,
             turbulence=Turbulence.BHR(), **base, **tr)
    c.set_bc(0, BCT.fixedValueCell, "all", (0.03, 0.0, 0.0))
    c.set_bc(1, BCT.slip)
    c.turbulence.bc_a[1] = BC(BCT.slip)
    c.turbulence.bc_b[0] = BC(BCT.fixedValueCell, (0.3, 0.0, 0.0))
    go("seeded_1d_bhr_vanleer", c, 254)
    go("seeded_3d_bhr_vanalbada",
       Case(dim=3, N=(8, 6, 5), L=(1.0, 0.8, 0.6), limiter=TL.vanAlbada,
            turbulence=Turbulence.BHR(), **base, **tr), 255, n=3, kp=(1, 3))
    # the reference decks' setting: star reconstruction
    go("seeded_2d_bhr_star_hll",
       slip_y(Case(dim=2, N=(12, 10, 1), L=(1.0, 0.8, 1.0), flux=FS.HLL, limiter=TL.minmod,
                   turbulence=Turbulence.BHR(), reconstruction="star", **base, **tr)), 256)


def star_cases(base):
    """#Reconstruction star: the diffusion stage takes the cell gradients of
    v, rho, p (and a) from the Riemann star states of the faces, averaged over
    the two SSP-RK2 stages (Diffusion.cpp:176-193, 529-537;
    secondOrderStepSolverRK.cpp:73-85; HLLCSolver.cpp:390-402,
    HLLSolver.cpp:302-314, pressureStarClass.cpp), or of the one
    MUSCL-Hancock stage, whose star velocity is replaced by the interpolated
    velocity under gravity (secondOrderStepSolver.cpp:74-76).  The
    reconstruction every diffusion-on reference deck uses."""
    from schemi_amd.case import Turbulence
    st = dict(diffusion=True, transport=(0.004, 0.02, 0.002), reconstruction="star")
    vdw = dict(gas_model=gasModel.vanDerWaals, Tc=[24.0, 3.2], Pc=[1.0, 1.0])
    c = Case(dim=2, N=(16, 12, 1), L=(0.5, 0.375, 0.1), limiter=TL.minmod, **base, **st)
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    run_const("seeded_2d_star_hllc", c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, 231))
    c = Case(dim=2, N=(16, 12, 1), L=(0.5, 0.375, 0.1), limiter=TL.minmod, flux=FS.HLL,
             **base, **st, **vdw)
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    run_const("seeded_2d_star_vdw_hll", c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, 232))
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), flux=FS.HLL, limiter=TL.vanLeer,
             **base, **st)
    c.set_bc(0, BCT.slip)
    c.set_bc(1, BCT.fixedValueCell, "p", (1.2, 0, 0))
    c.set_bc(1, BCT.fixedValueCell, "v", (0.05, 0, 0))
    c.set_bc(1, BCT.fixedValueCell, "c", (0.03, 0, 0))
    run_const("seeded_1d_star_vanleer_hll", c, 5, 2e-4, (1, 5), restart=seeded_primitives(c, 233))
    c = Case(dim=3, N=(8, 6, 6), L=(1.0, 0.5, 0.35), limiter=TL.vanAlbada, **base, **st)
    for pt in (2, 5):
        c.set_bc(pt, BCT.slip)
    run_const("seeded_3d_star_vanalbada", c, 3, 2e-4, (1, 3), restart=seeded_primitives(c, 234))
    # gravity with the SSP-RK2 step (the star velocity stays) and with
    # MUSCL-Hancock (replaced by the interpolated velocity)
    for name, extra, seed in (("seeded_2d_star_gravity", {}, 235),
                              ("seeded_2d_star_hancock_gravity", dict(step_solver=0), 236),
                              ("seeded_2d_star_hancock", dict(step_solver=0, gravity=False), 237)):
        kw = dict(gravity=True, g=(0.0, -1.0, 0.0))
        kw.update(extra)
        c = Case(dim=2, N=(16, 12, 1), L=(0.5, 0.375, 0.1), limiter=TL.minmod, flux=FS.HLL,
                 **base, **st, **kw)
        c.set_bc(3, BCT.slip)
        c.set_bc(4, BCT.slip)
        run_const(name, c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, seed))
    # the Jacobi solver's heating term
    c = Case(dim=2, N=(16, 12, 1), L=(0.5, 0.375, 0.1), limiter=TL.minmod,
             **base, **dict(st, transport=(0.02, 0.1, 0.01), sle_solver="Jacobi"))
    c.set_bc(3, BCT.slip)
    c.set_bc(4, BCT.slip)
    run_const("seeded_2d_star_jacobi", c, 4, 2e-4, (1, 4), restart=seeded_primitives(c, 238))
    # turbulence: the shear model's grad v terms and kEpsA's grad p, grad rho
    # and grad a
    tr = dict(diffusion=True, transport=(2e-3, 2e-3, 2e-3), reconstruction="star")

    def slip_y(c):
        for p in (3, 4):
            c.set_bc(p, BCT.slip)
            c.turbulence.bc_a[p] = BC(BCT.slip)
        return c

    c = slip_y(Case(dim=2, N=(12, 10, 1), L=(1.0, 0.8, 1.0), limiter=TL.minmod,
                    turbulence=Turbulence(), **base, **tr))
    run_const("seeded_2d_star_turb_shear_hllc", c, 4, 2e-4, (1, 4),
              restart=seeded_primitives(c, 239), restart_turb=seeded_turbulence(c, 239))
    c = slip_y(Case(dim=2, N=(12, 10, 1), L=(1.0, 0.8, 1.0), flux=FS.HLL, limiter=TL.minmod,
                    turbulence=Turbulence.kEpsA(), **base, **tr, **vdw))
    run_const("seeded_2d_star_kepsa_vdw_hll", c, 4, 2e-4, (1, 4),
              restart=seeded_primitives(c, 240), restart_turb=seeded_turbulence(c, 240),
              restart_a=seeded_a(c, 240))
    c = Case(dim=1, N=(40, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer,
             turbulence=Turbulence.kEpsA(), **base, **tr)
    c.set_bc(0, BCT.fixedValueCell, "all", (0.03, 0.0, 0.0))
    c.set_bc(1, BCT.slip)
    c.turbulence.bc_a[1] = BC(BCT.slip)
    run_const("seeded_1d_star_kepsa_vanleer", c, 4, 1e-3, (1, 4),
              restart=seeded_primitives(c, 241), restart_turb=seeded_turbulence(c, 241),
              restart_a=seeded_a(c, 241))


def turbulence_source_case(base, nsteps=6, keep=(1, 3, 6)):
    """"Courant-Source" with the shear model's source step (shearModel.cpp:
    107-150): epsilon over its growth rate caps the next Courant step.  The
    oracle finds the step sequence; the reference must take the same steps."""
    from schemi_amd.case import Turbulence
    c = Case(dim=1, N=(64, 1, 1), L=(1.0, 1.0, 1.0), limiter=TL.vanLeer, courant=0.3,
             min_time=1e-6, diffusion=True, transport=(2e-3, 2e-3, 2e-3),
             timestep_source=True, source_coeffs=(0.02, 0.5),
             turbulence=Turbulence(), **base)
    init, tinit = seeded_primitives(c, 207), seeded_turbulence(c, 207)
    o = Oracle(c)
    o.init_primitive(*init, k=tinit[0], eps=tinit[1])
    t, dt, dts, times = 0.0, c.min_time, [], []
    for _ in range(nsteps):
        assert o.step(dt, 0) == 0
        dts.append(dt)
        t = t + dt
        times.append(t)
        dt = o.courant_dt(c.courant, c.min_time)
    with tempfile.TemporaryDirectory(dir=os.path.join(ROOT, "oracle", "_ref")) as root:
        decks.write_deck(root, c, t_end=t, restart=init, restart_turb=tinit)
        r = decks.run(root)
        if r.returncode != 0:
            raise RuntimeError(f"reference failed\n{r.stdout[-2000:]}\n{r.stderr}")
        tt = decks.read_times(root)
        if tt.shape[0] != nsteps + 2:
            raise RuntimeError(f"reference took {tt.shape[0] - 2} steps, expected {nsteps}")
        snaps = {s: decks.read_output(root, s + 1, c.Ns) for s in keep}
        ref_times = tt[2:, 1]
    print("dts", dts, "reference times", ref_times)
    save("seeded_1d_turb_courant_source", c, dts, snaps, times=ref_times, restart=init,
         restart_turb=tinit)


if __name__ == "__main__":
    b = dict(M=[29.0, 4.0], Cv=[2.5, 1 / 0.67])
    if len(sys.argv) > 1 and sys.argv[1] == "lz":
        lz_cases(b)
    elif len(sys.argv) > 1 and sys.argv[1] == "vdw":
        vdw_cases(b)
    elif len(sys.argv) > 1 and sys.argv[1] == "hancock":
        hancock_cases(b)
    elif len(sys.argv) > 1 and sys.argv[1] == "rk":
        rk_cases(b)
    elif len(sys.argv) > 1 and sys.argv[1] == "stiffened":
        stiffened_cases(b)
    elif len(sys.argv) > 1 and sys.argv[1] == "kataoka":
        kataoka_cases(b)
    elif len(sys.argv) > 1 and sys.argv[1] == "diffusion":
        diffusion_cases(b)
    elif len(sys.argv) > 1 and sys.argv[1] == "jacobi":
        jacobi_cases(b)
    elif len(sys.argv) > 1 and sys.argv[1] == "gs":
        gauss_seidel_cases(b)
    elif len(sys.argv) > 1 and sys.argv[1] == "molmass":
        species_flux_cases(b)
    elif len(sys.argv) > 1 and sys.argv[1] == "vdwdiff":
        vdw_diffusion_cases(b)
    elif len(sys.argv) > 1 and sys.argv[1] == "source":
        courant_source_cases(b)
    elif len(sys.argv) > 1 and sys.argv[1] == "turbulence":
        turbulence_cases(b)
        turbulence_source_case(b)
    elif len(sys.argv) > 1 and sys.argv[1] == "kepsa":
        kepsa_cases(b)
    elif len(sys.argv) > 1 and sys.argv[1] == "star":
        star_cases(b)
    elif len(sys.argv) > 1 and sys.argv[1] == "bhr":
        bhr_cases(b)
    else:
        main()
