"""Loading of the golden fixtures written by tests/golden/make_golden.py."""
import glob
import json
import os

import numpy as np

from schemi_amd import BC, Case

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KEYS = ("c0", "c", "rho0", "rho", "v", "p", "U", "T", "m", "E")


def names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    d = json.loads(str(z["case"]))
    case = Case(dim=d["dim"], N=tuple(d["N"]), L=tuple(d["L"]), M=d["M"], Cv=d["Cv"],
                R=d["R"], flux=d["flux"], limiter=d["limiter"], gravity=d["gravity"],
                g=tuple(d["g"]), courant=d["courant"], min_time=d["min_time"],
                bc_c=[[BC(t, tuple(v)) for t, v in row] for row in d["bc_c"]],
                bc_v=[BC(t, tuple(v)) for t, v in d["bc_v"]],
                bc_p=[BC(t, tuple(v)) for t, v in d["bc_p"]],
                gas_model=d.get("gas_model", 0), Tc=d.get("Tc", ()), Pc=d.get("Pc", ()),
                p0=d.get("p0", ()), gammas=d.get("gammas", ()),
                epsLJ=d.get("epsLJ", ()), sigmaLJ=d.get("sigmaLJ", ()),
                kataoka_b=tuple(d.get("kataoka_b", ("critical", 1.0))),
                step_solver=d.get("step_solver", 1),
                diffusion=bool(d.get("diffusion", False)),
                transport=tuple(d.get("transport", (2e-5, 2e-5, 2e-2))),
                sle_solver=d.get("sle_solver", "explicit"),
                sle_iters=int(d.get("sle_iters", 2000)),
                enthalpy_flow=bool(d.get("enthalpy_flow", False)),
                molar_mass_correction=bool(d.get("molar_mass_correction", False)),
                reconstruction=d.get("reconstruction", "linear"),
                timestep_source=bool(d.get("timestep_source", False)),
                source_coeffs=tuple(d.get("source_coeffs", (0.25, 0.5))),
                turbulence=_turbulence(d.get("turbulence")))
    keys = KEYS + (("k", "eps") if case.turbulence is not None else ())
    if case.turbulence is not None and case.turbulence.a_field:
        keys += ("a", "b")
    snaps = {int(s): {k: z[f"s{s}_{k}"] for k in keys} for s in z["steps"]}
    out = {"case": case, "dts": z["dts"], "snaps": snaps}
    if "times" in z:
        out["times"] = z["times"]
    if "init_c" in z:
        out["restart"] = (z["init_c"], z["init_v"], z["init_p"])
    if "init_k" in z:
        out["restart_turb"] = (z["init_k"], z["init_eps"])
    if "init_a" in z:
        out["restart_a"] = z["init_a"]
    if "init_b" in z:
        out["restart_b"] = z["init_b"]
    return out


def _turbulence(t):
    if t is None:
        return None
    from schemi_amd.case import Turbulence
    p = t["params"]
    return Turbulence(model=t["model"], Cmu=p[0], sigmaSc=p[1], sigmaT=p[2], sigmaE=p[3],
                      sigmaK=p[4], sigmaEps=p[5], CMS_R=p[6], CMS_D=p[7], mink=p[8],
                      mineps=p[9], C=tuple(t["C"]),
                      bc_k=[BC(ty, tuple(v)) for ty, v in t["bc_k"]],
                      bc_eps=[BC(ty, tuple(v)) for ty, v in t["bc_eps"]],
                      sigmaa=t.get("sigmaa", 1.4), sigmab=t.get("sigmab", 0.7),
                      minb=t.get("minb", 1e-15),
                      bc_b=[BC(ty, tuple(v)) for ty, v in t.get("bc_b", [[1, [0, 0, 0]]] * 6)],
                      bc_a=[BC(ty, tuple(v)) for ty, v in t.get("bc_a", [[1, [0, 0, 0]]] * 6)])


def conserved(st):
    return np.concatenate([st["rho"], st["rho0"][None], st["m"], st["E"][None]])


# Parity bars of the FP64 path, per conserved field (rows rho_1..rho_Ns, rho,
# m_x, m_y, m_z, E):
#  * one step from a common state: max over cells of |got - ref| / |ref| <= 1e-12,
#    no floor (per_cell_rel with the default floors);
#  * after several steps against the reference's outputs, whose separately
#    averaged concentrations differ from the conserved-state model in the last
#    bits and drift apart step by step: per cell relative to max(|ref|, floor)
#    with floors LATER_FLOORS (fractions of the field's largest magnitude:
#    partial densities can be exactly zero where a species is absent, momentum
#    components pass through zero) <= 1e-10, and the field-normalised maximum
#    <= 1e-12.
LATER_FLOORS = {"rho_k": 1e-15, "m": 1e-4}


def per_cell_rel(got, ref, ns, rho_k_floor=0.0, m_floor=0.0):
    """Per-field maxima over cells of |got - ref| / max(|ref|, floor), the
    floor a fraction of the field's largest magnitude (rho_k_floor for the
    partial densities, m_floor for the momentum components, none for the
    total density and the energy)."""
    out = []
    for q in range(ref.shape[0]):
        frac = rho_k_floor if q < ns else (m_floor if ns + 1 <= q <= ns + 3 else 0.0)
        floor = frac * float(np.abs(ref[q]).max())
        den = np.maximum(np.abs(ref[q]), max(floor, 1e-300))
        out.append(float((np.abs(got[q] - ref[q]) / den).max()))
    return out


def field_rel(got, ref):
    """Per-field maxima of |got - ref| normalised by the field's largest magnitude."""
    return [float(np.abs(got[q] - ref[q]).max() / max(np.abs(ref[q]).max(), 1e-300))
            for q in range(ref.shape[0])]
