"""The drop-in: the reference's own, unmodified main loop (its deck reader,
initialisation, time step, output) running the GPU step through its
step-solver factory (oracle/_ref/schemi_gpu_dropin: the reference's objects,
schemi_amd/host/gpuStepSolverRK.cpp and libschemi_gpu.so, the key
registered by tests/host/dropin_wrap.cpp), against the reference's CPU run of
the same deck (the golden fixtures).

The GPU step works on the conserved state, the reference additionally on
separately averaged concentrations (last-bit differences that drift step by
step), so the bars are those of test_gpu_step.test_golden_fixtures: the
first step 1e-12 per cell (no floor), later snapshots 1e-10 per cell with
golden_io.LATER_FLOORS and 1e-12 of each field's largest magnitude.  With
"Courant-Source" the main loop forms its time step from the GPU diffusion
stage's source step (written to mesh::timestepSourceRef() by the adapter):
the times must be the reference's.  Turbulent decks (shear, zero, kEpsA and
BHR models, #Reconstruction star) carry rho k, rho epsilon, rho a and rho b
through the adapter; the k, epsilon, a and b the loop writes follow them to
the same bars."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from tests.golden_io import LATER_FLOORS, conserved, field_rel, load, per_cell_rel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "oracle", "_ref", "schemi_gpu_dropin")
REF = os.path.join(ROOT, "oracle", "_ref", "schemi")

# fixtures of restart decks with the ideal mixture (the drop-in maps it) and
# a constant step, plus the Courant-Source run
CASES = ["seeded_1d_vanLeer", "seeded_1d_hll_slip_fixed", "seeded_1d_hancock_minmod",
         "seeded_2d_rt_vanleer", "seeded_2d_linear_hll", "seeded_3d_rm_vanalbada",
         "seeded_1d_diffusion_vanleer", "seeded_2d_jacobi_hll", "seeded_1d_courant_source",
         # turbulence (shear, zero, kEpsA, BHR) and #Reconstruction star
         "seeded_1d_turb_shear_vanleer", "seeded_2d_turb_zero_hllc",
         "seeded_1d_turb_courant_source", "seeded_2d_kepsa_hllc", "seeded_3d_bhr_vanalbada",
         "seeded_2d_bhr_star_hll", "seeded_2d_star_hllc", "seeded_1d_star_kepsa_vanleer"]


def _run_dropin(name, key_suffix="_GPU", binary=DROPIN):
    from oracle import decks

    g = load(name)
    case = g["case"]
    with tempfile.TemporaryDirectory() as root:
        t = 0.0
        for dt in g["dts"]:
            t = t + dt
        const = g["dts"][0] if "times" not in g else -1.0
        t_end = g["times"][-1] if "times" in g else t
        decks.write_deck(root, case, t_end=t_end, restart=g["restart"], const_dt=const,
                         restart_turb=g.get("restart_turb"), restart_a=g.get("restart_a"),
                         restart_b=g.get("restart_b"))
        main = os.path.join(root, "set", "main.txt")
        txt = open(main).read()
        order = "SecondOrderRK" if case.step_solver == 1 else "SecondOrder"
        if binary == DROPIN:
            txt = txt.replace(f"#Approximation_order\n{order}\n",
                              f"#Approximation_order\n{order}{key_suffix}\n")
        open(main, "w").write(txt)
        r = subprocess.run([binary], cwd=root, stdin=subprocess.DEVNULL, capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
        snaps = {s: decks.read_output(root, s + 1, case.Ns) for s in g["snaps"]}
        times = decks.read_times(root)
    return g, snaps, times


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_reference_main_loop_runs_the_gpu_step(name, hip_device):
    if not os.path.exists(DROPIN):
        pytest.fail("oracle/_ref/schemi_gpu_dropin is not built (make -C oracle dropin)")
    g, snaps, times = _run_dropin(name)
    case = g["case"]
    for s, st in snaps.items():
        got, ref = conserved(st), conserved(g["snaps"][s])
        if case.turbulence is not None:
            # k, epsilon (and a, BHR's b) as the loop writes them: per cell,
            # 1e-12 after one step, 1e-10 later; a's components with a floor
            # of 1e-4 of their largest magnitude (they pass through zero)
            bar = 1e-12 if s == 1 else 1e-10
            for k in ("k", "eps"):
                r = g["snaps"][s][k]
                e = float((np.abs(st[k] - r) / np.abs(r)).max())
                assert e <= bar, (name, s, k, e)
            if case.turbulence.a_field:
                # BHR's b is part of the state; kEpsA's is the diffusion
                # stage's algebraic b (clipped at 0: floor 1e-4 of its maximum)
                r = g["snaps"][s]["b"]
                fl = max(1e-4 * float(np.abs(r).max()), 1e-300)
                e = float((np.abs(st["b"] - r) / np.maximum(np.abs(r), fl)).max())
                assert e <= bar, (name, s, "b", e)
            if case.turbulence.a_field:
                for c in range(case.dim):
                    r = g["snaps"][s]["a"][c]
                    fl = max(1e-4 * float(np.abs(r).max()), 1e-300)
                    e = float((np.abs(st["a"][c] - r) / np.maximum(np.abs(r), fl)).max())
                    assert e <= bar, (name, s, "a", c, e)
        if s == 1:
            for q, err in enumerate(per_cell_rel(got, ref, case.Ns)):
                assert err <= 1e-12, f"{name} step 1 field {q}: {err:.3e}"
        else:
            cell = per_cell_rel(got, ref, case.Ns, LATER_FLOORS["rho_k"], LATER_FLOORS["m"])
            for q, (e1, e2) in enumerate(zip(cell, field_rel(got, ref))):
                assert e1 <= 1e-10 and e2 <= 1e-12, f"{name} step {s} field {q}: {e1} {e2}"
        # the primitive fields the loop reads back (p, T, v) follow the state
        for k in ("p", "T"):
            e = np.abs(st[k] - g["snaps"][s][k]).max() / np.abs(g["snaps"][s][k]).max()
            assert e <= 1e-12, (name, s, k, e)
    if "times" in g:
        # the main loop's own time steps, Courant-Source with the GPU's source
        # step (the plain Courant step differs by 1.5x here: test_oracle_golden.
        # test_courant_source_fixture_discriminates); the loop forms them from
        # the fields, whose last bits follow the conserved-state model
        t_ref = g["times"]
        assert times.shape[0] - 2 == t_ref.shape[0]
        assert np.max(np.abs(times[2:, 1] - t_ref) / t_ref) <= 1e-12, (times[2:, 1], t_ref)


@pytest.mark.gpu
def test_reference_main_loop_runs_the_mixed_gpu_step(hip_device):
    """The "_GPU_MIXED" key: the mixed-precision path under the same loop,
    within the mixed path's 1e-5 relative L1 of the reference's run."""
    if not os.path.exists(DROPIN):
        pytest.fail("oracle/_ref/schemi_gpu_dropin is not built (make -C oracle dropin)")
    g, snaps, _ = _run_dropin("seeded_3d_rm_vanalbada", "_GPU_MIXED")
    for s, st in snaps.items():
        got, ref = conserved(st), conserved(g["snaps"][s])
        l1 = np.abs(got - ref).sum(axis=1) / np.maximum(np.abs(ref).sum(axis=1), 1e-300)
        live = np.abs(ref).sum(axis=1) > 0
        assert np.all(l1[live] <= 1e-5), l1


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["seeded_1d_turb_shear_vanleer", "seeded_2d_kepsa_hllc",
                                  "seeded_3d_bhr_vanalbada", "seeded_2d_bhr_star_hll"])
def test_dropin_turbulent_viscosity(name, hip_device):
    """The turbulent viscosity the loop writes (the phase's transport
    coefficients, formed by the diffusion stage from the k and epsilon it
    starts from) against the reference solver's run of the same deck
    (oracle/_ref/schemi), per cell."""
    for b in (DROPIN, REF):
        if not os.path.exists(b):
            pytest.fail(f"{b} is not built (make -C oracle ref dropin)")
    g, snaps, _ = _run_dropin(name)
    _, ref, _ = _run_dropin(name, binary=REF)
    for s, st in snaps.items():
        bar = 1e-12 if s == 1 else 1e-10
        r = ref[s]["nut"]
        assert np.abs(r).max() > 0, (name, s)
        e = float((np.abs(st["nut"] - r) / np.maximum(np.abs(r), 1e-300)).max())
        assert e <= bar, (name, s, e)
