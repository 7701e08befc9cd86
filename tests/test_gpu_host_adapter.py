"""The C++ drop-in (createStepSolver("SecondOrderRK_GPU" / "SecondOrder_GPU"),
gpuStepSolverRK) driven by the reference's time loop, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "schemi_amd", "host", "drive_shock_tube")


@pytest.mark.gpu
@pytest.mark.parametrize("key,step_solver", [("SecondOrderRK_GPU", 1), ("SecondOrder_GPU", 0)])
def test_cpp_adapter_shock_tube(hip_device, key, step_solver):
    from oracle.oracle import Oracle
    from schemi_amd import cases

    if not os.path.exists(DRIVER):
        pytest.fail("schemi_amd/host/drive_shock_tube is not built")
    nsteps = 50
    r = subprocess.run([DRIVER, str(nsteps), key], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = dict(line.split() for line in r.stdout.strip().splitlines())
    case, (c, v, p) = cases.shock_tube_1d(1000)
    case.step_solver = step_solver
    o = Oracle(case)
    o.init_primitive(c, v, p)
    dt, t = case.min_time, 0.0
    for _ in range(nsteps):
        assert o.step(dt, 1) == 0
        t = t + dt
        dt = o.courant_dt(case.courant, case.min_time)
    st = o.state()
    assert float(got["time"]) == t
    assert int(got["steps"]) == nsteps
    assert np.isclose(float(got["density0"]), st["rho0"].sum(), rtol=1e-13)
    assert np.isclose(float(got["density1"]), st["rho"][0].sum(), rtol=1e-13)
    assert np.isclose(float(got["density2"]), st["rho"][1].sum(), rtol=1e-13)
    assert np.isclose(float(got["totalEnergy"]), st["E"].sum(), rtol=1e-13)


def test_cpp_adapter_rejects_unknown_key():
    if not os.path.exists(DRIVER):
        pytest.skip("schemi_amd/host/drive_shock_tube is not built")
    # the key is checked before any device call
    r = subprocess.run([DRIVER, "1", "ThirdOrder_GPU"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode != 0
    assert "Unknown or CPU-only step solver" in r.stderr
