"""The host step core driven by a copy of the reference's time loop
(schemi_amd/host/drive_shock_tube: upload once, step, download after every
step, Courant step from the device) against the oracle: the final conserved
state bit for bit, field by field, and the same time."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "schemi_amd", "host", "drive_shock_tube")


@pytest.mark.gpu
@pytest.mark.parametrize("key,step_solver", [("SecondOrderRK_GPU", 1), ("SecondOrder_GPU", 0)])
def test_host_core_shock_tube(hip_device, key, step_solver):
    from oracle.oracle import Oracle
    from schemi_amd import cases

    if not os.path.exists(DRIVER):
        pytest.fail("schemi_amd/host/drive_shock_tube is not built")
    nsteps = 50
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "state.bin")
        r = subprocess.run([DRIVER, str(nsteps), key, out], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        got = dict(line.split() for line in r.stdout.strip().splitlines())
        U = np.fromfile(out, dtype=np.float64).reshape(7, 1000)
    case, (c, v, p) = cases.shock_tube_1d(1000)
    case.step_solver = step_solver
    o = Oracle(case)
    o.init_primitive(c, v, p)
    dt, t = case.min_time, 0.0
    for _ in range(nsteps):
        assert o.step(dt, 1) == 0
        t = t + dt
        dt = o.courant_dt(case.courant, case.min_time)
    assert float(got["time"]) == t
    assert int(got["steps"]) == nsteps
    ref = o.conserved()
    for q in range(ref.shape[0]):
        assert np.array_equal(U[q], ref[q]), q


def test_host_core_rejects_unknown_key():
    if not os.path.exists(DRIVER):
        pytest.skip("schemi_amd/host/drive_shock_tube is not built")
    # the key is checked before any device call
    r = subprocess.run([DRIVER, "1", "ThirdOrder_GPU", os.devnull], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode != 0
    assert "Unknown or CPU-only step solver" in r.stderr
