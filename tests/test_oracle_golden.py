"""The oracle against the reference's own outputs (golden fixtures), bit for bit.

The fixtures come from the compiled reference (tests/golden/make_golden.py):
the 1-D two-component shock tube with constant and Courant steps, and seeded
1-D/2-D/3-D states for every limiter, both fluxes, gravity, slip and
fixed-value boundaries.
"""
import dataclasses

import numpy as np
import pytest

from oracle.oracle import Oracle
from tests.golden_io import KEYS, load, names


def start(g):
    """Oracle at the fixture's initial state (restart fields, with k and
    epsilon when the case is turbulent and a with the kEpsA model, or the
    reference's output 0)."""
    o = Oracle(g["case"])
    if "restart" in g:
        kw = {}
        if "restart_turb" in g:
            kw = dict(k=g["restart_turb"][0], eps=g["restart_turb"][1])
        if "restart_a" in g:
            kw["a"] = g["restart_a"]
        if "restart_b" in g:
            kw["b"] = g["restart_b"]
        o.init_primitive(*g["restart"], **kw)
    else:
        o.set_full(g["snaps"][0])
    return o


@pytest.mark.parametrize("name", names())
def test_oracle_matches_reference(name):
    g = load(name)
    case = g["case"]
    o = start(g)
    last = max(g["snaps"])
    t = 0.0
    for s, dt in enumerate(g["dts"], start=1):
        assert o.step(dt, 0) == 0
        t = t + dt
        if "times" in g:
            assert t == g["times"][s - 1]
        if s in g["snaps"]:
            st = o.state()
            for k in g["snaps"][s]:
                assert np.array_equal(st[k], g["snaps"][s][k]), f"{name} step {s}: {k}"
        if s == last:
            break


@pytest.mark.parametrize("name", ["shock_tube_1d_cfl", "seeded_1d_courant_source",
                                  "seeded_1d_turb_courant_source"])
def test_courant_sequence_matches_reference(name):
    """The oracle's own step sequence reproduces the times the reference
    stepped through (Courant; Courant-Source with the enthalpy-flow or the
    turbulence model's source step capping it)."""
    g = load(name)
    case = g["case"]
    o = start(g)
    dt, t = case.min_time, 0.0
    for s in range(len(g["dts"])):
        assert dt == g["dts"][s]
        assert o.step(dt, 0) == 0
        t = t + dt
        assert t == g["times"][s]
        dt = o.courant_dt(case.courant, case.min_time)


def test_courant_source_fixture_discriminates():
    """The same run with the plain Courant step leaves the reference's times
    after the first step: the fixture pins the source cap and its order with
    #Minimal_timestep (the last step's source step is below it)."""
    g = load("seeded_1d_courant_source")
    case = dataclasses.replace(g["case"], timestep_source=False)
    o = Oracle(case)
    o.init_primitive(*g["restart"])
    assert o.step(case.min_time, 0) == 0
    dt = o.courant_dt(case.courant, case.min_time)
    assert dt > 1.5 * g["dts"][1]
    assert g["dts"][-1] == case.min_time
