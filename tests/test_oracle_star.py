"""The oracle's #Reconstruction star beyond the reference fixtures.

The fixtures tests/golden/seeded_*_star_* pin the oracle's star
reconstruction to the reference bit for bit (test_oracle_golden.py).  These
tests show that the star states are what those fixtures exercise: the same
runs with linear reconstruction leave every fixture but one, and that one is
the MUSCL-Hancock step under gravity, where the reference replaces the star
velocity by the interpolated velocity (secondOrderStepSolver.cpp:74-76), so
that a laminar run with star reconstruction equals the linear one bit for
bit.
"""
import dataclasses

import numpy as np
import pytest

from oracle.oracle import Oracle
from tests.golden_io import load

STAR = ["seeded_2d_star_hllc", "seeded_2d_star_vdw_hll", "seeded_1d_star_vanleer_hll",
        "seeded_3d_star_vanalbada", "seeded_2d_star_gravity", "seeded_2d_star_hancock",
        "seeded_2d_star_jacobi", "seeded_2d_star_turb_shear_hllc",
        "seeded_2d_star_kepsa_vdw_hll", "seeded_1d_star_kepsa_vanleer"]


def run(g, case):
    o = Oracle(case)
    kw = {}
    if "restart_turb" in g:
        kw = dict(k=g["restart_turb"][0], eps=g["restart_turb"][1])
    if "restart_a" in g:
        kw["a"] = g["restart_a"]
    if "restart_b" in g:
        kw["b"] = g["restart_b"]
    o.init_primitive(*g["restart"], **kw)
    last = max(g["snaps"])
    for dt in g["dts"][:last]:
        assert o.step(dt, 0) == 0
    return o.state(), g["snaps"][last]


@pytest.mark.parametrize("name", STAR)
def test_linear_reconstruction_leaves_the_fixture(name):
    g = load(name)
    assert g["case"].reconstruction == "star"
    got, ref = run(g, dataclasses.replace(g["case"], reconstruction="linear"))
    rel = float(np.abs(got["T"] - ref["T"]).max() / np.abs(ref["T"]).max())
    assert rel > 1e-9, rel


def test_hancock_gravity_star_is_linear():
    g = load("seeded_2d_star_hancock_gravity")
    got, ref = run(g, dataclasses.replace(g["case"], reconstruction="linear"))
    for key in ("rho", "m", "E", "T", "p"):
        assert np.array_equal(got[key], ref[key]), key
