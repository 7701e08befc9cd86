"""FP64 GPU step against the oracle, bit for bit (conserved-state mode)."""
import numpy as np
import pytest

from schemi_amd.gpu import GpuSolver, conserved_from_primitive
from tests.cases_small import cases, random_primitives

CASES = cases()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_fp64_step_bitexact(name, hip_device):
    from oracle.oracle import Oracle

    case = CASES[name]
    c, v, p = random_primitives(case, 7)
    U0 = conserved_from_primitive(case, c, v, p)
    o = Oracle(case)
    Ns = case.Ns
    o.set_conserved(U0[:Ns], U0[Ns], U0[Ns + 1:Ns + 4], U0[Ns + 4])
    g = GpuSolver(case)
    g.set_conserved(U0)
    dt = 2e-4
    for s in range(3):
        assert o.step(dt, 1) == 0
        g.step(dt)
        ref = o.conserved()
        got = g.conserved()
        diff = np.abs(got - ref)
        print(name, s, "max abs diff", diff.max(), "n differing", int((got != ref).sum()))
        assert np.array_equal(got, ref), f"{name}: step {s + 1} differs"
