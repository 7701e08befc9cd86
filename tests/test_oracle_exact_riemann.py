"""The oracle's two-gas shock tube against the exact Riemann solution
(tests/exact_riemann.py, Toro ch. 4, each wave family with the gamma of its
gas): the L1 error of density, pressure and velocity falls with refinement
at the order a second-order TVD scheme reaches on a solution with a shock
and a contact (about 0.7-0.9)."""
import numpy as np

from oracle.oracle import Oracle
from schemi_amd import cases
from tests.exact_riemann import sample, star_state

LEFT = (1.0, 0.0, 1.0, 1.4)      # species A: rho, u, p, gamma
RIGHT = (0.125, 0.0, 0.1, 1.67)  # species B


def test_exact_star_state_textbook():
    # Sod's problem (Toro table 4.2, test 1): p* = 0.30313, u* = 0.92745
    ps, us = star_state((1.0, 0.0, 1.0, 1.4), (0.125, 0.0, 0.1, 1.4))
    assert abs(ps - 0.30313) < 1e-5 and abs(us - 0.92745) < 1e-5


def run_to(N, t_end=0.2):
    case, (c, v, p) = cases.shock_tube_1d(N)
    o = Oracle(case)
    o.init_primitive(c, v, p)
    t, dt = 0.0, case.min_time
    while t < t_end:
        if t + dt > t_end:
            dt = t_end - t
        assert o.step(dt, 0) == 0
        t = t + dt
        dt = o.courant_dt(case.courant, case.min_time)
    return o.state(), t


def test_shock_tube_converges_to_exact_solution():
    errs = []
    for N in (100, 200, 400):
        st, t = run_to(N)
        x = (np.arange(N) + 0.5) / N
        rho, u, p, _ = sample(LEFT, RIGHT, x, t, 0.5)
        errs.append((np.abs(st["rho0"] - rho).mean(), np.abs(st["p"] - p).mean(),
                     np.abs(st["v"][0] - u).mean()))
    errs = np.array(errs)
    order = np.log2(errs[:-1] / errs[1:])
    print("L1 (rho, p, u):", errs.tolist(), "orders:", order.tolist())
    assert np.all(order > 0.5) and np.all(order < 1.3)
    assert errs[-1, 0] < 5e-3 and errs[-1, 1] < 5e-3 and errs[-1, 2] < 1.2e-2
