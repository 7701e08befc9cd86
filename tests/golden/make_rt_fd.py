"""Writes tests/golden/rt_fd_oracle_256x1024.json: central finite differences
of the mixing width J = sum_i V_i Y (1 - Y) of the 2-D Rayleigh-Taylor
gradient workload (256 x 1024, periodic x, slip y, HLLC + van Leer, gravity
-1; 16 seeded single-mode amplitudes a_k in [0, 0.005], seed 2024; 400 steps
of dt = 2.5e-4) computed with the oracle, the C restatement of the
reference's step (conserved-state mode), for modes 1, 4, 8 and 16 at
h = 1e-6 and h = 1e-8.  CPU only; about half an hour on 8 cores.

    python tests/golden/make_rt_fd.py [workers]
"""
import json
import os
import sys
import time
from multiprocessing import Pool

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

NSTEPS, DT, SEED = 400, 2.5e-4, 2024
MODES = (0, 3, 7, 15)
HS = (1e-6, 1e-8)
OUT = os.path.join(ROOT, "tests", "golden", "rt_fd_oracle_256x1024.json")


def amplitudes():
    import numpy as np

    return np.random.default_rng(SEED).uniform(0.0, 0.005, 16)


def oracle_J(delta):
    """J after NSTEPS oracle steps from the state with a + delta (mode, h)."""
    import numpy as np
    import torch

    from oracle.oracle import Oracle
    from schemi_amd import adjoint as adj

    case = adj.rt_case(256, 1024)
    a = amplitudes()
    if delta is not None:
        k, h = delta
        a = a.copy()
        a[k] += h
    U0 = adj.rt_modes_state(case, torch.as_tensor(a)).numpy()
    ns = case.Ns
    o = Oracle(case)
    o.set_conserved(U0[:ns], U0[ns], U0[ns + 1:ns + 4], U0[ns + 4])
    for _ in range(NSTEPS):
        if o.step(DT, 1) != 0:
            raise RuntimeError("oracle step failed")
    U = o.conserved()
    _, V = o.cell_geometry()
    Y = U[0] / U[ns]
    return float(np.sum(V * Y * (1 - Y)))


def main():
    workers = int(sys.argv[1]) if len(sys.argv) > 1 else 7
    jobs = [None] + [(k, s * h) for k in MODES for h in HS for s in (1, -1)]
    t0 = time.time()
    with Pool(workers) as pool:
        Js = pool.map(oracle_J, jobs, chunksize=1)
    res = {"workload": "2-D Rayleigh-Taylor 256x1024, periodic x, slip y, HLLC + van Leer, "
                       "gravity -1, 16 modes a_k ~ U[0, 0.005] (numpy default_rng(2024)), "
                       "400 steps of dt 2.5e-4, J = sum V Y (1 - Y) of species 0",
           "nsteps": NSTEPS, "dt": DT, "seed": SEED, "J": Js[0], "fd": {}}
    for k in MODES:
        for h in HS:
            jp = Js[jobs.index((k, h))]
            jm = Js[jobs.index((k, -h))]
            res["fd"][f"{k}:{h:g}"] = (jp - jm) / (2 * h)
    res["cpu_seconds_wall"] = time.time() - t0
    with open(OUT, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
