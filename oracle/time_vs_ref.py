"""Time the oracle's explicit operators against the reference's own.

The reference side is MPS_CAL_v4.0's CPU solver compiled in place into
``oracle/_ref/ref_v4`` (``make -C oracle ref``); its ``--time`` mode loads the
scene through the reference's readers (MPS_CAL_v4.0/MPS.cpp:62-300), runs
``MPS_CPU::Initial`` and times ``cal_n`` (MPS_CPU.cpp) and the per-particle
``d3_LapVel`` / ``d3_GradPress`` over every particle.  The oracle side times
``pnd``, ``lapvel`` and ``gradp_vec`` (oracle/mps_oracle.c) on the same scene
with the same OpenMP thread count.  The ratio relates the oracle's CPU
baseline, which is what the benchmark can time next to the GPU, to the
reference's own CPU code.

Runs only where the reference tree is mounted (it is not on the GPU host).

    python oracle/time_vs_ref.py [--threads T] [--reps R]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
REF_BIN = os.path.join(HERE, "_ref", "ref_v4")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count())
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args()
    # the OpenMP runtime reads this when the oracle library first runs a parallel loop
    os.environ["OMP_NUM_THREADS"] = str(a.threads)
    sys.path.insert(0, REPO)
    sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
    import numpy as np

    import make_golden as mg
    import oracle

    if not os.path.exists(REF_BIN):
        sys.exit("oracle/_ref/ref_v4 is missing: run make -C oracle ref (needs the reference tree)")

    # the config-0 dam break: 20x20x40 fluid in an 80x20x50 open-top tank, 2+2 layers
    prm = dict(dim=3, rho=1000.0, nu=1e-6, dp=0.01, re_n_mult=2.1, re_lap_mult=4.0, beta=0.97,
               cs=20.0, cfl=0.2, col_mult=0.5, col_rate=0.2, dt_max=1e-4, dt_min_mult=1e-3,
               tt=1.0, tout=0.1, max_dis=0.001)
    nb2, nb1, pos = mg.lattice_dam((20, 20, 40), (80, 20, 50), 2, 2, 0.01, 0, 0.01, 3)
    pos = pos.astype(np.float32).astype(np.float64)   # what the reference's %f reader keeps
    n = len(pos)
    rng = np.random.default_rng(1)
    vel = (0.05 * rng.standard_normal((n, 3))).astype(np.float32).astype(np.float64)
    p = (1000 * 9.81 * np.maximum(0.4 - pos[:, 2], 0.0)).astype(np.float32).astype(np.float64)

    with tempfile.TemporaryDirectory() as td:
        mg.write_input(os.path.join(td, "input.dat"), prm)
        mg.write_geo(os.path.join(td, "GEO.dat"), nb2, nb1, pos, vel, p)
        r = subprocess.run([REF_BIN, td, "--time", str(a.reps)], capture_output=True, text=True,
                           env=dict(os.environ, OMP_NUM_THREADS=str(a.threads)))
        if r.returncode != 0:
            sys.exit(r.stdout[-2000:] + r.stderr[-2000:])
        ref = json.loads([ln for ln in r.stdout.splitlines() if ln.startswith("{")][-1])

    P = oracle.params(f32=0, **{k: v for k, v in prm.items() if k not in ("tt", "tout")})
    x, y, z = pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy()
    G = oracle.grid_init(P, x, y, z)
    types = np.array([2] * nb2 + [1] * nb1 + [0] * (n - nb2 - nb1), np.int32)
    key = oracle.hash_cells(P, G, x, y, z, types)
    perm, _, cs = oracle.sort_cells(key, G.ncells)
    st = {"x": x[perm], "y": y[perm], "z": z[perm], "u": vel[perm, 0].copy(), "v": vel[perm, 1].copy(),
          "w": vel[perm, 2].copy(), "type": types[perm]}
    ps = p[perm]
    _, n0, lam, _, _ = oracle.calonce(P, G, st["x"], st["y"], st["z"], st["type"], cs)
    coef = 2.0 * 3 / (n0 * lam)

    def best(f):
        f()   # first call outside the timing (thread start-up)
        ts = []
        for _ in range(a.reps):
            t0 = time.perf_counter()
            f()
            ts.append((time.perf_counter() - t0) * 1e3)
        return sum(ts) / len(ts)

    orc = {
        "pnd_ms": best(lambda: oracle.pnd(P, G, st["x"], st["y"], st["z"], st["type"], cs)),
        "lapvel_ms": best(lambda: oracle.lapvel(P, G, st, cs, coef)),
        "gradp_ms": best(lambda: oracle.gradp_vec(P, G, st, ps, cs, n0)),
    }
    out = {"scene": f"dam break, 16,000 fluid / {n} particles, l0 = 10 mm, re 2.1 / 4.0 l0",
           "threads": a.threads, "reps": a.reps, "reference": ref, "oracle": orc,
           "oracle_over_reference": {
               "n": orc["pnd_ms"] / ref["cal_n_ms"],
               "lapvel": orc["lapvel_ms"] / ref["lapvel_ms"],
               "gradp": orc["gradp_ms"] / ref["gradpress_ms"]}}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
