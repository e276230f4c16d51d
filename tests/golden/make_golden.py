"""Write golden vectors from the reference's own MPS_CAL_v4.0 CPU solver.

Runs only in the build container, where the reference tree is mounted.  It
compiles the reference's CPU sources in place (``make -C oracle ref`` ->
``oracle/_ref/ref_v4``, which is git-ignored and never shipped), feeds each
scene through the reference's own ``input.dat`` / ``GEO.dat`` readers
(MPS_CAL_v4.0/MPS.cpp:62-300), and stores inputs + outputs as compressed
``.npz`` fixtures next to this script.  The fixtures are data only.

Scenes:
  v4_2d_geo   -- the reference's own MPS_CAL_v4.0 GEO.dat + input.dat (2-D, 2,634 particles)
  v4_3d_dam   -- a small jittered 3-D dam break (3,488 particles) with a smooth
                 velocity field and a hydrostatic-plus-noise pressure field
  v4_2d_jit   -- a 2-D dam break on the same lattice family with jitter

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_ROOT = "/root/reference"
REF_BIN = os.path.join(REPO, "oracle", "_ref", "ref_v4")

LABELS = ["DIMENTION", "FLUID RHO", "FLUID DYNAMICS VISCOSITY", "INITIAL PARTICLE DISTANCE(DP)",
          "R0(times of DP)", "RLAP(times of DP)", "BETA", "NUMERICAL SPEED OF SOUND",
          "CFL NUMBER", "COLLISION DISTANCE(times of DP)", "COLLISION RATE", "MAX TIME STEP",
          "MIN TIME STEP(times of MAX)", "TT TOTAL CALCULATION TIME", "TOUT TIME STEP OF OUTPUT",
          "MAXIMUM DISPLACEMENT OF THE GEOMETRY"]
KEYS = ["dim", "rho", "nu", "dp", "re_n_mult", "re_lap_mult", "beta", "cs", "cfl", "col_mult",
        "col_rate", "dt_max", "dt_min_mult", "tt", "tout", "max_dis"]


def read_input(path):
    lines = open(path).read().splitlines()
    vals = [float(lines[2 * k + 1].split()[0]) for k in range(16)]
    d = dict(zip(KEYS, vals))
    d["dim"] = int(d["dim"])
    return d


def write_input(path, prm):
    with open(path, "w") as f:
        for lab, k in zip(LABELS, KEYS):
            v = prm[k]
            f.write(f"{lab}\n{int(v) if k == 'dim' else repr(float(v))}\n")


def read_geo(path):
    tok = open(path).read().split()
    n, nb2, nb1 = int(tok[0]), int(tok[1]), int(tok[2])
    a = np.array(tok[3:3 + 7 * n], dtype=np.float64)
    pos = a[:3 * n].reshape(n, 3)
    vel = a[3 * n:6 * n].reshape(n, 3)
    p = a[6 * n:7 * n]
    return nb2, nb1, pos, vel, p


def write_geo(path, nb2, nb1, pos, vel, p):
    n = len(p)
    with open(path, "w") as f:
        f.write(f"{n} {nb2} {nb1}\n")
        for arr in (pos, vel):
            f.write(" ".join(repr(float(v)) for v in arr.reshape(-1)))
            f.write("\n")
        f.write(" ".join(repr(float(v)) for v in p))
        f.write("\n")


def lattice_dam(fluid, inner, layers_wall, layers_dummy, dp, seed, jitter, dim):
    """Cubic-lattice dam break: fluid block in the (0,0,0) corner of an open-top tank.
    Order b2 (dummy), b1 (wall), fluid, each x-fastest (MPS_INI/main.cpp:64-94)."""
    L = layers_wall + layers_dummy
    rng = np.random.default_rng(seed)
    nx, ny, nz = inner
    fx, fy, fz = fluid
    ys = range(-L, ny + L) if dim == 3 else [0]
    pts = {0: [], 1: [], 2: []}
    for k in range(-L, nz):
        for j in ys:
            for i in range(-L, nx + L):
                d = max(-i if i < 0 else (i - nx + 1 if i >= nx else 0),
                        -k if k < 0 else 0,
                        (-j if j < 0 else (j - ny + 1 if j >= ny else 0)) if dim == 3 else 0)
                if d == 0:
                    if i < fx and k < fz and (dim == 2 or j < fy):
                        pts[0].append((i, j, k))
                elif d <= layers_wall:
                    pts[1].append((i, j, k))
                else:
                    pts[2].append((i, j, k))
    allp = np.array(pts[2] + pts[1] + pts[0], dtype=np.float64) * dp
    allp += rng.uniform(-jitter, jitter, allp.shape) * dp
    if dim == 2:
        allp[:, 1] = 0.0
    return len(pts[2]), len(pts[1]), allp


def run_ref(prm, nb2, nb1, pos, vel, p):
    with tempfile.TemporaryDirectory() as td:
        write_input(os.path.join(td, "input.dat"), prm)
        write_geo(os.path.join(td, "GEO.dat"), nb2, nb1, pos, vel, p)
        r = subprocess.run([REF_BIN, td], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        # the reference's own reader decides what was read: re-read the scene it saw
        out = parse(os.path.join(td, "ref_out.bin"))
    return out


def parse(path):
    out = {}
    with open(path, "rb") as f:
        while True:
            nm = f.read(32)
            if len(nm) < 32:
                break
            name = nm.split(b"\0")[0].decode()
            kind = np.frombuffer(f.read(4), np.int32)[0]
            cnt = np.frombuffer(f.read(8), np.int64)[0]
            dt = {0: np.float64, 1: np.int32}[int(kind)]
            out[name] = np.frombuffer(f.read(int(cnt) * np.dtype(dt).itemsize), dt).copy()
    return out


def save(name, prm, nb2, nb1, pos, vel, p, ref):
    # the GEO.dat text is what the reference parsed: keep the parsed values
    fn = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(fn, params=json.dumps(prm), nb2=nb2, nb1=nb1, pos=pos, vel=vel, p=p,
                        **{f"ref_{k}": v for k, v in ref.items()})
    print(f"wrote {fn}: N={len(p)} nnz(A)={len(ref['A_v'])}")


def main():
    if not os.path.isdir(REF_ROOT):
        sys.exit("reference tree not mounted; fixtures are generated only where it is")
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)

    # 1. the reference's own 2-D scene and parameters
    v4 = os.path.join(REF_ROOT, "MPS_CAL_v4.0")
    prm = read_input(os.path.join(v4, "input.dat"))
    nb2, nb1, pos, vel, p = read_geo(os.path.join(v4, "GEO.dat"))
    # give the scene a smooth velocity and pressure field so that the
    # Laplacian and gradient operators are exercised
    rng = np.random.default_rng(11)
    vel = np.stack([0.3 * np.sin(pos[:, 0] * 40.0) + 0.01 * rng.standard_normal(len(p)),
                    np.zeros(len(p)),
                    0.2 * np.cos(pos[:, 2] * 30.0) + 0.01 * rng.standard_normal(len(p))], 1)
    vel[: nb2 + nb1] = 0.0
    p = 1000.0 * 9.81 * (pos[:, 2].max() - pos[:, 2]) + 5.0 * rng.standard_normal(len(p))
    save("v4_2d_geo", prm, nb2, nb1, pos, vel, p, run_ref(prm, nb2, nb1, pos, vel, p))

    # 2. small 3-D jittered dam break, config-0 constants
    prm3 = dict(dim=3, rho=1000.0, nu=1e-6, dp=0.01, re_n_mult=2.1, re_lap_mult=4.0, beta=0.97,
                cs=20.0, cfl=0.1, col_mult=0.8, col_rate=2.0, dt_max=1e-4, dt_min_mult=1e-3,
                tt=0.5, tout=0.001, max_dis=0.001)
    nb2, nb1, pos = lattice_dam((6, 6, 8), (12, 6, 10), 2, 2, 0.01, 3, 0.01, 3)
    n = len(pos)
    rng = np.random.default_rng(5)
    vel = np.stack([0.5 * np.sin(pos[:, 2] * 20.0), 0.3 * np.cos(pos[:, 0] * 25.0),
                    -0.4 * np.sin(pos[:, 1] * 30.0)], 1) + 0.01 * rng.standard_normal((n, 3))
    vel[: nb2 + nb1] = 0.0
    p = 1000.0 * 9.81 * np.maximum(0.08 - pos[:, 2], 0.0) + 20.0 * rng.standard_normal(n)
    save("v4_3d_dam", prm3, nb2, nb1, pos, vel, p, run_ref(prm3, nb2, nb1, pos, vel, p))

    # 3. 2-D jittered dam break with re_lap != re_n
    prm2 = dict(prm)
    prm2.update(re_lap_mult=3.1, dt_max=5e-4)
    nb2, nb1, pos = lattice_dam((20, 1, 30), (50, 1, 40), 1, 2, 0.004, 9, 0.01, 2)
    n = len(pos)
    rng = np.random.default_rng(8)
    vel = np.stack([0.2 * np.sin(pos[:, 2] * 50.0), np.zeros(n), 0.1 * np.cos(pos[:, 0] * 40.0)], 1)
    vel[: nb2 + nb1] = 0.0
    p = 1000.0 * 9.81 * np.maximum(0.12 - pos[:, 2], 0.0) + 3.0 * rng.standard_normal(n)
    save("v4_2d_jit", prm2, nb2, nb1, pos, vel, p, run_ref(prm2, nb2, nb1, pos, vel, p))


if __name__ == "__main__":
    main()
