"""Wave-tank scene generator and the oracle's prescribed piston motion (CPU).

Pinned by closed forms: particle counts of the lattice, the piston's layer
structure, and the piston law x(t) = x(0) + A sin(w t) reproduced exactly by
the oracle simulator's moving particles.
"""
import numpy as np

import mps_solvers_amd as mps
import oracle
from scenes import oracle_params


def _counts(nx, ny, nz, depth, travel, lw, ld):
    L = lw + ld
    fluid = nx * ny * depth
    piston = L * ny * nz
    floor = (nx + 2 * L + travel) * (ny + 2 * L) * L
    sides = 2 * L * (nx + 2 * L + travel) * nz
    far = L * ny * nz
    return fluid, piston, floor + sides + far + piston


def test_wavetank_counts_and_layers():
    nx, ny, nz, depth, travel = 20, 6, 12, 6, 3
    nb2, nb1, pos, pist = mps.wavetank(nx, ny, nz, depth, travel, 2, 2, 0.004, 0.0, 0)
    fluid, piston, solid = _counts(nx, ny, nz, depth, travel, 2, 2)
    assert len(pos) - nb2 - nb1 == fluid
    assert nb2 + nb1 == solid
    assert int(pist.sum()) == piston
    q = np.rint(pos / 0.004).astype(int)
    assert len(np.unique(q, axis=0)) == len(q)
    # piston: x in [-4, -1], inner y and z; first 2 layers walls, the others dummies
    qp = q[pist == 1]
    assert qp[:, 0].min() == -4 and qp[:, 0].max() == -1
    assert qp[:, 1].min() == 0 and qp[:, 1].max() == ny - 1 and qp[:, 2].min() == 0
    t = np.zeros(len(pos), int)
    t[:nb2] = 2
    t[nb2:nb2 + nb1] = 1
    assert np.all(t[pist == 1][qp[:, 0] >= -2] == 1) and np.all(t[pist == 1][qp[:, 0] < -2] == 2)
    # fluid fills [0, nx) x [0, ny) x [0, depth)
    qf = q[t == 0]
    assert qf[:, 0].min() == 0 and qf[:, 0].max() == nx - 1 and qf[:, 2].max() == depth - 1
    # the floor reaches travel sites behind the piston
    assert q[:, 0].min() == -4 - travel


def test_wavetank_jitter_is_seeded():
    a = mps.wavetank(10, 4, 8, 4, 2, 2, 2, 0.004, 0.01, 5)[2]
    b = mps.wavetank(10, 4, 8, 4, 2, 2, 2, 0.004, 0.01, 5)[2]
    c = mps.wavetank(10, 4, 8, 4, 2, 2, 2, 0.004, 0.01, 6)[2]
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.max(np.abs(a - np.rint(a / 0.004) * 0.004)) <= 0.01 * 0.004 + 1e-15


def test_oracle_piston_follows_the_law():
    nb2, nb1, pos, pist = mps.wavetank(16, 4, 10, 5, 3, 2, 2, 0.004, 0.0, 0)
    P = oracle_params(1, dp=0.004, dt_max=2e-4, max_dis=0.004 * 4, col_mult=0.8)
    sim = oracle.Sim(P, nb2, nb1, pos[:, 0], pos[:, 1], pos[:, 2])
    amp, freq = 0.01, 0.8
    sim.set_motion(pist, 0, amp, freq)
    sim.step(25)
    o = sim.state()
    t = sim.consts()["time"]
    w = 2 * np.pi * freq
    ids = o["id"]
    m = pist[ids] == 1
    assert np.allclose(o["x"][m], pos[ids[m], 0] + amp * np.sin(w * t), rtol=0, atol=1e-15)
    assert np.allclose(o["u"][m], amp * w * np.cos(w * t), rtol=0, atol=1e-15)
    assert np.all(o["v"][m] == 0) and np.all(o["w"][m] == 0)
    # other walls stay put; the water next to the piston is pushed along +x
    still = (pist[ids] == 0) & (o["type"] != 0)
    assert np.array_equal(o["x"][still], pos[ids[still], 0])
    near = (o["type"] == 0) & (pos[ids, 0] < 3 * 0.004)
    assert o["u"][near].mean() > 0.0
