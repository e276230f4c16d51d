"""Full-size checks on the benchmark scene (dam break, 1,049,600 fluid / 1,377,536
particles, l0 = 2.5 mm) through properties that do not depend on size:

* the re-binning: the permutation is a permutation, the cell keys recomputed
  on the host in FP32 exactly as the hash forms them are non-decreasing in the
  new order, cell_start[c] is the first slot with key >= c, and the sort is
  stable (inside each cell the old indices increase);
* the particle number density of 4096 sampled non-dummy particles equals an FP64
  restatement over their 27-cell stencils within 1e-12 (cutoff decided in FP32
  as sqrt(rr) < R0, terms from exact FP64 differences of the FP32 positions);
* after the steps, every value is finite and no particle has left the grid.

Also the re-binning schedule of MPS_CAL_expl_v1.0 (re-bin after every position
update) against the oracle over a short trajectory, at the parity-test size.
"""
import numpy as np
import pytest

import mps_solvers_amd as mps
import oracle
from scenes import gpu_params, normwise, oracle_params, small_dam

pytestmark = pytest.mark.gpu

DP = 0.0025


def _full_scene():
    nb2, nb1, pos = mps.dambreak((80, 80, 164), (120, 80, 170), 2, 2, DP, 0.01, 0)
    P = gpu_params(0, dp=DP, dt_max=2.5e-5, max_dis=0.1 * DP)
    return mps.Solver(P, nb2, nb1, pos), oracle_params(0, dp=DP, dt_max=2.5e-5, max_dis=0.1 * DP)


def _keys32(pos, g):
    """The cell hash in FP32: trunc((x - left) / h) per axis, z-major key."""
    h = np.float32(g["h"])
    c = [np.trunc((pos[:, a].astype(np.float32) - np.float32(o)) / h).astype(np.int64)
         for a, o in enumerate((g["left"], g["back"], g["bottom"]))]
    return (c[2] * g["dy"] + c[1]) * g["dx"] + c[0]


def test_full_size_sort_and_density():
    s, P = _full_scene()
    N = s.nt
    assert N == 1377536
    s.step(5)
    before = s.state()
    perm, cs = s.stage_sort()
    g = s.grid()
    st = s.state()
    # permutation
    assert np.array_equal(np.sort(perm), np.arange(N))
    assert np.array_equal(st["id"], before["id"][perm])
    # keys non-decreasing, cell_start consistent, stable inside cells
    keys = _keys32(st["pos"], g)
    assert s.np_active() == N and keys.min() >= 0 and keys.max() < g["ncells"]
    assert np.all(np.diff(keys) >= 0)
    assert np.array_equal(cs, np.searchsorted(keys, np.arange(g["ncells"] + 1), side="left"))
    same = np.diff(keys) == 0
    assert np.all(np.diff(perm)[same] > 0)

    # n of sampled particles against an FP64 restatement over the stencil
    s.stage_pnd_pres()
    st = s.state()
    re = np.float32(oracle.thresholds(P)[0])
    x = st["pos"].astype(np.float32)
    rng = np.random.default_rng(0)
    sample = rng.choice(np.nonzero(st["type"] != 2)[0], 4096, replace=False)   # n is not formed for dummies
    n_ref = np.zeros(len(sample))
    dx_, dy_ = g["dx"], g["dy"]
    for k, i in enumerate(sample):
        c = keys[i]
        cz, rem = divmod(c, dx_ * dy_)
        cy, cx = divmod(rem, dx_)
        cand = []
        for zz in range(max(cz - 1, 0), min(cz + 2, g["dz"])):
            for yy in range(max(cy - 1, 0), min(cy + 2, dy_)):
                row = (zz * dy_ + yy) * dx_
                a, b = cs[row + max(cx - 1, 0)], cs[row + min(cx + 1, dx_ - 1) + 1]
                cand.append(np.arange(a, b))
        j = np.concatenate(cand)
        j = j[j != i]
        d32 = x[j] - x[i]
        rr = (d32[:, 0] * d32[:, 0] + d32[:, 1] * d32[:, 1]) + d32[:, 2] * d32[:, 2]
        j = j[np.sqrt(rr) < re]
        d64 = x[j].astype(np.float64) - x[i].astype(np.float64)
        n_ref[k] = np.sum(float(re) / np.sqrt(np.sum(d64 * d64, axis=1)) - 1.0)
    assert normwise(st["n"][sample], n_ref) < 1e-12

    s.stage_gradp()
    st = s.state()
    for k in ("pos", "vel", "p", "n"):
        assert np.all(np.isfinite(st[k])), k
    s.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_rebin_after_every_move_trajectory(precision):
    """MPS_CAL_expl_v1.0's schedule (visc, re-bin, collision, re-bin, n, p,
    grad p, re-bin) for 20 steps against the oracle, by particle id."""
    nb2, nb1, pos = small_dam(8)
    s = mps.Solver(gpu_params(precision, rebin_every_move=1), nb2, nb1, pos)
    sim = oracle.Sim(oracle_params(precision), nb2, nb1, pos[:, 0], pos[:, 1], pos[:, 2])
    s.step(20)
    sim.step(20, rebin_every_move=True)
    g = s.state(by_id=True)
    o = sim.state()
    order = np.argsort(o["id"])
    for a, k in enumerate("xyz"):
        err = np.max(np.abs(g["pos"][:, a] - o[k][order]))
        assert err < (2e-6 if precision == 0 else 1e-12), (k, err)
    # velocity vectors against the largest oracle speed: in a dam at rest the
    # horizontal components are ~1e-19 m/s, so a per-component bar compares noise
    vo = np.stack([o[k][order] for k in "uvw"], 1)
    err = np.max(np.linalg.norm(g["vel"] - vo, axis=1)) / np.max(np.linalg.norm(vo, axis=1))
    assert err < (1e-4 if precision == 0 else 1e-6), err
    assert abs(s.scalars()["time"] - sim.consts()["time"]) < 1e-12
