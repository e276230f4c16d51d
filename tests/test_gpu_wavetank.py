"""GPU checks of the wave tank with a moving piston, and of the VTK writer.

Bars: the piston's wall and dummy particles sit exactly on the prescribed law
(FP32: the law rounded to float; FP64: 1e-12 m); a 20-step trajectory matches
the oracle simulator with the same motion within the explicit-step bars
(positions 2e-6 m FP32 / 1e-12 m FP64, velocities normwise 1e-4 / 1e-6); the
VTK file carries the state in id order at the writer's %lf precision.
"""
import numpy as np
import pytest

import mps_solvers_amd as mps
import oracle
from scenes import gpu_params, normwise, oracle_params

pytestmark = pytest.mark.gpu

DP = 0.004
KW = dict(dp=DP, dt_max=2e-4, max_dis=DP * 4)


def _tank(jitter=0.01):
    return mps.wavetank(16, 4, 10, 5, 3, 2, 2, DP, jitter, 4)


@pytest.mark.parametrize("precision", [0, 1])
def test_piston_trajectory_matches_oracle(precision):
    nb2, nb1, pos, pist = _tank()
    s = mps.Solver(gpu_params(precision, **KW), nb2, nb1, pos)
    sim = oracle.Sim(oracle_params(precision, **KW), nb2, nb1, pos[:, 0], pos[:, 1], pos[:, 2])
    amp, freq = 0.05, 0.8
    s.set_motion(pist, 0, amp, freq)
    sim.set_motion(pist, 0, amp, freq)
    s.step(20)
    sim.step(20)
    g = s.state(by_id=True)
    o = sim.state()
    order = np.argsort(o["id"])
    t = s.scalars()["time"]
    assert abs(t - sim.consts()["time"]) < 1e-12
    w = 2 * np.pi * freq
    m = pist == 1
    law = pos[m, 0].astype(np.float32 if precision == 0 else np.float64).astype(np.float64)
    law = law + amp * np.sin(w * t)
    if precision == 0:
        law = law.astype(np.float32).astype(np.float64)
        assert np.max(np.abs(g["pos"][m, 0] - law)) <= np.spacing(np.float32(0.1))
    else:
        assert np.max(np.abs(g["pos"][m, 0] - law)) < 1e-12
    assert np.allclose(g["vel"][m, 0], amp * w * np.cos(w * t), rtol=1e-6, atol=0)
    for a, k in enumerate("xyz"):
        err = np.max(np.abs(g["pos"][:, a] - o[k][order]))
        assert err < (2e-6 if precision == 0 else 1e-12), (k, err)
    for a, k in enumerate("uvw"):
        assert normwise(g["vel"][:, a], o[k][order]) < (1e-4 if precision == 0 else 1e-6), k
    # the piston moved into the water and pushed it along +x
    fl = g["type"] == 0
    assert g["vel"][fl, 0].mean() > 0.0


def test_semi_implicit_wave_tank_with_learned_guess():
    nb2, nb1, pos, pist = _tank(jitter=0.0)
    s = mps.Solver(gpu_params(0, re_lap_mult=3.1, **KW), nb2, nb1, pos)
    s.set_motion(pist, 0, 0.05, 0.8)
    alpha = 1.0 / (1000.0 * 20.0 ** 2)
    s.mlp_init(2)
    s.mlp_set_record(True)
    it0, _ = s.step_semi(6, alpha=alpha, x0_kind=1, tol=1e-8)
    s.mlp_set_record(False)
    s.mlp_train(100, batch=512, lr=1e-3)
    it1, _ = s.step_semi(3, alpha=alpha, x0_kind=3, tol=1e-8)
    assert it0 > 0 and it1 > 0
    st = s.state(by_id=True)
    assert np.all(np.isfinite(st["pos"])) and np.all(np.isfinite(st["p"]))


def _parse_vtk(path):
    lines = open(path).read().split("\n")
    assert lines[0] == "# vtk DataFile Version 3.0" and lines[1] == "unstructured grid"
    assert lines[2] == "ASCII" and lines[3] == "DATASET UNSTRUCTURED_GRID"
    n = int(lines[4].split()[1])
    assert lines[4] == f"POINTS {n} float"
    pts = np.array(lines[5].split(), float).reshape(n, 3)
    assert lines[6] == f"CELLS {n} {2 * n}"
    cells = np.array(lines[7].split(), int).reshape(n, 2)
    assert np.all(cells[:, 0] == 1) and np.array_equal(cells[:, 1], np.arange(n))
    assert lines[8] == f"CELL_TYPES {n}" and lines[9].split() == ["1"] * n
    assert lines[10] == f"POINT_DATA {n}"
    assert lines[11] == "SCALARS id int 1" and lines[12] == "LOOKUP_TABLE default"
    ids = np.array(lines[13].split(), int)
    assert lines[14] == "SCALARS pressure float 1" and lines[15] == "LOOKUP_TABLE default"
    pr = np.array(lines[16].split(), float)
    assert lines[17] == "VECTORS velocity float"
    vel = np.array(lines[18].split(), float).reshape(n, 3)
    return pts, ids, pr, vel


def test_vtk_writer(tmp_path):
    nb2, nb1, pos, pist = _tank()
    s = mps.Solver(gpu_params(0, **KW), nb2, nb1, pos)
    s.set_motion(pist, 0, 0.05, 0.8)
    s.step(5)
    path = tmp_path / "part0000.vtk"
    s.write_vtk(path)
    pts, ids, pr, vel = _parse_vtk(path)
    st = s.state(by_id=True)
    assert np.array_equal(ids, np.arange(len(pos)))
    assert np.max(np.abs(pts - st["pos"])) <= 5e-7
    assert np.max(np.abs(vel - st["vel"])) <= 5e-7
    assert np.max(np.abs(pr - st["p"])) <= 5e-7


@pytest.mark.parametrize("precision", [0, 1])
def test_semi_implicit_piston_matches_oracle(precision):
    """The config-4 step (semi-implicit, piston moving) for 5 steps against the
    oracle's semi-implicit simulator with the same motion; bars as the
    semi-implicit trajectory test (p, u normwise 1e-5 FP64 / 2e-3 FP32)."""
    nb2, nb1, pos, pist = _tank(jitter=0.0)
    kw = dict(KW, re_lap_mult=3.1)
    alpha = 1.0 / (1000.0 * 20.0 ** 2)
    s = mps.Solver(gpu_params(precision, **kw), nb2, nb1, pos)
    sim = oracle.Sim(oracle_params(precision, **kw), nb2, nb1, pos[:, 0], pos[:, 1], pos[:, 2])
    s.set_motion(pist, 0, 0.05, 0.8)
    sim.set_motion(pist, 0, 0.05, 0.8)
    s.step_semi(5, alpha=alpha, x0_kind=1, tol=1e-8)
    sim.step_semi(5, alpha=alpha, tol=1e-8)
    g = s.state(by_id=True)
    o = sim.state()
    order = np.argsort(o["id"])
    fl = g["type"] == 0
    bar = 2e-3 if precision == 0 else 1e-5
    assert np.max(o["p"]) > 0.0
    assert normwise(g["p"][fl], o["p"][order][fl]) < bar
    for a, k in enumerate("uvw"):
        assert normwise(g["vel"][fl, a], o[k][order][fl]) < bar, k
