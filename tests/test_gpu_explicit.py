"""GPU parity of the explicit MPS step against the CPU oracle.

Bars (per quantity, normwise: max |gpu - oracle| / max |oracle|):
  * permutations, cell_start, neighbour lists, normals, collision hit counts: bit-exact;
  * FP32 path (precision 0): 1e-4 on every operator output (positions: 4 ulp), n within 1e-12;
  * FP64 path (precision 1): 1e-6 (observed ~1e-13).
Every operator is driven from an identical sorted state read back from the
device, so the oracle and the kernels see the same inputs.
"""
import numpy as np
import pytest

import mps_solvers_amd as mps
import oracle
from scenes import (gpu_params, normwise, oracle_grid, oracle_params, oracle_view, small_dam)

pytestmark = pytest.mark.gpu

TOL = {0: 1e-4, 1: 1e-6}


def _pos_ok(out_pos, o, ref, precision):
    """Positions: FP64 path normwise on the displacement; FP32 path within
    4 ulp of the largest coordinate (a displacement dt * du can sit below the
    FP32 resolution of the coordinate itself)."""
    for a, k in enumerate("xyz"):
        if precision == 1:
            assert normwise(out_pos[:, a] - o[k], ref[k] - o[k]) < TOL[1], k
        else:
            ulp = np.spacing(np.float32(np.max(np.abs(ref[k]))))
            assert np.max(np.abs(out_pos[:, a] - ref[k])) <= 4 * ulp, k


def _solver(precision, seed=3, **kw):
    nb2, nb1, pos = small_dam(seed)
    P = gpu_params(precision, **kw)
    rng = np.random.default_rng(seed)
    vel = np.zeros_like(pos)
    nfl = len(pos) - nb2 - nb1
    vel[nb2 + nb1:] = 0.3 * rng.standard_normal((nfl, 3))
    return mps.Solver(P, nb2, nb1, pos, vel), oracle_params(precision, **kw)


def _perturb(s, precision, amp, seed=1):
    """Move fluid particles by up to amp*dp and give them a fresh pressure field."""
    st = s.state()
    rng = np.random.default_rng(seed)
    fl = st["type"] == 0
    pos = st["pos"].copy()
    pos[fl] += amp * s.P.dp * rng.uniform(-1, 1, (fl.sum(), 3))
    pr = 1000 * 9.81 * np.maximum(0.1 - pos[:, 2], 0) + 50 * rng.standard_normal(len(pos))
    s.set_state(pos=pos, press=pr)
    return s.state()


@pytest.mark.parametrize("precision", [0, 1])
def test_sort_permutation_bit_exact(precision):
    s, P = _solver(precision)
    before = _perturb(s, precision, 0.9)
    G = oracle_grid(s)
    key = oracle.hash_cells(P, G, before["pos"][:, 0], before["pos"][:, 1], before["pos"][:, 2],
                            before["type"])
    perm_o, index_o, cs_o = oracle.sort_cells(key, G.ncells)
    perm, cs = s.stage_sort()
    np.testing.assert_array_equal(perm, perm_o)
    np.testing.assert_array_equal(cs, cs_o)
    after = s.state()
    np.testing.assert_array_equal(after["id"], before["id"][perm_o])
    np.testing.assert_array_equal(after["pos"], before["pos"][perm_o])
    assert not np.array_equal(perm, np.arange(len(perm)))  # the perturbation moved particles


@pytest.mark.parametrize("precision", [0, 1])
def test_rebinning_from_a_sorted_order(precision):
    """Once sorted, a state is re-binned by gathering each new cell from the
    27 old cells around it; a particle that jumps further, or leaves the grid,
    makes its cell walk every slot.  Each case against the oracle's stable sort
    of the same keys: ordinary moves (up to ~1 l0), then one particle carried
    three cells in x and another out of the top of the grid."""
    s, P = _solver(precision)
    G = oracle_grid(s)
    h = s.grid()["h"]

    def check():
        before = s.state()
        key = oracle.hash_cells(P, G, before["pos"][:, 0], before["pos"][:, 1], before["pos"][:, 2],
                                before["type"])
        perm_o, _, cs_o = oracle.sort_cells(key, G.ncells)
        perm, cs = s.stage_sort()
        np.testing.assert_array_equal(perm, perm_o)
        np.testing.assert_array_equal(cs, cs_o)
        np.testing.assert_array_equal(s.state()["id"], before["id"][perm_o])
        return perm

    s.set_dt(0.03)    # 0.3 m/s random velocities: moves up to ~1 l0
    s.stage_visc()
    perm = check()
    assert not np.array_equal(perm, np.arange(len(perm)))

    st = s.state()
    fl = np.nonzero(st["type"] == 0)[0]
    vel = np.zeros_like(st["vel"])
    dt = 1e-3
    vel[fl[0], 0] = 3.2 * h / dt          # three cells along x within one update
    vel[fl[len(fl) // 2], 2] = 60 * h / dt  # out of the top of the grid
    s.set_state(vel=vel)
    s.set_dt(dt)
    nexcl = s.scalars()["nexcl"]
    s.stage_visc()
    check()
    assert s.scalars()["nexcl"] == nexcl + 1


@pytest.mark.parametrize("precision", [0, 1])
def test_calonce_constants_and_normals(precision):
    s, P = _solver(precision)
    st = s.state()
    G = oracle_grid(s)
    perm, index, cs = oracle.sort_cells(
        oracle.hash_cells(P, G, *[st["pos"][:, a] for a in range(3)], st["type"]), G.ncells)
    assert np.array_equal(perm, np.arange(len(perm)))  # state is already sorted
    nn, n0, lam, am, normal = oracle.calonce(P, G, *[st["pos"][:, a] for a in range(3)],
                                             st["type"], cs)
    sc = s.scalars()
    assert abs(sc["n0"] - n0) <= 1e-12 * n0
    assert abs(sc["lambda"] - lam) <= 1e-9 * lam
    assert normwise(st["n"], nn) < 1e-12
    # normals: p of every dummy equals p of its partner after one pressure stage
    s.stage_pnd_pres()
    st2 = s.state()
    d = st2["type"] == 2
    has = normal[d] >= 0
    np.testing.assert_array_equal(st2["p"][d][has], st2["p"][normal[d][has]])


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("which", [0, 1])
def test_neighbor_lists_bit_exact(precision, which):
    s, P = _solver(precision)
    _perturb(s, precision, 0.3)
    s.stage_sort()
    st = s.state()
    G = oracle_grid(s)
    cs = s.stage_sort()[1]   # already sorted: identity, returns cell_start
    st = s.state()
    rp_o, col_o = oracle.neighbors(P, G, *[st["pos"][:, a] for a in range(3)], cs, which)
    rp, col = s.neighbor_list(which)
    np.testing.assert_array_equal(rp, rp_o)
    np.testing.assert_array_equal(col, col_o)


@pytest.mark.parametrize("precision", [0, 1])
def test_pnd_and_pressure(precision):
    s, P = _solver(precision)
    _perturb(s, precision, 0.15)
    _, cs = s.stage_sort()
    st = s.state()
    G = oracle_grid(s)
    o = oracle_view(st)
    n_o = oracle.pnd(P, G, o["x"], o["y"], o["z"], o["type"], cs, n_init=o["n"])
    s.stage_pnd_pres()
    out = s.state()
    assert normwise(out["n"], n_o) < 1e-12
    n0 = s.scalars()["n0"]
    nrm = np.array([-1] * len(n_o), np.int32)
    p_o = oracle.pres_expl(P, n_o, n0, o["type"], nrm, p_init=o["p"])
    nd = out["type"] != 2
    assert np.max(p_o[nd]) > 0
    assert normwise(out["p"][nd], p_o[nd]) < TOL[precision]


@pytest.mark.parametrize("precision", [0, 1])
def test_viscosity_operator(precision):
    # nu large enough that the Laplacian term dominates the velocity change
    s, P = _solver(precision, nu=0.05)
    _perturb(s, precision, 0.2)
    _, cs = s.stage_sort()
    dt = 1e-4
    s.set_dt(dt)
    st = s.state()
    o = oracle_view(st)
    coef = s.scalars()["coef"]
    ref = oracle.visc(P, oracle_grid(s), o, cs, coef, dt)
    s.stage_visc()
    out = s.state()
    for a, k in enumerate("uvw"):
        du, du_o = out["vel"][:, a] - o[k], ref[k] - o[k]
        assert normwise(du, du_o) < TOL[precision], k
    _pos_ok(out["pos"], o, ref, precision)
    # viscous part vs gravity-only: the sum really contributed
    assert np.max(np.abs(ref["u"] - o["u"])) > 1e-4


@pytest.mark.parametrize("precision", [0, 1])
def test_collision_operator(precision):
    s, P = _solver(precision, col_mult=1.0)
    _perturb(s, precision, 0.1)
    _, cs = s.stage_sort()
    dt = 1e-4
    s.set_dt(dt)
    st = s.state()
    o = oracle_view(st)
    ref, ncol = oracle.collision(P, oracle_grid(s), o, cs, dt)
    s.stage_collision()
    out = s.state()
    assert ncol > 50
    assert int(s.scalars()["ncol"]) == ncol
    for a, k in enumerate("uvw"):
        assert normwise(out["vel"][:, a] - o[k], ref[k] - o[k]) < TOL[precision], k
    _pos_ok(out["pos"], o, ref, precision)


@pytest.mark.parametrize("precision", [0, 1])
def test_pressure_gradient_operator(precision):
    s, P = _solver(precision)
    _perturb(s, precision, 0.1)
    _, cs = s.stage_sort()
    dt = 1e-4
    s.set_dt(dt)
    st = s.state()
    o = oracle_view(st)
    n0 = s.scalars()["n0"]
    ref = oracle.gradp(P, oracle_grid(s), o, o["p"], cs, n0, dt)
    s.stage_gradp()
    out = s.state()
    for a, k in enumerate("uvw"):
        assert normwise(out["vel"][:, a] - o[k], ref[k] - o[k]) < TOL[precision], k
    _pos_ok(out["pos"], o, ref, precision)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("lap", [2.5, 5.0])
def test_sweeps_at_other_cell_sizes(precision, lap):
    """The cell edge is max(re_n, re_lap).  At re_lap = 5 l0 a cell holds up to
    ~125 particles, so a cell is swept in two chunks (64 and the rest); at
    2.5 l0 it holds at most 27, so the viscosity sweep keeps each chunk in one
    lane group.  Viscosity, collision, PND + p and the pressure gradient, each
    against the oracle from an identical state."""
    kw = dict(nu=0.05, re_lap_mult=lap, col_mult=1.0)

    def state():
        s, P = _solver(precision, **kw)
        _perturb(s, precision, 0.15)
        _, cs = s.stage_sort()
        s.set_dt(1e-4)
        return s, P, cs, s.state()

    s, P, cs, st = state()
    occ = np.diff(cs)
    assert (occ.max() > 64) if lap == 5.0 else (occ.max() <= 32), occ.max()
    o = oracle_view(st)
    ref = oracle.visc(P, oracle_grid(s), o, cs, s.scalars()["coef"], 1e-4)
    s.stage_visc()
    out = s.state()
    for a, k in enumerate("uvw"):
        assert normwise(out["vel"][:, a] - o[k], ref[k] - o[k]) < TOL[precision], ("visc", k)
    _pos_ok(out["pos"], o, ref, precision)

    s, P, cs, st = state()
    o = oracle_view(st)
    ref, ncol = oracle.collision(P, oracle_grid(s), o, cs, 1e-4)
    s.stage_collision()
    assert ncol > 0 and int(s.scalars()["ncol"]) == ncol
    out = s.state()
    for a, k in enumerate("uvw"):
        assert normwise(out["vel"][:, a] - o[k], ref[k] - o[k]) < TOL[precision], ("collision", k)

    s, P, cs, st = state()
    o = oracle_view(st)
    n_o = oracle.pnd(P, oracle_grid(s), o["x"], o["y"], o["z"], o["type"], cs, n_init=o["n"])
    s.stage_pnd_pres()
    assert normwise(s.state()["n"], n_o) < 1e-12

    s, P, cs, st = state()
    o = oracle_view(st)
    ref = oracle.gradp(P, oracle_grid(s), o, o["p"], cs, s.scalars()["n0"], 1e-4)
    s.stage_gradp()
    out = s.state()
    for a, k in enumerate("uvw"):
        assert normwise(out["vel"][:, a] - o[k], ref[k] - o[k]) < TOL[precision], ("gradp", k)
    _pos_ok(out["pos"], o, ref, precision)


@pytest.mark.parametrize("precision", [0, 1])
def test_short_trajectory(precision):
    """20 full steps on the device against the oracle simulator, by particle id."""
    nb2, nb1, pos = small_dam(7)
    P = gpu_params(precision)
    s = mps.Solver(P, nb2, nb1, pos)
    sim = oracle.Sim(oracle_params(precision), nb2, nb1, pos[:, 0], pos[:, 1], pos[:, 2])
    s.step(20)
    sim.step(20)
    g = s.state(by_id=True)
    o = sim.state()
    order = np.argsort(o["id"])
    for a, k in enumerate("xyz"):
        err = np.max(np.abs(g["pos"][:, a] - o[k][order]))
        assert err < (2e-6 if precision == 0 else 1e-12), (k, err)
    for a, k in enumerate("uvw"):
        assert normwise(g["vel"][:, a], o[k][order]) < TOL[precision], k
    fl = g["type"] == 0
    assert normwise(g["p"][fl], o["p"][order][fl]) < (1e-3 if precision == 0 else 1e-6)
    assert abs(s.scalars()["time"] - sim.consts()["time"]) < 1e-12
