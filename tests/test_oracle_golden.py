"""Pin the CPU oracle against golden vectors from the reference's own
MPS_CAL_v4.0 CPU solver (tests/golden/make_golden.py).

Covers MPS::allocateCellList, MPS_CPU::divideCell (hash, stable sort,
cell_start/end), MPS::calOnce (n0, lambda), MPS_CPU::cal_n, d3_LapVel,
d3_GradPress, update_dt, buildPoisson and DoCG.
"""
import numpy as np
import pytest

import oracle
from golden_io import CASES, load


def _params(prm, **kw):
    keys = ["dim", "rho", "nu", "dp", "re_n_mult", "re_lap_mult", "beta", "cs", "cfl",
            "col_mult", "col_rate", "dt_max", "dt_min_mult", "max_dis"]
    d = {k: prm[k] for k in keys}
    d.update(kw)
    return oracle.params(**d)


def _setup(case):
    fx = load(case)
    prm = fx["params"]
    P = _params(prm, f32=0, gradp_skip_dummy=1)
    pos = fx["pos"]
    n = len(pos)
    nb2, nb1 = int(fx["nb2"]), int(fx["nb1"])
    type0 = np.where(np.arange(n) < nb2, 2, np.where(np.arange(n) < nb2 + nb1, 1, 0)).astype(np.int32)
    G = oracle.grid_init(P, pos[:, 0], pos[:, 1], pos[:, 2])
    key = oracle.hash_cells(P, G, pos[:, 0], pos[:, 1], pos[:, 2], type0)
    perm, index, cs = oracle.sort_cells(key, G.ncells)
    st = {k: fx["pos"][perm, a] for a, k in enumerate("xyz")}
    st.update({k: fx["vel"][perm, a] for a, k in enumerate("uvw")})
    st["type"] = type0[perm]
    st["p"] = fx["p"][perm]
    return fx, P, G, perm, index, cs, st


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("case", CASES)
def test_grid_matches_allocateCellList(case):
    fx, P, G, *_ = _setup(case)
    c = fx["ref_consts"]
    assert (G.left, G.back, G.bottom, G.h) == (c[1], c[2], c[3], c[4])
    assert (G.dx, G.dy, G.dz, G.ncells) == (int(c[5]), int(c[6]), int(c[7]), int(c[8]))


@pytest.mark.parametrize("case", CASES)
def test_sort_permutation_bit_exact(case):
    fx, P, G, perm, index, cs, st = _setup(case)
    np.testing.assert_array_equal(perm, fx["ref_id"])
    np.testing.assert_array_equal(index, fx["ref_index_last"])
    np.testing.assert_array_equal(cs[:-1], fx["ref_cell_start"])
    np.testing.assert_array_equal(cs[1:], fx["ref_cell_end"])
    np.testing.assert_array_equal(st["type"], fx["ref_type"])
    for k in "xyz":
        np.testing.assert_array_equal(st[k], fx[f"ref_{k}"])


@pytest.mark.parametrize("case", CASES)
def test_calonce_n0_lambda(case):
    fx, P, G, perm, index, cs, st = _setup(case)
    nn, n0, lam, am, normal = oracle.calonce(P, G, st["x"], st["y"], st["z"], st["type"], cs)
    c = fx["ref_consts"]
    assert abs(n0 - c[9]) <= 1e-13 * c[9]
    assert abs(lam - c[10]) <= 1e-12 * c[10]
    # brute-force n of calOnce (all particles) vs the stencil sum
    assert _rel(nn, fx["ref_n_calonce"]) < 1e-13


@pytest.mark.parametrize("case", CASES)
def test_pnd_matches_cal_n(case):
    fx, P, G, perm, index, cs, st = _setup(case)
    n = oracle.pnd(P, G, st["x"], st["y"], st["z"], st["type"], cs, n_init=fx["ref_n_calonce"])
    np.testing.assert_allclose(n, fx["ref_n"], rtol=1e-14, atol=0)


@pytest.mark.parametrize("case", CASES)
def test_laplacian_and_gradient(case):
    fx, P, G, perm, index, cs, st = _setup(case)
    c = fx["ref_consts"]
    lap = oracle.lapvel(P, G, st, cs, c[12])
    for a, k in enumerate("uvw"):
        assert _rel(lap[a], fx[f"ref_lap_{k}"]) < 1e-13, k
    g = oracle.gradp_vec(P, G, st, st["p"], cs, c[9])
    for a, k in enumerate("xyz"):
        if np.max(np.abs(fx[f"ref_grad_{k}"])) == 0:
            assert np.max(np.abs(g[a])) == 0
        else:
            assert _rel(g[a], fx[f"ref_grad_{k}"]) < 1e-13, k


@pytest.mark.parametrize("case", CASES)
def test_update_dt(case):
    fx, P, G, perm, index, cs, st = _setup(case)
    dt, small = oracle.update_dt(P, st["u"], st["v"], st["w"])
    assert dt == pytest.approx(fx["ref_dt_update"][0], rel=1e-15)


@pytest.mark.parametrize("case", CASES)
def test_build_poisson_and_cg(case):
    fx, P, G, perm, index, cs, st = _setup(case)
    c = fx["ref_consts"]
    prm = fx["params"]
    alpha = 1.0 / (prm["rho"] * prm["cs"] ** 2)
    rp, col, val, b, x0, bc = oracle.build_poisson(P, G, st, cs, fx["ref_n"], st["p"], c[9],
                                                   c[10], c[11], alpha)
    N = len(b)
    rows = np.repeat(np.arange(N), np.diff(rp))
    ours = {(int(i), int(j)): v for i, j, v in zip(rows, col, val) if v != 0.0}
    ref = {(int(i), int(j)): v for i, j, v in zip(fx["ref_A_i"], fx["ref_A_j"], fx["ref_A_v"])}
    assert ours.keys() == ref.keys()
    diff = max(abs(ours[k] - ref[k]) / abs(ref[k]) for k in ref)
    assert diff < 1e-13
    np.testing.assert_allclose(b, fx["ref_b"], rtol=1e-14, atol=0)
    np.testing.assert_array_equal(x0, fx["ref_x0"])
    np.testing.assert_array_equal(bc.astype(np.int32), fx["ref_bc"])
    # DoCG stops at r.r <= 1e-20 absolute (MPS_CAL_v4.0/common.h:45-48)
    x, it, rr = oracle.cg(rp, col, val, b, x0, 1e-20, relative=False)
    assert it > 0
    assert _rel(x, fx["ref_x_cg"]) < 1e-10
    np.testing.assert_allclose(np.maximum(x, 0.0), fx["ref_p_corr"],
                               atol=1e-10 * np.max(np.abs(fx["ref_p_corr"])))
