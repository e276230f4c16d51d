"""GPU parity of the pressure Poisson assembly and CG against the oracle.

Bars: CSR pattern and Dirichlet flags bit-exact; FP64 values and b within
1e-12 normwise; every solve reaches |b - A x| / |b| <= 1e-8 measured by the
oracle's FP64 SpMV on the oracle's own matrix; solutions agree with the
oracle's FP64 CG (1e-13 relative residual) within 1e-5 normwise.
"""
import numpy as np
import pytest

import mps_solvers_amd as mps
import oracle
from scenes import gpu_params, normwise, oracle_grid, oracle_params, oracle_view, small_dam

pytestmark = pytest.mark.gpu

DT = 5e-4


def _setup(precision, alpha, seed=3, re_lap=3.1):
    nb2, nb1, pos = small_dam(seed)
    kw = dict(re_lap_mult=re_lap)
    s = mps.Solver(gpu_params(precision, **kw), nb2, nb1, pos)
    st = s.state()
    rng = np.random.default_rng(seed)
    p = st["pos"].copy()
    fl = st["type"] == 0
    p[fl] += 0.08 * s.P.dp * rng.uniform(-1, 1, (fl.sum(), 3))
    pr = 1000 * 9.81 * np.maximum(0.08 - p[:, 2], 0)
    s.set_state(pos=p, press=pr)
    _, cs = s.stage_sort()
    s.set_dt(DT)
    s.stage_pnd()
    nnz = s.ppe_assemble(alpha)
    st = s.state()
    sc = s.scalars()
    ref = oracle.build_poisson(oracle_params(precision, **kw), oracle_grid(s), oracle_view(st), cs,
                               st["n"], st["p"], sc["n0"], sc["lambda"], DT, alpha)
    return s, nnz, ref


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("alpha", [0.0, 1.0 / (1000.0 * 20.0 ** 2)])
def test_assembly_matches_buildPoisson(precision, alpha):
    s, nnz, (rp, col, val, b, x0, bc) = _setup(precision, alpha)
    A = s.ppe_get()
    assert nnz == len(col)
    np.testing.assert_array_equal(A["row_ptr"], rp)
    np.testing.assert_array_equal(A["col"], col)
    np.testing.assert_array_equal(A["bc"], bc)
    assert normwise(A["val"], val) < 1e-12
    assert normwise(A["b"], b) < 1e-12
    assert bc.sum() > 0 and (bc == 0).sum() > 0
    # interior rows carry ~ the lattice neighbour count at 3.1 l0
    nrow = np.diff(rp)
    assert nrow.max() >= 100


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("mixed", [True, False])
@pytest.mark.parametrize("x0_kind", [0, 1])
def test_cg_reaches_1e8(precision, mixed, x0_kind):
    s, nnz, (rp, col, val, b, x0, bc) = _setup(precision, 0.0)
    it, rel = s.ppe_solve(mixed=mixed, x0_kind=x0_kind, tol=1e-8)
    assert rel <= 1e-8 and it > 0
    x = s.ppe_get()["x"]
    r = b - oracle.spmv(rp, col, val, x)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.05e-8
    xo, ito, rro = oracle.cg(rp, col, val, b, np.zeros_like(b), 1e-13, relative=True)
    assert normwise(x, xo) < 1e-5
    s.ppe_press_corr()
    p = s.state()["p"][: len(x)]
    np.testing.assert_allclose(p, np.maximum(x, 0.0), rtol=1e-6 if precision == 0 else 1e-15,
                               atol=1e-6 * np.max(np.abs(x)))


def test_cg_matches_dense_direct_solve():
    """Small system: the oracle CG and numpy's dense solve agree (pins the oracle's CG)."""
    s, nnz, (rp, col, val, b, x0, bc) = _setup(1, 0.0)
    n = len(b)
    A = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(rp))
    A[rows, col] = val
    assert np.allclose(A, A.T, rtol=0, atol=1e-12 * np.abs(A).max())
    xd = np.linalg.solve(A, b)
    xo, _, _ = oracle.cg(rp, col, val, b, np.zeros_like(b), 1e-13, relative=True)
    assert normwise(xo, xd) < 1e-9
    it, rel = s.ppe_solve(mixed=True, x0_kind=0, tol=1e-8)
    assert normwise(s.ppe_get()["x"], xd) < 1e-5


@pytest.mark.parametrize("precision", [0, 1])
def test_semi_implicit_steps_run_and_converge(precision):
    # unjittered lattice at rest: n0 is the interior value, so gravity compresses
    # the bottom layers at once and the pressure field builds from below
    nb2, nb1, pos = small_dam(5, fluid=(6, 6, 12), inner=(12, 6, 14), jitter=0.0)
    s = mps.Solver(gpu_params(precision, re_lap_mult=3.1, dt_max=1e-3), nb2, nb1, pos)
    it, ms = s.step_semi(10, alpha=0.0, x0_kind=1, tol=1e-8)
    assert it > 0 and ms > 0
    st = s.state()
    fl = st["type"] == 0
    assert np.all(np.isfinite(st["pos"])) and np.max(st["p"][fl]) > 0
    # hydrostatic check: mean pressure grows with depth in the fluid column
    z = st["pos"][fl, 2]
    p = st["p"][fl]
    lo, hi = z < np.percentile(z, 25), z > np.percentile(z, 75)
    assert p[lo].mean() > p[hi].mean()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("rebin", [0, 1])
def test_semi_implicit_trajectory_matches_oracle(precision, rebin):
    """5 semi-implicit steps (alpha = 1/(rho c^2)) on the device against the
    oracle's semi-implicit simulator (FP64 CG to the same relative residual),
    by particle id, with one re-binning per step or one after every position
    update.  The two solvers stop at different points inside the 1e-8
    residual ball, so the bars are looser than the explicit step's: FP64 path
    p and u within 1e-5 normwise, FP32 path within 2e-3."""
    nb2, nb1, pos = small_dam(5, jitter=0.0)
    kw = dict(re_lap_mult=3.1, dt_max=DT)
    alpha = 1.0 / (1000.0 * 20.0 ** 2)
    s = mps.Solver(gpu_params(precision, rebin_every_move=rebin, **kw), nb2, nb1, pos)
    sim = oracle.Sim(oracle_params(precision, **kw), nb2, nb1, pos[:, 0], pos[:, 1], pos[:, 2])
    s.step_semi(5, alpha=alpha, x0_kind=1, tol=1e-8)
    sim.step_semi(5, alpha=alpha, tol=1e-8, rebin_every_move=bool(rebin))
    g = s.state(by_id=True)
    o = sim.state()
    order = np.argsort(o["id"])
    assert abs(s.scalars()["time"] - sim.consts()["time"]) < 1e-12
    fl = g["type"] == 0
    bar = 2e-3 if precision == 0 else 1e-5
    assert np.max(o["p"]) > 0.0
    assert normwise(g["p"][fl], o["p"][order][fl]) < bar
    for a, k in enumerate("uvw"):
        assert normwise(g["vel"][fl, a], o[k][order][fl]) < bar, k
    for a, k in enumerate("xyz"):
        err = np.max(np.abs(g["pos"][:, a] - o[k][order]))
        assert err < (4e-6 if precision == 0 else 1e-9), (k, err)
