"""GPU checks of the learned CG warm start (BF16-MFMA MLP).

Bars: forward within 3e-3 normwise of the float64 oracle that rounds the same
operands to BF16; gradients within 2e-2 normwise per parameter block (FP32
accumulation order and rare BF16 rounding flips); features within 1e-5 of a
numpy restatement from the assembled system; training lowers the loss; the
learned guess reaches 1e-8 like every other guess.
"""
import numpy as np
import pytest

import mps_solvers_amd as mps
from oracle import mlp
from scenes import gpu_params, normwise, small_dam

pytestmark = pytest.mark.gpu

BLOCKS = [("W1", 0, 1024), ("b1", 1024, 1088), ("W2", 1088, 5184), ("b2", 5184, 5248),
          ("W3", 5248, 9344), ("b3", 9344, 9408), ("W4", 9408, 9472), ("b4", 9472, 9473)]


def _solver():
    nb2, nb1, pos = small_dam(3)
    return mps.Solver(gpu_params(0, re_lap_mult=3.1, dt_max=5e-4), nb2, nb1, pos)


def test_forward_matches_oracle():
    s = _solver()
    theta = mlp.init(4).astype(np.float32)
    s.mlp_set_params(theta)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((1000, 16)).astype(np.float32)
    y = s.mlp_forward(X)
    yo, _ = mlp.forward(theta.astype(np.float64), X.astype(np.float64), bf16=True)
    assert normwise(y, yo) < 3e-3
    y64, _ = mlp.forward(theta.astype(np.float64), X.astype(np.float64), bf16=False)
    assert normwise(y, y64) < 5e-2      # BF16 operands vs exact


def test_gradient_matches_oracle():
    s = _solver()
    theta = mlp.init(5).astype(np.float32)
    s.mlp_set_params(theta)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((777, 16)).astype(np.float32)
    t = rng.standard_normal(777).astype(np.float32)
    loss, g = s.mlp_grad(X, t)
    lo, go, _ = mlp.loss_and_grad(theta.astype(np.float64), X.astype(np.float64), t.astype(np.float64),
                                  bf16=True)
    assert abs(loss - lo) <= 3e-3 * lo
    for name, a, b in BLOCKS:
        assert normwise(g[a:b], go[a:b]) < 2e-2, name


def test_training_lowers_loss():
    s = _solver()
    s.mlp_init(7)
    # synthetic regression target from the oracle's own network
    rng = np.random.default_rng(3)
    X = rng.standard_normal((20000, 16)).astype(np.float32)
    teacher = mlp.init(11)
    t, _ = mlp.forward(teacher, X.astype(np.float64))
    # train through the device dataset by recording from this synthetic set
    l_first = None
    for it in range(300):
        idx = rng.choice(len(X), 1024, replace=False)
        loss, g = s.mlp_grad(X[idx], t[idx].astype(np.float32))
        if l_first is None:
            l_first = loss
        th = s.mlp_get_params()
        s.mlp_set_params(th - 0.05 * g)
    loss_end, _ = s.mlp_grad(X[:4096], t[:4096].astype(np.float32))
    assert loss_end < 0.5 * l_first


def _features_np(A, st, n0, rho_g, dp, ps):
    rp, col, val, b = A["row_ptr"], A["col"], A["val"], A["b"]
    n = len(b)
    x = st["p"][:n].astype(np.float64)
    z = st["pos"][:n, 2]
    nn = st["n"][:n]
    rows = np.repeat(np.arange(n), np.diff(rp))
    diag = col == rows
    D = np.bincount(rows[diag], val[diag], n)
    W = np.bincount(rows[~diag], val[~diag], n)
    ax = np.bincount(rows, val * x[col], n)
    ident = (np.diff(rp) == 1) & (D == 1.0)
    c = np.where(ident, 0.0, (b - ax) / np.where(D == 0, 1, D))
    od = ~diag
    r_, j_ = rows[od], col[od]
    a_ = val[od]
    dz = z[j_] - z[r_]
    dx = x[j_] - x[r_]
    sc = np.bincount(r_, a_ * c[j_], n)
    sx = np.bincount(r_, a_ * x[j_], n)
    sdev = np.bincount(r_, a_ * (nn[j_] - n0) / n0, n)
    sxz = np.bincount(r_, a_ * dx * dz, n)
    szz = np.bincount(r_, a_ * dz * dz, n)
    scz = np.bincount(r_, a_ * (c[j_] - c[r_]) * dz, n)
    sxx = np.bincount(r_, a_ * dx * dx, n)
    iw = np.where(W > 0, 1 / np.where(W > 0, W, 1), 0)
    izz = np.where(szz > 0, 1 / np.where(szz > 0, szz, 1), 0)
    F = np.stack([x / ps, c / ps, sc * iw / ps, (sx * iw - np.where(W > 0, x, 0)) / ps,
                  100 * (nn - n0) / n0, 100 * sdev * iw, b / (D * ps), W / np.abs(D), -sc / D / ps,
                  sxz * izz / rho_g, scz * izz * dp / ps, np.abs(D) / 60.0,
                  (st["type"][:n] == 1).astype(float), np.diff(rp) / 128.0, sxx * iw / ps ** 2,
                  np.ones(n)], 1)
    F[ident] = 0.0
    return F


def test_features_match_restatement():
    s = _solver()
    s.step_semi(5, alpha=0.0, x0_kind=1, tol=1e-8)
    s.stage_pnd()
    s.ppe_assemble(0.0)
    F = s.mlp_features()
    A = s.ppe_get()
    st = s.state()
    ps = 1000.0 * 9.81 * 64 * s.P.dp
    Fo = _features_np(A, st, s.scalars()["n0"], 1000.0 * 9.81, s.P.dp, ps)
    for k in range(16):
        assert normwise(F[:, k], Fo[:, k]) < 1e-5, k


def test_learned_guess_converges_and_records():
    s = _solver()
    s.mlp_init(1)
    s.mlp_set_record(True)
    s.step_semi(6, alpha=0.0, x0_kind=1, tol=1e-8)
    X, T = s.mlp_dataset()
    assert len(X) > 1000 and np.all(np.isfinite(X)) and np.all(np.isfinite(T))
    l0, l1 = s.mlp_train(200, batch=512, lr=1e-3)
    assert l1 < l0
    s.mlp_set_record(False)
    it, ms = s.step_semi(2, alpha=0.0, x0_kind=3, tol=1e-8)
    assert it > 0
