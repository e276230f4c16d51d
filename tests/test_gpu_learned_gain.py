"""What the learned CG warm start gains, measured on a held-out scene.

The network is trained on (features, converged correction) pairs recorded from
one seeded dam break and evaluated on another.  Asserted: its guess leaves a
smaller initial residual |b - A x0| / |b| than the previous-step pressure,
which leaves a smaller one than zero; every guess still reaches the same FP64
residual of 1e-8; the learned guess needs no more CG iterations than zero.
Reported (printed): the three initial residuals and iteration counts.  The
iteration gain is small by construction: CG on these matrices spends ~9
iterations per decade of residual whatever the start, and a guess only saves
the decades its initial residual already has.
"""
import numpy as np
import pytest

import mps_solvers_amd as mps
from scenes import normwise

pytestmark = pytest.mark.gpu

DP = 0.0025
ALPHA = 1.0 / (1000.0 * 20.0 ** 2)      # the reference's compressibility 1/(rho c^2)


def _params():
    return mps.params(dim=3, precision=0, rho=1000.0, nu=1e-6, dp=DP, re_n_mult=2.1, re_lap_mult=3.1,
                      beta=0.97, cs=20.0, cfl=0.1, col_mult=0.8, col_rate=2.0, dt_max=2.5e-4,
                      dt_min_mult=1e-3, max_dis=0.1 * DP, g=(0.0, 0.0, -9.81))


def _solver(fluid, inner, seed):
    nb2, nb1, pos = mps.dambreak(fluid, inner, 2, 2, DP, 0.0, seed)
    return mps.Solver(_params(), nb2, nb1, pos)


def _spmv(A, x):
    rp, col, val = A["row_ptr"], A["col"], A["val"]
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    return np.bincount(rows, val * x[col], len(rp) - 1)


def test_learned_guess_gain():
    train = _solver((24, 24, 40), (36, 24, 44), 11)
    train.mlp_init(0)
    train.step_semi(8, alpha=ALPHA, x0_kind=1, tol=1e-8)
    train.mlp_set_record(True)
    train.step_semi(12, alpha=ALPHA, x0_kind=1, tol=1e-8)
    X, T = train.mlp_dataset()
    train.close()
    assert len(X) > 10000 and np.all(np.isfinite(X)) and np.all(np.isfinite(T))

    s = _solver((28, 28, 44), (42, 28, 48), 12)
    s.mlp_init(0)
    s.mlp_dataset_append(X, T)
    l0, l1 = s.mlp_train(600, batch=2048, lr=1e-3, seed=0)
    assert l1 < 0.5 * l0
    s.step_semi(15, alpha=ALPHA, x0_kind=1, tol=1e-8)
    # the next step's system, as mps_step_semi builds it
    s.stage_dt()
    s.stage_sort()
    s.stage_visc()
    s.stage_collision()
    s.stage_pnd()
    s.ppe_assemble(ALPHA)
    A = s.ppe_get()
    b, bc = A["b"], A["bc"].astype(bool)
    n = len(b)
    p_prev = s.state()["p"][:n].astype(np.float64)
    rp = A["row_ptr"]
    rows = np.repeat(np.arange(n), np.diff(rp))
    diag = np.bincount(rows[A["col"] == rows], A["val"][A["col"] == rows], n)
    ident = (np.diff(rp) == 1) & (diag == 1.0)

    # the learned guess, rebuilt on the host: x0 = p_prev + P_s y, identity rows 0
    s.mlp_infer()
    F = s.mlp_features()
    y = s.mlp_forward(F).astype(np.float64)
    ps = 1000.0 * 9.81 * 64 * DP
    x_learned = np.where(ident, 0.0, p_prev + ps * y)
    x_prev = np.where(bc, 0.0, p_prev)

    bn = np.linalg.norm(b)
    r_zero = 1.0
    r_prev = np.linalg.norm(b - _spmv(A, x_prev)) / bn
    r_learned = np.linalg.norm(b - _spmv(A, x_learned)) / bn

    it_zero, rel_zero = s.ppe_solve(mixed=True, x0_kind=0, tol=1e-8)
    it_prev, rel_prev = s.ppe_solve(mixed=True, x0_kind=1, tol=1e-8)
    s.mlp_infer()
    it_learned, rel_learned = s.ppe_solve(mixed=True, x0_kind=3, tol=1e-8)
    it_host, _ = s.ppe_solve(mixed=True, x0_kind=2, x0=x_learned, tol=1e-8)
    x_sol = s.ppe_get()["x"]

    print(f"\n[learned guess] rows {n}: initial |r|/|b| zero {r_zero:.3f}, previous p {r_prev:.3f}, "
          f"learned {r_learned:.3f}; CG iterations to 1e-8: zero {it_zero}, previous {it_prev}, "
          f"learned {it_learned}; guess error vs converged p: previous "
          f"{normwise(x_prev, x_sol):.3f}, learned {normwise(x_learned, x_sol):.3f}")

    assert max(rel_zero, rel_prev, rel_learned) <= 1e-8
    assert r_prev < r_zero
    assert r_learned < r_prev
    assert normwise(x_learned, x_sol) < normwise(x_prev, x_sol)
    assert it_learned <= it_zero
    assert abs(it_host - it_learned) <= 2      # the host rebuild is the guess the device uses
    s.close()
