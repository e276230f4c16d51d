"""Scenes and oracle plumbing shared by the tests."""
import numpy as np

import mps_solvers_amd as mps
import oracle

# config-0 constants (re_n = 2.1 l0, re_lap = 4.0 l0, rho 1000, nu 1e-6, c 20)
CFG0 = dict(dim=3, rho=1000.0, nu=1e-6, dp=0.01, re_n_mult=2.1, re_lap_mult=4.0, beta=0.97,
            cs=20.0, cfl=0.1, col_mult=0.8, col_rate=2.0, dt_max=1e-4, dt_min_mult=1e-3,
            max_dis=0.001, g=(0.0, 0.0, -9.81))


def small_dam(seed=3, fluid=(6, 6, 8), inner=(12, 6, 10), dp=0.01, dim=3, jitter=0.01):
    return mps.dambreak(fluid, inner, 2, 2, dp, jitter, seed, dim)


def gpu_params(precision=0, **kw):
    d = dict(CFG0)
    d.update(kw)
    d["precision"] = precision
    return mps.params(**d)


def oracle_params(precision=0, **kw):
    d = dict(CFG0)
    d.update(kw)
    d.pop("tt", None)
    d.pop("tout", None)
    d.pop("rebin_every_move", None)
    d.pop("precision", None)
    d["f32"] = 1 if precision == 0 else 0
    return oracle.params(**d)


def to_float_exact(a, precision):
    """Values as the device holds them (float for the mixed path)."""
    return np.asarray(a, np.float32).astype(np.float64) if precision == 0 else np.asarray(a, np.float64)


def oracle_view(st):
    """Solver.state() (sorted order) -> dict of oracle arrays."""
    d = {k: st["pos"][:, a].copy() for a, k in enumerate("xyz")}
    d.update({k: st["vel"][:, a].copy() for a, k in enumerate("uvw")})
    d["type"] = st["type"].astype(np.int32)
    d["p"] = st["p"].copy()
    d["n"] = st["n"].copy()
    return d


def oracle_grid(solver):
    g = solver.grid()
    G = oracle.Grid()
    G.left, G.back, G.bottom, G.h = g["left"], g["back"], g["bottom"], g["h"]
    G.dx, G.dy, G.dz = g["dx"], g["dy"], g["dz"]
    G.sheet = g["dx"] * g["dy"]
    G.ncells = g["ncells"]
    return G


def normwise(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a - b)))
