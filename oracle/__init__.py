"""CPU oracle for the MPS particle time step -- TEST INFRASTRUCTURE ONLY.

ctypes bindings to ``oracle/liboracle.so`` (built from ``mps_oracle.c`` by
``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker or the timed CPU baseline; the product package ``mps_solvers_amd``
never imports it.

Each wrapped C function cites the reference file:line it restates (see
``mps_oracle.c``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)


class Params(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("f32", C.c_int), ("gradp_skip_dummy", C.c_int), ("pad_", C.c_int),
        ("rho", C.c_double), ("nu", C.c_double), ("dp", C.c_double),
        ("re_n_mult", C.c_double), ("re_lap_mult", C.c_double),
        ("beta", C.c_double), ("cs", C.c_double), ("cfl", C.c_double),
        ("col_mult", C.c_double), ("col_rate", C.c_double),
        ("dt_max", C.c_double), ("dt_min_mult", C.c_double),
        ("max_dis", C.c_double), ("g", C.c_double * 3),
    ]


class Grid(C.Structure):
    _fields_ = [
        ("left", C.c_double), ("back", C.c_double), ("bottom", C.c_double), ("h", C.c_double),
        ("dx", C.c_int), ("dy", C.c_int), ("dz", C.c_int), ("sheet", C.c_int),
        ("ncells", C.c_int), ("pad_", C.c_int),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _lib.orc_update_dt.restype = C.c_double
        _lib.orc_build_poisson.restype = C.c_long
        _lib.orc_sim_create.restype = C.c_void_p
        _lib.orc_sim_create.argtypes = [C.POINTER(Params), C.c_int, C.c_int, C.c_int] + [D] * 7
        for f in ("orc_sim_destroy", "orc_sim_np"):
            getattr(_lib, f).argtypes = [C.c_void_p]
        _lib.orc_sim_step.argtypes = [C.c_void_p, C.c_int, C.c_int]
        _lib.orc_sim_step_semi.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, I]
        _lib.orc_sim_get.argtypes = [C.c_void_p] + [D] * 8 + [I, I]
        _lib.orc_sim_consts.argtypes = [C.c_void_p, D]
        _lib.orc_sim_grid.argtypes = [C.c_void_p, C.POINTER(Grid)]
        _lib.orc_sim_set_motion.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double]
    return _lib


def params(**kw) -> Params:
    """Reference defaults of MPS_CAL_expl_v1.0/input.dat, overridable by keyword."""
    base = dict(dim=3, f32=0, gradp_skip_dummy=0, rho=1000.0, nu=1e-6, dp=0.01,
                re_n_mult=2.1, re_lap_mult=4.0, beta=0.97, cs=20.0, cfl=0.1,
                col_mult=0.8, col_rate=2.0, dt_max=1e-4, dt_min_mult=1e-3,
                max_dis=0.001, g=(0.0, 0.0, -9.81))
    base.update(kw)
    P = Params()
    for k, v in base.items():
        if k == "g":
            P.g[:] = list(v)
        else:
            setattr(P, k, v)
    return P


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(D)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(I)


def thresholds(P: Params) -> np.ndarray:
    o = (C.c_double * 6)()
    lib().orc_thresholds(C.byref(P), o)
    return np.array(list(o))


def grid_init(P, x, y, z) -> Grid:
    G = Grid()
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    rc = lib().orc_grid_init(C.byref(P), len(x), _d(x), _d(y), _d(z), C.byref(G))
    if rc:
        raise RuntimeError(f"orc_grid_init failed with code {rc}")
    return G


def hash_cells(P, G, x, y, z, type_=None):
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    key = np.zeros(len(x), np.int32)
    t = None if type_ is None else np.ascontiguousarray(type_, np.int32)
    lib().orc_hash(C.byref(P), C.byref(G), len(x), _d(x), _d(y), _d(z),
                   None if t is None else _i(t), key.ctypes.data_as(I))
    return key


def sort_cells(key, ncells):
    key = np.ascontiguousarray(key, np.int32)
    n = len(key)
    perm = np.zeros(n, np.int32)
    index = np.zeros(n, np.int32)
    cs = np.zeros(ncells + 1, np.int32)
    lib().orc_sort(n, ncells, _i(key), perm.ctypes.data_as(I), index.ctypes.data_as(I),
                   cs.ctypes.data_as(I))
    return perm, index, cs


def neighbors(P, G, x, y, z, cs, which):
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    n = len(x)
    rp = np.zeros(n + 1, np.int32)
    cap = n * 64
    while True:
        col = np.zeros(cap, np.int32)
        rc = lib().orc_neighbors(C.byref(P), C.byref(G), n, _d(x), _d(y), _d(z), _i(cs),
                                 which, rp.ctypes.data_as(I), col.ctypes.data_as(I),
                                 C.c_long(cap))
        if rc == 0:
            return rp, col[: rp[-1]]
        cap = int(rp[-1]) + 16


def calonce(P, G, x, y, z, type_, cs):
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    n = len(x)
    nn = np.zeros(n)
    n0 = C.c_double()
    lam = C.c_double()
    am = C.c_int()
    normal = np.zeros(n, np.int32)
    lib().orc_calonce(C.byref(P), C.byref(G), n, _d(x), _d(y), _d(z), _i(type_), _i(cs),
                      nn.ctypes.data_as(D), C.byref(n0), C.byref(lam), C.byref(am),
                      normal.ctypes.data_as(I))
    return nn, n0.value, lam.value, am.value, normal


def pnd(P, G, x, y, z, type_, cs, n_init=None):
    x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
    n = np.zeros(len(x)) if n_init is None else np.array(n_init, np.float64)
    lib().orc_pnd(C.byref(P), C.byref(G), len(x), _d(x), _d(y), _d(z), _i(type_), _i(cs),
                  n.ctypes.data_as(D))
    return n


def pres_expl(P, n, n0, type_, normal, p_init=None):
    p = np.zeros(len(n)) if p_init is None else np.array(p_init, np.float64)
    lib().orc_pres_expl(C.byref(P), len(n), _d(n), C.c_double(n0), _i(type_), _i(normal),
                        p.ctypes.data_as(D))
    return p


def _six(n):
    return [np.zeros(n) for _ in range(6)]


def visc(P, G, st, cs, coef, dt):
    out = _six(len(st["x"]))
    lib().orc_visc(C.byref(P), C.byref(G), len(st["x"]), *[_d(st[k]) for k in "xyzuvw"],
                   _i(st["type"]), _i(cs), C.c_double(coef), C.c_double(dt),
                   *[o.ctypes.data_as(D) for o in out])
    return dict(zip("xyzuvw", out))


def lapvel(P, G, st, cs, coef):
    out = [np.zeros(len(st["x"])) for _ in range(3)]
    lib().orc_lapvel(C.byref(P), C.byref(G), len(st["x"]), *[_d(st[k]) for k in "xyzuvw"],
                     _i(cs), C.c_double(coef), *[o.ctypes.data_as(D) for o in out])
    return out


def gradp_vec(P, G, st, p, cs, n0):
    out = [np.zeros(len(st["x"])) for _ in range(3)]
    lib().orc_gradp_vec(C.byref(P), C.byref(G), len(st["x"]), *[_d(st[k]) for k in "xyz"],
                        _d(p), _i(st["type"]), _i(cs), C.c_double(n0),
                        *[o.ctypes.data_as(D) for o in out])
    return out


def collision(P, G, st, cs, dt):
    out = _six(len(st["x"]))
    ncol = lib().orc_collision(C.byref(P), C.byref(G), len(st["x"]),
                               *[_d(st[k]) for k in "xyzuvw"], _i(st["type"]), _i(cs),
                               C.c_double(dt), *[o.ctypes.data_as(D) for o in out])
    return dict(zip("xyzuvw", out)), ncol


def gradp(P, G, st, p, cs, n0, dt):
    out = _six(len(st["x"]))
    lib().orc_gradp(C.byref(P), C.byref(G), len(st["x"]), *[_d(st[k]) for k in "xyzuvw"],
                    _d(p), _i(st["type"]), _i(cs), C.c_double(n0), C.c_double(dt),
                    *[o.ctypes.data_as(D) for o in out])
    return dict(zip("xyzuvw", out))


def update_dt(P, u, v, w):
    small = C.c_int()
    dt = lib().orc_update_dt(C.byref(P), len(u), _d(u), _d(v), _d(w), C.byref(small))
    return dt, bool(small.value)


def build_poisson(P, G, st, cs, n, p_prev, n0, lam, dt, alpha):
    x, y, z = (np.ascontiguousarray(st[k], np.float64) for k in "xyz")
    N = len(x)
    rp = np.zeros(N + 1, np.int32)
    b = np.zeros(N)
    x0 = np.zeros(N)
    bc = np.zeros(N, np.uint8)
    cap = N * 200
    while True:
        col = np.zeros(cap, np.int32)
        val = np.zeros(cap)
        k = lib().orc_build_poisson(
            C.byref(P), C.byref(G), N, _d(x), _d(y), _d(z), _i(st["type"]), _i(cs), _d(n),
            None if p_prev is None else _d(p_prev), C.c_double(n0), C.c_double(lam),
            C.c_double(dt), C.c_double(alpha), rp.ctypes.data_as(I), col.ctypes.data_as(I),
            val.ctypes.data_as(D), b.ctypes.data_as(D), x0.ctypes.data_as(D),
            bc.ctypes.data_as(C.POINTER(C.c_ubyte)), C.c_long(cap))
        if k >= 0:
            return rp, col[:k], val[:k], b, x0, bc
        cap = -k + 16


def cg(rp, col, val, b, x0, tol, relative=True, maxit=100000):
    x = np.array(x0, np.float64)
    rr = C.c_double()
    it = lib().orc_cg(len(b), _i(rp), _i(col), _d(val), _d(b), x.ctypes.data_as(D),
                      C.c_double(tol), int(relative), maxit, C.byref(rr))
    return x, it, rr.value


def spmv(rp, col, val, x):
    y = np.zeros(len(rp) - 1)
    lib().orc_spmv(len(y), _i(rp), _i(col), _d(val), _d(x), y.ctypes.data_as(D))
    return y


class Sim:
    """Whole-step explicit simulator (MPS_GPU_EXPL semantics, two-pass)."""

    def __init__(self, P, nb2, nb1, x, y, z, u=None, v=None, w=None, p=None):
        n = len(x)
        z0 = np.zeros(n)
        self._keep = [np.ascontiguousarray(a, np.float64) for a in
                      (x, y, z, z0 if u is None else u, z0 if v is None else v,
                       z0 if w is None else w, z0 if p is None else p)]
        self.P = P
        self.n = n
        self.h = lib().orc_sim_create(C.byref(P), n, nb2, nb1, *[a.ctypes.data_as(D) for a in self._keep])
        if not self.h:
            raise RuntimeError("orc_sim_create failed")

    def step(self, nsteps=1, rebin_every_move=False):
        rc = lib().orc_sim_step(self.h, nsteps, int(rebin_every_move))
        if rc:
            raise RuntimeError("dt below dt_min")

    def step_semi(self, nsteps=1, alpha=0.0, tol=1e-8, rebin_every_move=False):
        """Semi-implicit steps; returns the summed CG iterations."""
        it = C.c_int()
        rc = lib().orc_sim_step_semi(self.h, nsteps, int(rebin_every_move), float(alpha), float(tol), C.byref(it))
        if rc == 1:
            raise RuntimeError("dt below dt_min")
        if rc == 2:
            raise RuntimeError("CG did not converge")
        return it.value

    def state(self):
        n = self.n
        out = {k: np.zeros(n) for k in ("x", "y", "z", "u", "v", "w", "p", "n")}
        t = np.zeros(n, np.int32)
        idv = np.zeros(n, np.int32)
        lib().orc_sim_get(self.h, *[out[k].ctypes.data_as(D) for k in ("x", "y", "z", "u", "v", "w", "p", "n")],
                          t.ctypes.data_as(I), idv.ctypes.data_as(I))
        out["type"] = t
        out["id"] = idv
        out["np"] = lib().orc_sim_np(self.h)
        return out

    def consts(self):
        o = (C.c_double * 6)()
        lib().orc_sim_consts(self.h, o)
        return dict(zip(("n0", "lambda", "dt", "time", "coef", "ncells"), list(o)))

    def grid(self):
        G = Grid()
        lib().orc_sim_grid(self.h, C.byref(G))
        return G

    def set_motion(self, moving, axis=0, amplitude=0.05, freq=0.8):
        """Prescribed sinusoidal motion of the particles flagged by id (None: off)."""
        if moving is None:
            lib().orc_sim_set_motion(self.h, None, int(axis), 0.0, 0.0)
            return
        self._mov = np.ascontiguousarray(moving, np.uint8)
        lib().orc_sim_set_motion(self.h, self._mov.ctypes.data_as(C.c_void_p), int(axis), float(amplitude),
                                 float(freq))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_destroy(self.h)
            self.h = None
