"""MI355X-native MPS particle time step (host-side mirror of the reference's
MPS_GPU_EXPL / MPS_CPU_EXPL surface).

The compute path is the gfx950 HIP library ``libmps_amd.so`` built in-tree by
``mps_solvers_amd.build``; this module binds its C ABI (``include/mps_amd.h``)
with ctypes.  There is no CPU fallback: every entry point raises when the
library or a gfx950 device is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmps_amd.so")

MPS_OK, MPS_ERR_ARG, MPS_ERR_IO, MPS_ERR_GRID, MPS_ERR_DT = 0, 1, 3, 5, 6
MPS_ERR_HIP, MPS_ERR_NODEV, MPS_ERR_CAP, MPS_ERR_STATE = 7, 8, 9, 10
_ERRORS = {1: "bad argument", 3: "file missing or malformed", 5: "cells too thin",
           6: "dt below dt_min", 7: "HIP runtime error", 8: "no gfx950 device",
           9: "buffer too small", 10: "invalid state"}


class MPSError(RuntimeError):
    def __init__(self, code, where):
        super().__init__(f"{where}: {_ERRORS.get(code, 'error')} (code {code})")
        self.code = code


class Params(C.Structure):
    """mps_params of include/mps_amd.h: the 16 entries of input.dat plus the
    precision and step-order switches."""
    _fields_ = [
        ("dim", C.c_int), ("precision", C.c_int), ("gradp_skip_dummy", C.c_int),
        ("rebin_every_move", C.c_int),
        ("rho", C.c_double), ("nu", C.c_double), ("dp", C.c_double),
        ("re_n_mult", C.c_double), ("re_lap_mult", C.c_double),
        ("beta", C.c_double), ("cs", C.c_double), ("cfl", C.c_double),
        ("col_mult", C.c_double), ("col_rate", C.c_double),
        ("dt_max", C.c_double), ("dt_min_mult", C.c_double),
        ("tt", C.c_double), ("tout", C.c_double), ("max_dis", C.c_double),
        ("g", C.c_double * 3),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "g"}
        d["g"] = tuple(self.g)
        return d


_lib = None
D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)


def lib():
    """Load libmps_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m mps_solvers_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.mps_create.argtypes = [C.POINTER(Params), C.c_int, C.c_int, C.c_int, D, D, D, C.POINTER(vp)]
        for f in ("mps_destroy", "mps_sync", "mps_stage_visc", "mps_stage_collision",
                  "mps_stage_pnd_pres", "mps_stage_gradp", "mps_stage_dt"):
            getattr(L, f).argtypes = [vp]
        L.mps_step.argtypes = [vp, C.c_int]
        L.mps_step_async.argtypes = [vp, C.c_int]
        L.mps_get_np.argtypes = [vp, I, I]
        L.mps_get_state.argtypes = [vp, C.c_int, D, D, D, D, I, I]
        L.mps_set_state.argtypes = [vp, D, D, D]
        L.mps_get_scalars.argtypes = [vp, D]
        L.mps_get_grid.argtypes = [vp, D, I]
        L.mps_write_out.argtypes = [vp, C.c_char_p]
        L.mps_write_vtk.argtypes = [vp, C.c_char_p]
        L.mps_set_motion.argtypes = [vp, C.c_void_p, C.c_int, C.c_double, C.c_double]
        L.mps_scene_wavetank.argtypes = [C.c_int] * 7 + [C.c_double, C.c_double, C.c_ulonglong, I, I, I, D,
                                                         C.c_void_p]
        L.mps_stage_sort.argtypes = [vp, I, I]
        L.mps_set_dt.argtypes = [vp, C.c_double]
        L.mps_neighbor_list.argtypes = [vp, C.c_int, I, I, C.c_long, C.POINTER(C.c_long)]
        L.mps_stage_times.argtypes = [vp, D, C.c_int]
        L.mps_enable_timing.argtypes = [vp, C.c_int]
        L.mps_scene_dambreak.argtypes = [C.c_int] * 9 + [C.c_double, C.c_double, C.c_ulonglong,
                                                         I, I, I, D]
        L.mps_geo_read.argtypes = [C.c_char_p, C.c_int, I, I, I, D, D, D]
        L.mps_geo_write.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, D, D, D]
        L.mps_load_case.argtypes = [C.c_char_p, C.POINTER(Params)]
        L.mps_write_case.argtypes = [C.c_char_p, C.POINTER(Params)]
        L.mps_params_default.argtypes = [C.POINTER(Params)]
        L.mps_stage_pnd.argtypes = [vp]
        L.mps_ppe_assemble.argtypes = [vp, C.c_double, C.POINTER(C.c_long)]
        L.mps_ppe_info.argtypes = [vp, I, C.POINTER(C.c_long)]
        L.mps_ppe_get.argtypes = [vp, I, I, D, D, D, C.POINTER(C.c_ubyte)]
        L.mps_ppe_solve.argtypes = [vp, C.c_int, C.c_int, D, C.c_double, C.c_int, I, D]
        L.mps_ppe_press_corr.argtypes = [vp]
        L.mps_step_semi.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.c_double, I, D]
        F = C.POINTER(C.c_float)
        L.mps_mlp_init.argtypes = [vp, C.c_ulonglong]
        L.mps_mlp_set_params.argtypes = [vp, F]
        L.mps_mlp_get_params.argtypes = [vp, F]
        L.mps_mlp_forward.argtypes = [vp, F, C.c_int, F]
        L.mps_mlp_grad.argtypes = [vp, F, F, C.c_int, F, D]
        L.mps_mlp_features.argtypes = [vp, F, I]
        L.mps_mlp_infer.argtypes = [vp]
        L.mps_mlp_set_record.argtypes = [vp, C.c_int]
        L.mps_mlp_dataset.argtypes = [vp, C.POINTER(C.c_long), F, F]
        L.mps_mlp_dataset_clear.argtypes = [vp]
        L.mps_mlp_dataset_append.argtypes = [vp, F, F, C.c_long]
        L.mps_mlp_train.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_ulonglong, D, D]
        _lib = L
    return _lib


def _check(rc, where):
    if rc != MPS_OK:
        raise MPSError(rc, where)


def _dp(a):
    return None if a is None else a.ctypes.data_as(D)


def _ip(a):
    return None if a is None else a.ctypes.data_as(I)


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def params(**kw) -> Params:
    """Defaults of MPS_CAL_expl_v1.0/input.dat, overridden by keyword."""
    P = Params()
    _check(lib().mps_params_default(C.byref(P)), "mps_params_default")
    for k, v in kw.items():
        if k == "g":
            P.g[:] = list(v)
        else:
            setattr(P, k, v)
    return P


def load_case(path) -> Params:
    """MPS::loadCase (MPS_CAL_expl_v1.0/MPS.cpp:62-223)."""
    P = Params()
    _check(lib().mps_load_case(os.fsencode(path), C.byref(P)), "mps_load_case")
    return P


def write_case(path, P: Params):
    _check(lib().mps_write_case(os.fsencode(path), C.byref(P)), "mps_write_case")


def geo_read(path, as_float=False):
    """MPS::createGeo (MPS_CAL_expl_v1.0/MPS.cpp:225-300) -> (nb2, nb1, pos, vel, press)."""
    n, a, b = C.c_int(), C.c_int(), C.c_int()
    L = lib()
    _check(L.mps_geo_read(os.fsencode(path), int(as_float), C.byref(n), C.byref(a), C.byref(b),
                          None, None, None), "mps_geo_read")
    pos = np.zeros((n.value, 3))
    vel = np.zeros((n.value, 3))
    pr = np.zeros(n.value)
    _check(L.mps_geo_read(os.fsencode(path), int(as_float), C.byref(n), C.byref(a), C.byref(b),
                          _dp(pos), _dp(vel), _dp(pr)), "mps_geo_read")
    return a.value, b.value, pos, vel, pr


def geo_write(path, nb2, nb1, pos, vel=None, press=None):
    pos = np.ascontiguousarray(pos, np.float64)
    n = len(pos)
    vel = None if vel is None else np.ascontiguousarray(vel, np.float64)
    press = None if press is None else np.ascontiguousarray(press, np.float64)
    _check(lib().mps_geo_write(os.fsencode(path), n, nb2, nb1, _dp(pos), _dp(vel), _dp(press)),
           "mps_geo_write")


def dambreak(fluid, inner, layers_wall=2, layers_dummy=2, dp=0.01, jitter=0.01, seed=0, dim=3):
    """Seeded lattice dam break (mps_scene_dambreak): returns (nb2, nb1, pos[np, 3])."""
    fx, fy, fz = fluid
    nx, ny, nz = inner
    L = lib()
    n, a, b = C.c_int(), C.c_int(), C.c_int()
    args = [dim, fx, fy, fz, nx, ny, nz, layers_wall, layers_dummy, dp, jitter, seed]
    _check(L.mps_scene_dambreak(*args, C.byref(n), C.byref(a), C.byref(b), None), "mps_scene_dambreak")
    pos = np.zeros((n.value, 3))
    _check(L.mps_scene_dambreak(*args, C.byref(n), C.byref(a), C.byref(b), _dp(pos)),
           "mps_scene_dambreak")
    return a.value, b.value, pos


def wavetank(nx, ny, nz, depth, travel, layers_wall=2, layers_dummy=2, dp=0.004, jitter=0.01, seed=0):
    """Seeded 3-D wave tank with a piston wall (mps_scene_wavetank): returns
    (nb2, nb1, pos[np, 3], piston[np] as uint8 flags)."""
    L = lib()
    n, a, b = C.c_int(), C.c_int(), C.c_int()
    args = [nx, ny, nz, depth, travel, layers_wall, layers_dummy, dp, jitter, seed]
    _check(L.mps_scene_wavetank(*args, C.byref(n), C.byref(a), C.byref(b), None, None), "mps_scene_wavetank")
    pos = np.zeros((n.value, 3))
    piston = np.zeros(n.value, np.uint8)
    _check(L.mps_scene_wavetank(*args, C.byref(n), C.byref(a), C.byref(b), _dp(pos),
                                piston.ctypes.data_as(C.c_void_p)), "mps_scene_wavetank")
    return a.value, b.value, pos, piston


class Solver:
    """One device context (mps_create ... mps_destroy)."""

    def __init__(self, P: Params, nb2, nb1, pos, vel=None, press=None):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        n = len(pos)
        vel = None if vel is None else np.ascontiguousarray(vel, np.float64).reshape(-1, 3)
        press = None if press is None else np.ascontiguousarray(press, np.float64)
        self.P = P
        self.h = C.c_void_p()
        _check(lib().mps_create(C.byref(P), n, nb2, nb1, _dp(pos), _dp(vel), _dp(press),
                                C.byref(self.h)), "mps_create")
        self.nt = n
        self.nb2, self.nb1 = nb2, nb1

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().mps_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- step
    def step(self, nsteps=1):
        _check(lib().mps_step(self.h, nsteps), "mps_step")

    def step_async(self, nsteps=1):
        _check(lib().mps_step_async(self.h, nsteps), "mps_step_async")

    def sync(self):
        _check(lib().mps_sync(self.h), "mps_sync")

    # ------------------------------------------------------------- getters
    def np_active(self):
        t, a = C.c_int(), C.c_int()
        _check(lib().mps_get_np(self.h, C.byref(t), C.byref(a)), "mps_get_np")
        return a.value

    def state(self, by_id=False):
        n = self.nt
        out = dict(pos=np.zeros((n, 3)), vel=np.zeros((n, 3)), p=np.zeros(n), n=np.zeros(n),
                   id=np.zeros(n, np.int32), type=np.zeros(n, np.int32))
        _check(lib().mps_get_state(self.h, int(by_id), _dp(out["pos"]), _dp(out["vel"]),
                                   _dp(out["p"]), _dp(out["n"]), _ip(out["id"]), _ip(out["type"])),
               "mps_get_state")
        return out

    def set_state(self, pos=None, vel=None, press=None):
        cv = lambda a, s: None if a is None else np.ascontiguousarray(a, np.float64).reshape(s)
        pos, vel, press = cv(pos, (-1, 3)), cv(vel, (-1, 3)), cv(press, (-1,))
        _check(lib().mps_set_state(self.h, _dp(pos), _dp(vel), _dp(press)), "mps_set_state")

    def scalars(self):
        o = np.zeros(10)
        _check(lib().mps_get_scalars(self.h, _dp(o)), "mps_get_scalars")
        keys = ("n0", "lambda", "coef", "dt", "time", "ncells", "ncol", "nexcl", "h", "steps")
        return dict(zip(keys, o.tolist()))

    def grid(self):
        g = np.zeros(4)
        d = np.zeros(4, np.int32)
        _check(lib().mps_get_grid(self.h, _dp(g), _ip(d)), "mps_get_grid")
        return dict(left=g[0], back=g[1], bottom=g[2], h=g[3], dx=int(d[0]), dy=int(d[1]),
                    dz=int(d[2]), ncells=int(d[3]))

    def write_out(self, path):
        _check(lib().mps_write_out(self.h, os.fsencode(path)), "mps_write_out")

    def write_vtk(self, path):
        _check(lib().mps_write_vtk(self.h, os.fsencode(path)), "mps_write_vtk")

    def set_motion(self, moving, axis=0, amplitude=0.05, freq=0.8):
        """Prescribed sinusoidal motion of the flagged particles (flags by
        particle id, e.g. the piston of wavetank()); moving=None switches it off."""
        if moving is None:
            _check(lib().mps_set_motion(self.h, None, int(axis), 0.0, 0.0), "mps_set_motion")
            return
        m = np.ascontiguousarray(moving, np.uint8)
        if m.shape != (self.nt,):
            raise ValueError("moving needs one flag per particle")
        _check(lib().mps_set_motion(self.h, m.ctypes.data_as(C.c_void_p), int(axis), float(amplitude),
                                    float(freq)), "mps_set_motion")

    # -------------------------------------------------------------- stages
    def stage_sort(self):
        perm = np.zeros(self.nt, np.int32)
        cs = np.zeros(self.grid()["ncells"] + 1, np.int32)
        _check(lib().mps_stage_sort(self.h, _ip(perm), _ip(cs)), "mps_stage_sort")
        return perm, cs

    def stage_visc(self):
        _check(lib().mps_stage_visc(self.h), "mps_stage_visc")

    def stage_collision(self):
        _check(lib().mps_stage_collision(self.h), "mps_stage_collision")

    def stage_pnd_pres(self):
        _check(lib().mps_stage_pnd_pres(self.h), "mps_stage_pnd_pres")

    def stage_gradp(self):
        _check(lib().mps_stage_gradp(self.h), "mps_stage_gradp")

    def stage_dt(self):
        _check(lib().mps_stage_dt(self.h), "mps_stage_dt")

    def set_dt(self, dt):
        _check(lib().mps_set_dt(self.h, float(dt)), "mps_set_dt")

    def neighbor_list(self, which):
        na = self.np_active()
        rp = np.zeros(na + 1, np.int32)
        nnz = C.c_long()
        rc = lib().mps_neighbor_list(self.h, which, _ip(rp), None, 0, C.byref(nnz))
        if rc not in (MPS_OK, MPS_ERR_CAP):
            _check(rc, "mps_neighbor_list")
        col = np.zeros(max(nnz.value, 1), np.int32)
        _check(lib().mps_neighbor_list(self.h, which, _ip(rp), _ip(col), nnz.value, C.byref(nnz)),
               "mps_neighbor_list")
        return rp, col[: nnz.value]

    # ------------------------------------------------ pressure Poisson
    def stage_pnd(self):
        _check(lib().mps_stage_pnd(self.h), "mps_stage_pnd")

    def ppe_assemble(self, alpha):
        nnz = C.c_long()
        _check(lib().mps_ppe_assemble(self.h, float(alpha), C.byref(nnz)), "mps_ppe_assemble")
        return nnz.value

    def ppe_get(self):
        n, nnz = C.c_int(), C.c_long()
        _check(lib().mps_ppe_info(self.h, C.byref(n), C.byref(nnz)), "mps_ppe_info")
        rp = np.zeros(n.value + 1, np.int32)
        col = np.zeros(max(nnz.value, 1), np.int32)
        val = np.zeros(max(nnz.value, 1))
        b = np.zeros(n.value)
        x = np.zeros(n.value)
        bc = np.zeros(n.value, np.uint8)
        _check(lib().mps_ppe_get(self.h, _ip(rp), _ip(col), _dp(val), _dp(b), _dp(x),
                                 bc.ctypes.data_as(C.POINTER(C.c_ubyte))), "mps_ppe_get")
        return dict(row_ptr=rp, col=col[: nnz.value], val=val[: nnz.value], b=b, x=x, bc=bc)

    def ppe_solve(self, mixed=True, x0_kind=0, x0=None, tol=1e-8, maxit=100000, check=True):
        """Returns (iterations, relative residual |b - A x| / |b| in FP64)."""
        it, rel = C.c_int(), C.c_double()
        x0 = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        rc = lib().mps_ppe_solve(self.h, int(mixed), int(x0_kind), _dp(x0), float(tol), int(maxit),
                                 C.byref(it), C.byref(rel))
        if check:
            _check(rc, "mps_ppe_solve")
        return it.value, rel.value

    def ppe_press_corr(self):
        _check(lib().mps_ppe_press_corr(self.h), "mps_ppe_press_corr")

    def step_semi(self, nsteps=1, alpha=0.0, x0_kind=1, tol=1e-8):
        """Semi-implicit steps; returns (total CG iterations, solve milliseconds)."""
        it, ms = C.c_int(), C.c_double()
        _check(lib().mps_step_semi(self.h, nsteps, float(alpha), int(x0_kind), float(tol),
                                   C.byref(it), C.byref(ms)), "mps_step_semi")
        return it.value, ms.value

    # ------------------------------------------- learned warm start (MLP)
    def mlp_init(self, seed=0):
        _check(lib().mps_mlp_init(self.h, seed), "mps_mlp_init")

    def mlp_set_params(self, theta):
        th = np.ascontiguousarray(theta, np.float32)
        _check(lib().mps_mlp_set_params(self.h, _fp(th)), "mps_mlp_set_params")

    def mlp_get_params(self):
        th = np.zeros(lib().mps_mlp_nparams(), np.float32)
        _check(lib().mps_mlp_get_params(self.h, _fp(th)), "mps_mlp_get_params")
        return th

    def mlp_forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        y = np.zeros(len(X), np.float32)
        _check(lib().mps_mlp_forward(self.h, _fp(X), len(X), _fp(y)), "mps_mlp_forward")
        return y

    def mlp_grad(self, X, t):
        X = np.ascontiguousarray(X, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(lib().mps_mlp_nparams(), np.float32)
        loss = C.c_double()
        _check(lib().mps_mlp_grad(self.h, _fp(X), _fp(t), len(X), _fp(g), C.byref(loss)), "mps_mlp_grad")
        return loss.value, g

    def mlp_features(self):
        n = C.c_int()
        _check(lib().mps_mlp_features(self.h, None, C.byref(n)), "mps_mlp_features")
        F = np.zeros((n.value, 16), np.float32)
        _check(lib().mps_mlp_features(self.h, _fp(F), C.byref(n)), "mps_mlp_features")
        return F

    def mlp_infer(self):
        _check(lib().mps_mlp_infer(self.h), "mps_mlp_infer")

    def mlp_set_record(self, on=True):
        _check(lib().mps_mlp_set_record(self.h, int(on)), "mps_mlp_set_record")

    def mlp_dataset(self):
        n = C.c_long()
        _check(lib().mps_mlp_dataset(self.h, C.byref(n), None, None), "mps_mlp_dataset")
        X = np.zeros((n.value, 16), np.float32)
        T = np.zeros(n.value, np.float32)
        if n.value:
            _check(lib().mps_mlp_dataset(self.h, C.byref(n), _fp(X), _fp(T)), "mps_mlp_dataset")
        return X, T

    def mlp_dataset_append(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        _check(lib().mps_mlp_dataset_append(self.h, _fp(X), _fp(T), len(X)), "mps_mlp_dataset_append")

    def mlp_dataset_clear(self):
        _check(lib().mps_mlp_dataset_clear(self.h), "mps_mlp_dataset_clear")

    def mlp_train(self, iters, batch=4096, lr=1e-3, seed=0):
        l0, l1 = C.c_double(), C.c_double()
        _check(lib().mps_mlp_train(self.h, iters, batch, lr, seed, C.byref(l0), C.byref(l1)),
               "mps_mlp_train")
        return l0.value, l1.value

    def enable_timing(self, on=True):
        _check(lib().mps_enable_timing(self.h, int(on)), "mps_enable_timing")

    def stage_times(self, reset=False):
        ms = np.zeros(8)
        _check(lib().mps_stage_times(self.h, _dp(ms), int(reset)), "mps_stage_times")
        keys = ("dt", "sort", "visc", "collision", "pnd_pres", "dummy_p", "gradp", "step")
        return dict(zip(keys, ms.tolist()))


class MPS_GPU_EXPL:
    """Mirror of the reference class MPS_GPU_EXPL (MPS_CAL_expl_v1.0/MPS_GPU_EXPL.h:17-37):
    Initial() reads ./input.dat and ./GEO.dat from `workdir`, step() advances one
    explicit step, mps_cal() runs to TT writing ./out/%04d.out every TOUT, and the
    getters of MPS_CAL_/MPS_GPU.h:158-166 hand the particle arrays out."""

    def __init__(self, workdir=".", precision=0, rebin_every_move=0):
        self.workdir = workdir
        self.precision = precision
        self.rebin_every_move = rebin_every_move
        self.solver = None
        self.i_step = 0
        self.d_time = 0.0

    def Initial(self):
        P = load_case(os.path.join(self.workdir, "input.dat"))
        P.precision = self.precision
        P.rebin_every_move = self.rebin_every_move
        nb2, nb1, pos, vel, pr = geo_read(os.path.join(self.workdir, "GEO.dat"),
                                          as_float=(self.precision == 0))
        self.P = P
        self.solver = Solver(P, nb2, nb1, pos, vel, pr)

    def step(self):
        self.solver.step(1)
        self.i_step += 1
        self.d_time = self.solver.scalars()["time"]

    def writeCase(self, path):
        self.solver.write_out(path)

    def mps_cal(self, max_steps=None):
        """MPS_GPU_EXPL::mps_cal (MPS_CAL_expl_v1.0/MPS_GPU_EXPL.cpp:24-80)."""
        if self.solver is None:
            self.Initial()
        out = os.path.join(self.workdir, "out")
        os.makedirs(out, exist_ok=True)
        count, threshold = 0, 0.0
        while self.d_time < self.P.tt and (max_steps is None or self.i_step < max_steps):
            if self.d_time - threshold >= 0.0:
                self.writeCase(os.path.join(out, f"{count:04d}.out"))
                threshold += self.P.tout
                count += 1
            self.step()

    def getNp(self):
        return self.solver.np_active()

    def _st(self):
        return self.solver.state()

    def getPos(self):
        return self._st()["pos"]

    def getVel(self):
        return self._st()["vel"]

    def getPres(self):
        return self._st()["p"]

    def getId(self):
        return self._st()["id"]

    def getType(self):
        return self._st()["type"]
