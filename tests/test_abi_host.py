"""CPU checks of the C ABI: the library loads, exports every declared symbol,
and its host-only entry points (scene generator, input.dat / GEO.dat I/O)
behave; context creation fails loudly without a gfx950 device."""
import ctypes
import os
import re

import numpy as np
import pytest

import mps_solvers_amd as mps

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(REPO, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(REPO, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:int|long|void|double)\s+\**(mps_\w+)\s*\(", txt, re.M))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    lib = mps.lib()
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_create_without_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    nb2, nb1, pos = mps.dambreak((2, 2, 2), (6, 4, 4), 1, 1, 0.01, 0.0, 0)
    with pytest.raises(mps.MPSError) as e:
        mps.Solver(mps.params(), nb2, nb1, pos)
    assert e.value.code == mps.MPS_ERR_NODEV


def test_create_rejects_bad_arguments_before_touching_a_device():
    """Empty, inconsistent or wrongly dimensioned scenes are MPS_ERR_ARG (the
    reference would exit or read out of bounds); checked before any device call."""
    lib = mps.lib()
    P = mps.params()
    pos = np.zeros((4, 3))
    D = ctypes.POINTER(ctypes.c_double)
    h = ctypes.c_void_p()
    pp = pos.ctypes.data_as(D)
    assert lib.mps_create(ctypes.byref(P), 0, 0, 0, pp, None, None, ctypes.byref(h)) == mps.MPS_ERR_ARG
    assert lib.mps_create(ctypes.byref(P), 4, 3, 2, pp, None, None, ctypes.byref(h)) == mps.MPS_ERR_ARG
    assert lib.mps_create(ctypes.byref(P), 4, 0, 0, None, None, None, ctypes.byref(h)) == mps.MPS_ERR_ARG
    P.dim = 4
    assert lib.mps_create(ctypes.byref(P), 4, 0, 0, pp, None, None, ctypes.byref(h)) == mps.MPS_ERR_ARG
    n, a, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert lib.mps_scene_wavetank(10, 4, 8, 9, 2, 2, 2, 0.004, 0.0, 0, ctypes.byref(n), ctypes.byref(a),
                                  ctypes.byref(b), None, None) == mps.MPS_ERR_ARG   # depth > nz
    assert lib.mps_set_motion(None, None, 0, 0.0, 0.0) == mps.MPS_ERR_ARG


def _np_dambreak(fluid, inner, lw, ld, dp, jitter, seed, dim):
    """numpy restatement of mps_scene_dambreak (order dummy, wall, fluid; x fastest)."""
    L = lw + ld
    nx, ny, nz = inner
    fx, fy, fz = fluid
    pts = {0: [], 1: [], 2: []}
    for k in range(-L, nz):
        for j in (range(-L, ny + L) if dim == 3 else [0]):
            for i in range(-L, nx + L):
                di = -i if i < 0 else (i - nx + 1 if i >= nx else 0)
                dj = (-j if j < 0 else (j - ny + 1 if j >= ny else 0)) if dim == 3 else 0
                dk = -k if k < 0 else 0
                d = max(di, dj, dk)
                if d == 0:
                    if i < fx and k < fz and (dim == 2 or j < fy):
                        pts[0].append((i, j, k))
                else:
                    pts[1 if d <= lw else 2].append((i, j, k))
    lat = np.array(pts[2] + pts[1] + pts[0], np.float64)
    pos = lat * dp
    M = (1 << 64) - 1
    n = len(pos)
    for idx in range(n):
        for a in range(3):
            if a == 1 and dim == 2:
                continue
            z = (seed + (3 * idx + a + 1) * 0x9E3779B97F4A7C15) & M
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
            z ^= z >> 31
            u = (z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0
            pos[idx, a] += jitter * dp * u
    return len(pts[2]), len(pts[1]), pos


@pytest.mark.parametrize("dim", [2, 3])
def test_scene_generator_matches_restatement(dim):
    args = ((5, 3, 4), (9, 4, 7), 2, 2, 0.01, 0.01, 17)
    nb2, nb1, pos = mps.dambreak(*args[:2], *args[2:], dim=dim)
    nb2r, nb1r, posr = _np_dambreak(*args, dim)
    assert (nb2, nb1) == (nb2r, nb1r)
    np.testing.assert_array_equal(pos, posr)


def test_config_scenes_have_pinned_counts():
    # config 0: 16,000 fluid + 53,056 wall/dummy; config 1: 1,049,600 + 327,936
    nb2, nb1, pos = mps.dambreak((20, 20, 40), (80, 20, 50), 2, 2, 0.01, 0.01, 0)
    assert (len(pos) - nb2 - nb1, nb2 + nb1) == (16000, 53056)
    L = mps.lib()
    n, a, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert L.mps_scene_dambreak(3, 80, 80, 164, 120, 80, 170, 2, 2, 0.0025, 0.01, 0,
                                ctypes.byref(n), ctypes.byref(a), ctypes.byref(b), None) == 0
    assert (n.value - a.value - b.value, a.value + b.value) == (1049600, 327936)


def test_input_and_geo_round_trip(tmp_path):
    P = mps.params(dim=3, dp=0.0025, re_lap_mult=4.0, cs=20.0)
    mps.write_case(tmp_path / "input.dat", P)
    Q = mps.load_case(tmp_path / "input.dat")
    for k in ("dim", "rho", "nu", "dp", "re_n_mult", "re_lap_mult", "beta", "cs", "cfl",
              "col_mult", "col_rate", "dt_max", "dt_min_mult", "tt", "tout", "max_dis"):
        assert getattr(Q, k) == getattr(P, k), k
    nb2, nb1, pos = mps.dambreak((3, 3, 3), (6, 4, 5), 1, 1, 0.01, 0.01, 1)
    rng = np.random.default_rng(0)
    vel = rng.standard_normal(pos.shape)
    pr = rng.standard_normal(len(pos))
    mps.geo_write(tmp_path / "GEO.dat", nb2, nb1, pos, vel, pr)
    a, b, p2, v2, r2 = mps.geo_read(tmp_path / "GEO.dat")
    assert (a, b) == (nb2, nb1)
    np.testing.assert_allclose(p2, pos, rtol=1e-8)
    np.testing.assert_allclose(v2, vel, rtol=1e-8)
    np.testing.assert_allclose(r2, pr, rtol=1e-8)
    # float parsing rounds each token to the nearest float, as the float build's %f does
    a, b, p3, _, _ = mps.geo_read(tmp_path / "GEO.dat", as_float=True)
    np.testing.assert_array_equal(p3, p2.astype(np.float32).astype(np.float64))


def test_reference_input_files_parse():
    # formats of the reference's own files (copied values, not files): MPS_CAL_expl_v1.0/input.dat
    txt = ("DIMENTION\n3\nFLUID RHO\n1000.0\nFLUID DYNAMICS VISCOSITY\n1e-6\n"
           "INITIAL PARTICLE DISTANCE(DP)\n0.006\nR0(times of DP)\n2.1\nRLAP(times of DP)\n2.1\n"
           "BETA\n0.97\nNUMERICAL SPEED OF SOUND\n10\nCFL NUMBER\n0.1\n"
           "COLLISION DISTANCE(times of DP)\n0.8\nCOLLISION RATE\n2.0\nMAX TIME STEP\n0.0001\n"
           "MIN TIME STEP(times of MAX)\n0.001\nTT TOTAL CALCULATION TIME\n0.5\n"
           "TOUT TIME STEP OF OUTPUT\n0.001\nMAXIMUM DISPLACEMENT OF THE GEOMETRY\n0.001\n")
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        fn = os.path.join(td, "input.dat")
        open(fn, "w").write(txt)
        P = mps.load_case(fn)
    assert (P.dim, P.dp, P.re_n_mult, P.cs, P.col_mult, P.dt_max) == (3, 0.006, 2.1, 10.0, 0.8, 1e-4)
