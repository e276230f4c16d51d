#!/usr/bin/env python3
"""Benchmark of the MI355X MPS particle time step.

Workload (default): the 3-D dam break scaled to 1,049,600 fluid particles
(80 x 80 x 164 lattice, l0 = 2.5 mm) in an open-top tank of 120 x 80 x 170
inner lattice sites lined with 2 wall + 2 dummy layers: N = 1,377,536
particles.  re_n = 2.1 l0, re_lap = 4.0 l0, rho = 1000, nu = 1e-6,
g = -9.81 z, c = 20 m/s, dt_max = 2.5e-5 s (acoustic Courant number 0.2 as
at l0 = 10 mm with dt = 1e-4 s), FP32 state with FP64 n and (n - n0)/n0.
A step is one explicit MPS step (re-binning, viscosity, collision, PND +
pressure, pressure gradient + update) over the resident state.

Prints ONE JSON line: value = fluid particle-updates/s over the timed steps
(inputs resident in HBM), roofline of the dominant kernel from HIP events on
the library's stream (HBM, with the VALU issue rate beside it), the FP32 path
against the FP64 path, the pressure Poisson solve (zero / previous-step /
learned guess), and the CPU oracle timed on this host on a bounded sample of
the same workload.

    python bench.py [--gpus 1] [--steps 200] [--warmup 10] [--no-cpu] [--no-ppe]
    python bench.py --config 4 [--wavetank-nx 1000]   # semi-implicit wave tank line
    python bench.py --ppe-only                          # the pressure Poisson part only
    python bench.py --copy-calib 1024                   # 1 GiB device copy (counter scale)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per particle per launch (FP32 pos/vel 12 B, FP64 n 8 B,
# FP32 p 4 B, type/id/key/perm 4 B; neighbour gathers counted as cache hits)
ALG_BYTES = {"visc": 52, "collision": 52, "pnd_pres": 28, "gradp": 56, "sort": 100}
STEP_BYTES = 288

WORKLOADS = {
    1: dict(slug="dambreak_1M", name="dam break 1.05M fluid (80x80x164, l0=2.5mm), tank 120x80x170 inner, 2+2 wall layers",
            fluid=(80, 80, 164), inner=(120, 80, 170), dp=0.0025, dt_max=2.5e-5, max_dis=0.00025),
    0: dict(slug="dambreak_16k", name="dam break 16k fluid (20x20x40, l0=10mm), tank 80x20x50 inner, 2+2 wall layers",
            fluid=(20, 20, 40), inner=(80, 20, 50), dp=0.01, dt_max=1e-4, max_dis=0.001),
}


def params_for(mps, w, precision=0):
    return mps.params(dim=3, precision=precision, rho=1000.0, nu=1e-6, dp=w["dp"], re_n_mult=2.1,
                      re_lap_mult=4.0, beta=0.97, cs=20.0, cfl=0.1, col_mult=0.8, col_rate=2.0,
                      dt_max=w["dt_max"], dt_min_mult=1e-3, max_dis=w["max_dis"],
                      g=(0.0, 0.0, -9.81))


def cpu_baseline(w, nb2, nb1, pos, budget_s=20.0):
    """Oracle (C/OpenMP restatement of the reference CPU path) on this host's
    cores, on a bounded sample: whole explicit steps of the same scene."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    P = oracle.params(dim=3, f32=1, rho=1000.0, nu=1e-6, dp=w["dp"], re_n_mult=2.1,
                      re_lap_mult=4.0, beta=0.97, cs=20.0, cfl=0.1, col_mult=0.8, col_rate=2.0,
                      dt_max=w["dt_max"], dt_min_mult=1e-3, max_dis=w["max_dis"])
    sim = oracle.Sim(P, nb2, nb1, pos[:, 0], pos[:, 1], pos[:, 2])
    nfl = len(pos) - nb2 - nb1
    steps, t0 = 0, time.perf_counter()
    while True:
        sim.step(1)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 50:
            break
    return {"value": nfl * steps / el, "unit": "particle-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{steps} explicit step(s) of the same scene after initialisation, "
                      f"{el:.1f} s on {threads} OpenMP threads (oracle/mps_oracle.c, FP64 sums)"}


# Unperturbed lattices for the semi-implicit runs: n0 is the largest initial
# number density (as the reference's calOnce takes it), so on a perturbed
# lattice nearly every particle starts below n0, the source term is
# negative almost everywhere and p = max(x, 0) stays zero.
PPE_JITTER = 0.0
# compressibility of the Poisson matrix: alpha = 1 / (rho c^2) (calOnce), rho 1000, c 20
PPE_ALPHA = 1.0 / (1000.0 * 20.0 ** 2)
PPE_SCENES = {
    # config 2: same scene family at ~0.5M fluid (64 x 64 x 128), re_lap = 3.1 l0
    "eval": dict(fluid=(64, 64, 128), inner=(96, 64, 134), dp=0.0025, seed=2),
    # training runs for the learned guess (seeded dam breaks of ~100k-200k fluid)
    "train": [dict(fluid=(40, 40, 64), inner=(60, 40, 70), dp=0.0025, seed=11),
              dict(fluid=(48, 48, 80), inner=(72, 48, 86), dp=0.0025, seed=12)],
}


def ppe_params(mps, dp):
    return mps.params(dim=3, precision=0, rho=1000.0, nu=1e-6, dp=dp, re_n_mult=2.1,
                      re_lap_mult=3.1, beta=0.97, cs=20.0, cfl=0.1, col_mult=0.8, col_rate=2.0,
                      dt_max=2.5e-4, dt_min_mult=1e-3, max_dis=dp * 0.1, g=(0.0, 0.0, -9.81))


def ppe_bench(mps, spin_steps=20, train_iters=2000, log=print):
    """Pressure Poisson solve to |b - Ax| <= 1e-8 |b| (FP64 residual) on the
    config-2 scene: mixed-precision CG from zero (plain), from the previous
    pressure, and from the learned guess (inference time included); full FP64
    CG from zero for reference."""
    t0 = time.perf_counter()
    Xs, Ts = [], []
    for sc in PPE_SCENES["train"]:
        nb2, nb1, pos = mps.dambreak(sc["fluid"], sc["inner"], 2, 2, sc["dp"], PPE_JITTER, sc["seed"])
        s = mps.Solver(ppe_params(mps, sc["dp"]), nb2, nb1, pos)
        s.mlp_init(0)
        s.step_semi(10, alpha=PPE_ALPHA, x0_kind=1, tol=1e-8)
        s.mlp_set_record(True)
        s.step_semi(spin_steps, alpha=PPE_ALPHA, x0_kind=1, tol=1e-8)
        X, T = s.mlp_dataset()
        Xs.append(X)
        Ts.append(T)
        s.close()
        log(f"[bench] recorded {len(X)} samples from a {len(pos)}-particle dam break")
    sc = PPE_SCENES["eval"]
    nb2, nb1, pos = mps.dambreak(sc["fluid"], sc["inner"], 2, 2, sc["dp"], PPE_JITTER, sc["seed"])
    s = mps.Solver(ppe_params(mps, sc["dp"]), nb2, nb1, pos)
    s.mlp_init(0)
    s.mlp_dataset_append(np.concatenate(Xs), np.concatenate(Ts))
    t1 = time.perf_counter()
    l0, l1 = s.mlp_train(train_iters, batch=4096, lr=1e-3, seed=0)
    train_s = time.perf_counter() - t1
    log(f"[bench] trained {train_iters} Adam steps in {train_s:.1f} s: loss {l0:.3e} -> {l1:.3e}")
    s.step_semi(10 + spin_steps, alpha=PPE_ALPHA, x0_kind=1, tol=1e-8)
    # the system of the next step, built as mps_step_semi builds it
    s.stage_dt()
    s.stage_sort()
    s.stage_visc()
    s.stage_collision()
    s.stage_pnd()
    nnz = s.ppe_assemble(PPE_ALPHA)
    n = len(s.ppe_get()["b"])
    res = {}
    for name, mixed, kind in (("plain_zero", True, 0), ("prev_step", True, 1), ("learned", True, 3),
                              ("fp64_zero", False, 0)):
        best = None
        for rep in range(3):
            t2 = time.perf_counter()
            if kind == 3:
                s.mlp_infer()
            it, rel = s.ppe_solve(mixed=mixed, x0_kind=kind, tol=1e-8)
            ms = 1e3 * (time.perf_counter() - t2)
            if best is None or ms < best["ms"]:
                best = {"ms": round(ms, 3), "iters": it, "relres": rel}
        res[name] = best
    s.close()
    # inner-iteration roofline of the plain solve: 8 nnz + 56 N algorithmic bytes per
    # FP32 CG iteration (CSR values + columns, six N-vectors of 4 B, two FP64 partial sets)
    pz = res["plain_zero"]
    it_ms = pz["ms"] / max(pz["iters"], 1)
    it_bytes = 8.0 * nnz + 56.0 * n
    return {"cg_iteration": {"ms": it_ms, "alg_bytes": it_bytes,
                             "achieved_GBs": it_bytes / (it_ms * 1e-3) / 1e9,
                             "frac_hbm": it_bytes / (it_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                             "note": "solve ms / inner iterations: includes the FP64 residual passes"},
            "workload": "dam break 0.52M fluid (64x64x128, l0=2.5mm), re_lap=3.1 l0, PPE "
                        f"(alpha = 1/(rho c^2) as MPS_CPU_EXPL), after {10 + spin_steps} semi-implicit "
                        "steps of dt=2.5e-4 s",
            "rows": n, "nnz": int(nnz), "solves": res,
            "learned_over_plain_time": res["learned"]["ms"] / res["plain_zero"]["ms"],
            "learned_over_prev_time": res["learned"]["ms"] / res["prev_step"]["ms"],
            "train": {"samples": int(sum(len(x) for x in Xs)), "adam_steps": train_iters,
                      "loss_first": l0, "loss_last": l1, "train_s": round(train_s, 2),
                      "total_s": round(time.perf_counter() - t0, 2)}}


# Config 4: wave tank 4.0 x 0.3 x 0.6 m inner at l0 = 4 mm (1000 x 75 x 150 sites),
# water 0.3 m deep (5.625M fluid: the stated tank and depth at this l0), piston
# x(t) = 0.05 sin(2 pi 0.8 t) on the x = 0 wall (13 sites of travel), 2 + 2 wall
# layers, unperturbed lattice, semi-implicit step with the learned CG guess.
WAVETANK = dict(name="wave tank 4.0x0.3x0.6 m inner, l0=4mm, water depth 0.3 m (1000x75x75 fluid), "
                     "piston 0.05 sin(2 pi 0.8 t), 2+2 wall layers",
                inner=(1000, 75, 150), depth=75, travel=13, dp=0.004, amp=0.05, freq=0.8)


def wavetank_params(mps, dp, precision=0):
    return mps.params(dim=3, precision=precision, rho=1000.0, nu=1e-6, dp=dp, re_n_mult=2.1,
                      re_lap_mult=3.1, beta=0.97, cs=20.0, cfl=0.1, col_mult=0.8, col_rate=2.0,
                      dt_max=4e-4, dt_min_mult=1e-3, max_dis=dp * 4, g=(0.0, 0.0, -9.81))


def wavetank_bench(mps, steps, warmup, nx=None, log=print):
    """Semi-implicit wave tank with the moving piston and the learned guess:
    warm-up steps record (features, solution) pairs from the previous-pressure
    solves, the network is trained on them, then `steps` timed steps run with
    the learned guess.  The FP32 run is checked against the FP64 device path
    over a short window from the same state."""
    w = dict(WAVETANK)
    if nx and nx != w["inner"][0]:
        w["inner"] = (nx,) + w["inner"][1:]
        w["name"] = w["name"].replace("4.0x0.3x0.6 m", f"{nx * w['dp']:.3g}x0.3x0.6 m").replace(
            "(1000x75x75 fluid)", f"({nx}x75x75 fluid)")
    nx, ny, nz = w["inner"]
    nb2, nb1, pos, pist = mps.wavetank(nx, ny, nz, w["depth"], w["travel"], 2, 2, w["dp"], 0.0, 0)
    N = len(pos)
    nfl = N - nb2 - nb1
    log(f"[bench] wave tank N={N} fluid={nfl}")
    alpha = 1.0 / (1000.0 * 20.0 ** 2)
    s = mps.Solver(wavetank_params(mps, w["dp"]), nb2, nb1, pos)
    s.set_motion(pist, 0, w["amp"], w["freq"])
    s.mlp_init(0)
    s.mlp_set_record(True)
    t0 = time.perf_counter()
    it_w, ms_w = s.step_semi(warmup, alpha=alpha, x0_kind=1, tol=1e-8)
    s.mlp_set_record(False)
    l0, l1 = s.mlp_train(1000, batch=4096, lr=1e-3, seed=0)
    log(f"[bench] {warmup} recorded steps ({it_w} CG iterations) and training in "
        f"{time.perf_counter() - t0:.1f} s: loss {l0:.3e} -> {l1:.3e}")
    t1 = time.perf_counter()
    it, ms = s.step_semi(steps, alpha=alpha, x0_kind=3, tol=1e-8)
    el = time.perf_counter() - t1
    log(f"[bench] {steps} learned-guess steps in {el:.2f} s, {it} CG iterations")
    # plain CG on the system of the next step, for the same-matrix comparison
    s.stage_dt(); s.stage_sort(); s.stage_visc(); s.stage_collision(); s.stage_pnd()
    nnz = s.ppe_assemble(alpha)
    cmp = {}
    for name, kind in (("plain_zero", 0), ("prev_step", 1), ("learned", 3)):
        best = None
        for rep in range(3):
            t2 = time.perf_counter()
            if kind == 3:
                s.mlp_infer()
            i2, rel = s.ppe_solve(mixed=True, x0_kind=kind, tol=1e-8)
            ms2 = 1e3 * (time.perf_counter() - t2)
            if best is None or ms2 < best["ms"]:
                best = {"ms": round(ms2, 3), "iters": i2, "relres": rel}
        cmp[name] = best
    st32 = s.state(by_id=True)
    sc = s.scalars()
    s.close()
    # FP64 device path over a short window from the initial state: error of p and u
    k = 5
    s32 = mps.Solver(wavetank_params(mps, w["dp"], 0), nb2, nb1, pos)
    s64 = mps.Solver(wavetank_params(mps, w["dp"], 1), nb2, nb1, pos)
    for q in (s32, s64):
        q.set_motion(pist, 0, w["amp"], w["freq"])
        q.step_semi(k, alpha=alpha, x0_kind=1, tol=1e-8)
    a, b = s32.state(by_id=True), s64.state(by_id=True)
    fl = a["type"] == 0
    du = np.linalg.norm(a["vel"][fl] - b["vel"][fl], axis=1)
    err = {"steps": k,
           "p": float(np.max(np.abs(a["p"][fl] - b["p"][fl])) / max(np.max(b["p"][fl]), 1e-300)),
           "p_max_Pa": float(np.max(b["p"][fl])),
           "u": float(np.max(du) / max(np.max(np.linalg.norm(b["vel"][fl], axis=1)), 1e-300))}
    s32.close()
    s64.close()
    return {"metric": "wave tank semi-implicit particle-updates/s; PPE ms/solve learned vs plain",
            "value": nfl * steps / el, "unit": "particle-updates/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": 1e3 * el / steps, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": w["name"], "n_total": N, "n_fluid": nfl, "dt_last": sc["dt"],
                       "time": sc["time"], "re_lap": 3.1, "alpha": alpha},
            "ppe": {"rows": N, "nnz": int(nnz), "solves_last_system": cmp,
                    "timed_steps": {"cg_iterations": it, "solve_ms_total": ms,
                                    "ms_per_solve": ms / max(steps, 1), "iters_per_solve": it / max(steps, 1)},
                    "learned_over_plain_time": cmp["learned"]["ms"] / cmp["plain_zero"]["ms"]},
            "fp64_error": err}


def copy_calib(mib, reps=3):
    """A plain device-to-device copy of known size: the reference point that
    scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": src.numel() * 4, "reps": reps}), flush=True)


def traffic_from_profile(dom, slug):
    """HBM bytes per launch of kernel `dom` on workload `slug` from the committed
    FETCH_SIZE / WRITE_SIZE summary (a line 'bytes_per_launch: B'), or None."""
    tf = os.path.join(HERE, "profiles", f"traffic_{dom}_{slug}.txt")
    if not os.path.exists(tf):
        return None
    for line in open(tf):
        if line.startswith("bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


# VALU issue peak: 157.3 TFLOP/s FP32 vector / (64 lanes x 2 flop per FMA).  That
# figure needs packed FP32 (two FMAs per lane per instruction); the kernels here are
# built without packed FP32, so their ceiling is one wave64 instruction per SIMD
# every 4 cycles: 256 CUs x 4 SIMDs x 2.4 GHz / 4.
VALU_PEAK_INSTR_S = 157.3e12 / 128
VALU_UNPACKED_INSTR_S = 256 * 4 * 2.4e9 / 4


def valu_from_profile(dom, slug):
    """VALU wave64 instructions per launch of kernel `dom` (SQ_INSTS_VALU) from
    the committed counter summary, or None."""
    tf = os.path.join(HERE, "profiles", f"valu_step_{slug}.txt")
    if not os.path.exists(tf):
        return None
    key = f"valu_instr_per_launch_{dom}:"
    for line in open(tf):
        if line.startswith(key):
            return float(line.split(":", 1)[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", type=int, default=1, choices=sorted(WORKLOADS) + [4],
                    help="1 (default): explicit step at 1.05M fluid; 0: 16k fluid; 4: semi-implicit wave tank")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-ppe", action="store_true")
    ap.add_argument("--wavetank-nx", type=int, default=1000,
                    help="config 4: tank length in lattice sites (1000 = the 4.0 m tank)")
    ap.add_argument("--ppe-only", action="store_true",
                    help="only the pressure Poisson part (one JSON line with the ppe object)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter scaling) and exit")
    a = ap.parse_args()
    if a.copy_calib:
        copy_calib(a.copy_calib)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the explicit "
              "step would shard by spatial slabs, see the design notes)", file=sys.stderr)
        sys.exit(2)

    import mps_solvers_amd as mps

    log = lambda m: print(m, file=sys.stderr, flush=True)
    if a.ppe_only:
        print(json.dumps({"ppe": ppe_bench(mps, log=log)}), flush=True)
        return
    if a.config == 4:
        print(json.dumps(wavetank_bench(mps, a.steps if a.steps != 200 else 100, a.warmup, a.wavetank_nx,
                                        log=log)), flush=True)
        return
    w = WORKLOADS[a.config]
    nb2, nb1, pos = mps.dambreak(w["fluid"], w["inner"], 2, 2, w["dp"], 0.01, 0)
    N = len(pos)
    nfl = N - nb2 - nb1
    print(f"[bench] scene N={N} fluid={nfl} walls+dummies={nb2 + nb1}", file=sys.stderr, flush=True)
    s = mps.Solver(params_for(mps, w), nb2, nb1, pos)
    s.step(a.warmup)
    t0 = time.perf_counter()
    s.step(a.steps)
    el = time.perf_counter() - t0
    ms_step = 1e3 * el / max(a.steps, 1)
    print(f"[bench] {a.steps} steps in {el:.3f} s ({ms_step:.3f} ms/step)", file=sys.stderr, flush=True)

    # per-stage device time from HIP events on the library's stream
    nt = min(20, max(a.steps, 1))
    s.enable_timing(True)
    s.stage_times(reset=True)
    s.step(nt)
    st = s.stage_times(reset=True)
    s.enable_timing(False)
    per = {k: v / nt for k, v in st.items()}
    dom = max(("visc", "collision", "pnd_pres", "gradp", "sort"), key=lambda k: per[k])
    ach = ALG_BYTES[dom] * N / (per[dom] * 1e-3) / 1e9
    traffic = traffic_from_profile(dom, w["slug"])
    vi = valu_from_profile(dom, w["slug"])
    valu = None
    if vi is not None:
        va = vi / (per[dom] * 1e-3)
        valu = {"instr_per_launch": vi, "achieved": va, "peak": VALU_PEAK_INSTR_S, "unit": "wave64 instr/s",
                "frac": va / VALU_PEAK_INSTR_S, "peak_unpacked": VALU_UNPACKED_INSTR_S,
                "frac_unpacked": va / VALU_UNPACKED_INSTR_S}
    sc = s.scalars()
    s.close()
    # FP32 trajectory vs the FP64 device path (itself at ~1e-13 of the FP64 oracle in
    # the parity tests) after the timed number of steps from the same initial state
    # (the FP32-rounded lattice, so that both start from identical values and the
    # same n0): max |p32 - p64| / max |p64| over fluid, the same for |u| as a
    # vector, and the largest position difference
    k = a.steps
    pos32 = pos.astype(np.float32).astype(np.float64)
    a32 = mps.Solver(params_for(mps, w, 0), nb2, nb1, pos32)
    a64 = mps.Solver(params_for(mps, w, 1), nb2, nb1, pos32)
    a32.step(k)
    a64.step(k)
    g32, g64 = a32.state(by_id=True), a64.state(by_id=True)
    # per-operator error at full size: before each operator the FP64 path takes
    # the FP32 path's state (exact FP32 values, mapped by particle id into its own
    # order), both run the operator, and the outputs are compared -- the
    # parity-test measure at 1.38M particles (FP64 path ~1e-13 from the oracle)
    def sync():
        gs = a32.state(by_id=True)
        ids = a64.state()["id"]
        a64.set_state(pos=gs["pos"][ids], vel=gs["vel"][ids], press=gs["p"][ids])
        return gs

    def rel(d, ref):
        return float(np.max(np.linalg.norm(d, axis=-1)) / max(np.max(np.linalg.norm(ref, axis=-1)), 1e-300))

    sync()
    dt = float(w["dt_max"])
    for sv in (a32, a64):
        sv.set_dt(dt)
        sv.stage_sort()
    operators = {}
    for name, fn in (("visc", "stage_visc"), ("collision", "stage_collision"), ("pnd_pres", "stage_pnd_pres"),
                     ("gradp", "stage_gradp")):
        g0 = sync()
        getattr(a32, fn)()
        getattr(a64, fn)()
        h32, h64 = a32.state(by_id=True), a64.state(by_id=True)
        fl0 = g0["type"] == 0
        if name == "pnd_pres":
            operators["n"] = rel((h32["n"] - h64["n"])[fl0, None], h64["n"][fl0, None])
            operators["p"] = rel((h32["p"] - h64["p"])[fl0, None], h64["p"][fl0, None])
        else:
            dv32, dv64 = h32["vel"][fl0] - g0["vel"][fl0], h64["vel"][fl0] - g0["vel"][fl0]
            operators[name + "_du"] = rel(dv32 - dv64, dv64)
    a32.close()
    a64.close()
    flm = g32["type"] == 0
    du = np.linalg.norm(g32["vel"][flm] - g64["vel"][flm], axis=1)
    fp64_error = {"steps": k,
                  "p": float(np.max(np.abs(g32["p"][flm] - g64["p"][flm])) / max(np.max(g64["p"][flm]), 1e-300)),
                  "p_max_Pa": float(np.max(g64["p"][flm])),
                  "u": float(np.max(du) / max(np.max(np.linalg.norm(g64["vel"][flm], axis=1)), 1e-300)),
                  "x_max_abs_m": float(np.max(np.abs(g32["pos"] - g64["pos"]))),
                  "operators_same_state": operators}
    out = {
        "metric": "particle-updates/s at 1M particles; PPE ms/solve to 1e-8 residual, learned vs plain",
        "value": nfl * a.steps / el,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": w["name"], "n_total": N, "n_fluid": nfl, "l0": w["dp"],
                   "dt": sc["dt"], "re_n": 2.1, "re_lap": 4.0,
                   "all_particle_updates_per_s": N * a.steps / el,
                   "step_roofline_frac": STEP_BYTES * N / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                   "stage_ms": {k: round(v, 4) for k, v in per.items()}},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                     # the sweeps issue far more VALU work than their bytes need: the VALU
                     # issue rate is what bounds them (SQ_INSTS_VALU from the profile)
                     "valu": valu},
        "fp64_error": fp64_error,
    }
    if not a.no_ppe:
        print("[bench] pressure Poisson: plain vs previous-step vs learned guess", file=sys.stderr, flush=True)
        out["ppe"] = ppe_bench(mps, log=lambda m: print(m, file=sys.stderr, flush=True))
    if not a.no_cpu:
        print("[bench] timing the CPU oracle on a bounded sample", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(w, nb2, nb1, pos, a.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
