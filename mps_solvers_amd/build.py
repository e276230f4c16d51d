"""Build the gfx950 HIP library ``mps_solvers_amd/libmps_amd.so`` in-tree.

    python -m mps_solvers_amd.build          # compile every source, link
    python -m mps_solvers_amd.build --check  # print the hipcc command lines only

All device code is compiled for gfx950 only, without FMA contraction (the
neighbour cutoff and cell-hash decisions must round like the reference), with
correctly rounded FP32 division/sqrt, and with packed-FP32 instructions
disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libmps_amd.so")
SOURCES = ["mps_sort.hip", "mps_sweeps.hip", "mps_sweeps_wave.hip", "mps_ppe.hip", "mps_mlp.hip", "mps_motion.hip",
           "mps_capi.hip", "mps_ppe_capi.hip", "mps_host.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-fhip-fp32-correctly-rounded-divide-sqrt",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument", "-Wno-unused-result"]


def _sources():
    return [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = _sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # command-line driver linked against the library (rpath: its own folder)
    exe = os.path.join(HERE, "mps_cal")
    cmd = ["g++", "-O2", "-std=c++17", "-o", exe + ".tmp", os.path.join(CSRC, "mps_cal_main.cpp"),
           "-L" + HERE, "-lmps_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(exe + ".tmp", exe)
    if verbose:
        print(f"built {LIB} and {exe}")
    return LIB


if __name__ == "__main__":
    if "--check" in sys.argv:
        for s in _sources():
            print(" ".join([HIPCC] + FLAGS + ["-c", s]))
    else:
        build(verbose=True)
