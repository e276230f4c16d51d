"""The benchmark's output line, on the small dam break (16,000 fluid particles).

bench.py prints one JSON line whose fields are consumed by other programs: the
metric and its unit, the whole-job value with the step time it came from, the
roofline object of the dominant kernel (HBM bytes against 8 TB/s, plus the
VALU issue figures), and the CPU baseline.  This runs it as a separate process,
as it is run in practice, with a short timed region and a one-second CPU
sample, and checks those fields and their arithmetic.
"""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    assert len(lines) == 1, r.stdout[-2000:]
    return json.loads(lines[0])


def test_bench_line_fields_and_arithmetic():
    d = _run("--config", "0", "--steps", "5", "--warmup", "2", "--no-ppe", "--cpu-budget", "1")
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
              "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 5 and d["warmup"] == 2
    assert d["unit"] == "particle-updates/s" and d["higher_is_better"] is True
    assert d["vs_baseline"] is None and d["data"] == "synthetic"
    cfg = d["config"]
    assert "workload" in cfg and cfg["n_fluid"] == 16000
    # value = fluid particles x steps / wall seconds of the timed steps
    assert d["value"] == pytest.approx(cfg["n_fluid"] / (d["ms_per_step"] * 1e-3), rel=1e-6)
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert rf["frac"] == pytest.approx(rf["achieved"] / rf["peak"], rel=1e-9)
    assert 0.0 < rf["frac"] < 1.0
    cb = d["cpu_baseline"]
    assert cb["unit"] == "particle-updates/s" and cb["kind"] in ("port", "reference")
    assert cb["value"] > 0 and cb["cores"] >= 1 and cb["sample"]


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
