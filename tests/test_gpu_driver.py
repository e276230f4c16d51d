"""The C++ host surface (include/mps_solvers.hpp) through the mps_cal driver:
input.dat / GEO.dat in, out/%04d.out in the reference's writeCase format out."""
import os
import subprocess

import numpy as np
import pytest

import mps_solvers_amd as mps
from scenes import CFG0, small_dam

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(mps.__file__), "mps_cal")


@pytest.mark.parametrize("mode", [[], ["--semi"], ["--fp64"], ["--vtk"]])
def test_driver_runs_reference_files(tmp_path, mode):
    nb2, nb1, pos = small_dam(2)
    kw = {k: v for k, v in CFG0.items() if k != "g"}
    kw.update(tt=0.0012, tout=0.0005)
    if "--semi" in mode:
        kw.update(re_lap_mult=3.1, dt_max=2e-4, tt=0.0009, tout=0.0004)
    mps.write_case(tmp_path / "input.dat", mps.params(**kw))
    mps.geo_write(tmp_path / "GEO.dat", nb2, nb1, pos)
    r = subprocess.run([EXE, str(tmp_path)] + mode, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    outs = sorted(f for f in os.listdir(tmp_path / "out") if f.endswith(".out"))
    assert outs[0] == "0000.out" and len(outs) >= 2
    vtks = sorted(f for f in os.listdir(tmp_path / "out") if f.endswith(".vtk"))
    assert len(vtks) == (len(outs) if "--vtk" in mode else 0)
    if vtks:
        head = open(tmp_path / "out" / vtks[0]).read().split("\n")[:5]
        assert head[0] == "# vtk DataFile Version 3.0" and head[4] == f"POINTS {len(pos)} float"
    tok = open(tmp_path / "out" / outs[-1]).read().split()
    assert len(tok) == 9 * len(pos)
    rec = np.array(tok, float).reshape(-1, 9)
    assert sorted(rec[:, 0].astype(int).tolist()) == list(range(len(pos)))
    assert set(rec[:, 1].astype(int).tolist()) <= {0, 1, 2}
    # the run log (the reference's out/LOG.txt): inputs, geometry, one line per output, the end
    log = open(tmp_path / "out" / "LOG.txt").read().splitlines()
    assert log[0].startswith("input.dat: dim 3") and log[1].startswith(f"GEO.dat: {len(pos)} particles")
    assert sum(line.startswith("output ") for line in log) == len(outs)
    assert log[-1].startswith("done: ")
    if "--semi" in mode:
        assert "CG" in log[2]
