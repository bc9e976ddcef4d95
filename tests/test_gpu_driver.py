"""The headless driver nbody_run (the reference's main.cpp frame loop without the window):
its steps equal the Python path's bit for bit, a run resumed from its checkpoint equals an
uninterrupted run, the elastic pool recovers a too-small pool, and its outputs (JSON lines,
checkpoint, vertex export) are complete."""
import json
import os
import subprocess

import numpy as np
import pytest

from nbodysimulation_amd import _lib as L
from nbodysimulation_amd import ic
from nbodysimulation_amd.engine import NBodySim, make_config

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "nbodysimulation_amd", "nbody_run")


def run(*args):
    assert os.path.exists(EXE), "nbody_run is not built (make -C nbodysimulation_amd)"
    r = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]


def state(path):
    with NBodySim.from_checkpoint(path) as s:
        return s.download(), s.step_count()


def test_driver_equals_python_path(gpu, tmp_path):
    n = 3000
    out = run("--bodies", n, "--algorithm", "bh", "--rules", "standard", "--ic", "plummer", "--seed", 3,
              "--steps", 5, "--checkpoint", tmp_path / "a.nbck")
    assert out[-1]["steps"] == 5 and out[-1]["step_count"] == 5
    (p, v, a), k = state(tmp_path / "a.nbck")
    assert k == 5
    with NBodySim(make_config(n, L.BARNES_HUT, rules=L.RULES_STANDARD, law=L.LAW_PLUMMER)) as s:
        s.upload(*ic.plummer(n, 3))
        s.step(5)
        ref = s.download()
    for x, y in zip((p, v, a), ref):
        assert x.tobytes() == y.tobytes()


def test_driver_resume_equals_uninterrupted(gpu, tmp_path):
    common = ["--bodies", 2000, "--algorithm", "bh", "--rules", "reference", "--ic", "normal", "--seed", 4]
    run(*common, "--steps", 10, "--checkpoint", tmp_path / "full.nbck")
    run(*common, "--steps", 4, "--checkpoint", tmp_path / "part.nbck")
    out = run("--resume", tmp_path / "part.nbck", "--steps", 6, "--report-every", 3, "--checkpoint",
              tmp_path / "resumed.nbck")
    assert [r["step"] for r in out if "step" in r] == [7, 10]
    (f, kf), (r, kr) = state(tmp_path / "full.nbck"), state(tmp_path / "resumed.nbck")
    assert kf == kr == 10
    for x, y in zip(f, r):
        assert x.tobytes() == y.tobytes()


def test_driver_grows_a_small_pool(gpu, tmp_path):
    common = ["--bodies", 4000, "--algorithm", "bh", "--ic", "plummer", "--seed", 5, "--steps", 4]
    out = run(*common, "--cells", 16, "--grow", "--checkpoint", tmp_path / "g.nbck")
    assert out[-1]["pool_grown"] > 0 and out[-1]["cell_capacity"] >= out[-1]["cells_used"] > 16
    run(*common, "--checkpoint", tmp_path / "d.nbck")
    for x, y in zip(state(tmp_path / "g.nbck")[0], state(tmp_path / "d.nbck")[0]):
        assert x.tobytes() == y.tobytes()


def test_driver_outputs(gpu, tmp_path):
    n = 1500
    out = run("--bodies", n, "--algorithm", "bf", "--law", "unit", "--ic", "reference", "--pos", "sphere",
              "--vel", "tangent", "--massive", 2, "--steps", 6, "--energy", "--timing", "--vertices",
              tmp_path / "v.bin")
    summary = out[-1]
    assert set(summary["stage_ms_per_step"]) >= {"brute", "update"} and summary["stage_ms_per_step"]["brute"] > 0
    assert "energy_start" in out[0] and "energy_end" in out[1]
    assert abs(summary["relative_energy_drift"]) < 1e-2
    v = np.fromfile(tmp_path / "v.bin", np.float32).reshape(-1, 4)
    assert v.shape == (n, 4) and (v[:, 3] == 1.0).all() and np.isfinite(v).all()
    out = run("--bodies", 3000, "--algorithm", "bh", "--method", "leapfrog", "--collision", "--ic", "ball",
              "--steps", 3, "--timing")
    assert out[-1]["step_count"] == 3 and out[-1]["stage_ms_per_step"]["force"] > 0
