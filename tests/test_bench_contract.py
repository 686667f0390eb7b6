"""bench.py keeps its output contract: one JSON line with the metric, the
whole-job value, the roofline object of the dominant kernel and the CPU
baseline beside it; more than one GPU is refused with a message."""
import json
import os
import subprocess
import sys

import pytest
from conftest import ROOT

BENCH = os.path.join(ROOT, "bench.py")


def test_more_than_one_gpu_is_refused():
    run = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True,
                         timeout=120)
    assert run.returncode == 2
    assert "one GPU" in run.stderr


@pytest.mark.gpu
def test_bench_line_has_the_contract_fields():
    run = subprocess.run([sys.executable, BENCH, "--size", "64", "--spheres", "20", "--steps", "8",
                          "--warmup", "2", "--no-adjoint", "--no-cpu"],
                         capture_output=True, text=True, timeout=600)
    assert run.returncode == 0, run.stderr[-2000:]
    lines = [ln for ln in run.stdout.splitlines() if ln.startswith('{"metric"')]
    assert len(lines) == 1
    out = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
                "roofline", "cpu_baseline"):
        assert key in out, key
    assert out["n_gpus"] == 1 and out["steps"] == 8 and out["warmup"] == 2
    assert out["value"] > 0 and out["ms_per_step"] > 0 and out["higher_is_better"] is True
    assert out["dtype"] == "f64"
    roof = out["roofline"]
    assert roof["bound"] == "hbm" and roof["unit"] == "GB/s" and roof["peak"] == 8000.0
    assert abs(roof["frac"] - roof["achieved"] / roof["peak"]) < 1e-3
    # 304 algorithmic bytes per fluid-cell update, at the measured rate
    assert abs(roof["achieved"] - out["value"] * 304 / 1e3) < 0.01 * roof["achieved"]
    assert out["config"]["fluid_cells"] > 0
    assert out["fp16"]["mlups"] > 0 and out["fp16"]["velocity_l2_rel_err_vs_fp64"] < 1e-2
