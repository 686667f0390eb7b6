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
    # the dominant kernel: 304 algorithmic bytes per fluid-cell update over
    # its own mean launch time (hipEvents on the library's stream)
    nf = out["config"]["fluid_cells"]
    assert roof["bytes_per_launch"] == 304 * nf
    assert abs(roof["achieved"] - roof["bytes_per_launch"] / roof["kernel_avg_ms"] / 1e6) \
        < 1e-3 * roof["achieved"]
    assert roof["kernel_avg_ms"] == max(k["avg_ms"] for k in out["kernels"].values())
    # the step pair over the whole timed run, at the line's value
    assert abs(roof["step_pair_achieved"] - out["value"] * 304 / 1e3) < 0.01 * roof["achieved"]
    assert roof["achievable_gbs"] == max(r["gbs"] for r in out["copy_sweep_gbs"])
    assert abs(roof["frac_of_achievable"] - roof["achieved"] / roof["achievable_gbs"]) < 1e-3
    assert out["adjoint_mlups"] is None and out["adjoint_sweep_mlups"] is None
    assert out["config"]["fluid_cells"] > 0
    assert out["fp16"]["mlups"] > 0 and out["fp16"]["velocity_l2_rel_err_vs_fp64"] < 1e-2


@pytest.mark.gpu
def test_profiling_options():
    """--geometry replaces the sphere packing, --kernel picks the variant and
    --no-sweep leaves out the copy sweep and the pattern: the achievable
    line is then null rather than invented"""
    run = subprocess.run([sys.executable, BENCH, "--size", "64", "--geometry", "channel",
                          "--kernel", "pull-aos", "--steps", "4", "--warmup", "2", "--no-adjoint",
                          "--no-cpu", "--no-fp16", "--no-sweep"],
                         capture_output=True, text=True, timeout=600)
    assert run.returncode == 0, run.stderr[-2000:]
    out = json.loads([ln for ln in run.stdout.splitlines() if ln.startswith('{"metric"')][0])
    assert "'channel' geometry" in out["config"]["workload"]
    assert out["config"]["kernel"] == "pull-aos"
    assert out["config"]["fluid_cells"] == 64 * 62 * 64
    roof = out["roofline"]
    assert roof["achievable_gbs"] is None and roof["frac_of_achievable"] is None
    assert out["copy_sweep_gbs"] is None and out["pattern_bandwidth_gbs"] is None
    assert roof["kernel_avg_ms"] > 0 and out["fp16"] is None
