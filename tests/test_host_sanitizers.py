"""The library's host C under AddressSanitizer and UndefinedBehaviorSanitizer.

lbm-bench_amd/csrc/lbm_host.c (geometryCreate's parsing and sphere
parameters for every type, the reference's -V initialisers, statistics and
verify; the device entry points it calls have host stand-ins)
is compiled by gcc with tests/host_sanitize_driver.c, which drives it
through a mock KernelDataType over a host array.  Ten lattices, including
a one-cell-wide axis and a sphere packing with no fluid left.  A sanitizer
report aborts the program, so a clean exit is the test.  Host code only;
nothing here touches a GPU.
"""
import os
import shutil
import subprocess

import pytest
from conftest import ROOT

HOST_C = os.path.join(ROOT, "lbm-bench_amd", "csrc", "lbm_host.c")
DRIVER = os.path.join(ROOT, "tests", "host_sanitize_driver.c")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is not installed")
def test_host_code_runs_clean_under_sanitizers(tmp_path):
    exe = str(tmp_path / "host_sanitize")
    sanitize = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                "-fno-omit-frame-pointer"]
    cmd = ["gcc", "-O1", "-g", "-std=c99", "-D_GNU_SOURCE", "-Wall", "-Wextra",
           "-Wno-unused-parameter", *sanitize, "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "lbm-bench_amd", "csrc"), HOST_C, DRIVER, "-o", exe, "-lm"]
    build = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    run = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert run.returncode == 0, run.stdout[-2000:] + run.stderr[-4000:]
    assert "runtime error" not in run.stderr
    assert run.stdout.count(": ok") == 10
