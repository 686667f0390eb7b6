"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

oracle/lbm_oracle.c is compiled with tests/oracle_sanitize_driver.c by gcc
(serial: the OpenMP pragmas are left out) and every oracle function runs on
small lattices of six geometries, all seven variants, 1-3 steps, plus the
Brinkman adjoint.  A sanitizer report aborts the program, so a clean exit is
the test.  Host code only; nothing here touches a GPU.
"""
import os
import shutil
import subprocess

import pytest
from conftest import ROOT

ORACLE = os.path.join(ROOT, "oracle")
DRIVER = os.path.join(ROOT, "tests", "oracle_sanitize_driver.c")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is not installed")
def test_oracle_runs_clean_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_sanitize")
    sanitize = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                "-fno-omit-frame-pointer"]
    cmd = ["gcc", "-O1", "-g", "-std=c99", "-D_GNU_SOURCE", "-Wall", "-Wextra",
           "-Wno-unused-parameter", "-Wno-unknown-pragmas", *sanitize, "-I", ORACLE,
           os.path.join(ORACLE, "lbm_oracle.c"), DRIVER, "-o", exe, "-lm"]
    build = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    run = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "runtime error" not in run.stderr
    assert run.stdout.count(": ok") == 7
