"""The benchmark driver binary on the GPU: LBM-Bench's -V verification run
(src/test.sh:61-103) for every kernel in both builds, and a benchmark run
whose output lines keep the reference's format (src/main.c:154-171)."""
import json
import math
import os
import re
import subprocess

import pytest
from conftest import GOLDEN

import lbmhip

pytestmark = pytest.mark.gpu


def _run(precision, *args):
    return subprocess.run([lbmhip.driver_path(precision), *args], capture_output=True, text=True,
                          timeout=300)


@pytest.mark.parametrize("precision", ["dp", "sp"])
def test_verification_runs(precision):
    with open(os.path.join(GOLDEN, "verification.json")) as f:
        ver = json.load(f)[precision]
    spread = max(ver.values()) - min(ver.values())
    for kernel in lbmhip.KERNELS:
        r = _run(precision, "-V", "-k", kernel)
        assert r.returncode == 0, r.stdout[-2000:]
        assert "VERIFICATION SUCCEEDED" in r.stdout
        norm = float(re.search(r"deviation from analytical solution: (\S+)", r.stdout).group(1))
        tol = 1e-6 * ver[kernel] if precision == "dp" else 2 * spread
        assert abs(norm - ver[kernel]) <= tol, (kernel, norm, ver[kernel])


def test_benchmark_output_format():
    r = _run("dp", "-d", "64x64x64", "-g", "channel", "-z", "-k", "aa-soa", "-i", "50", "-o",
             "1.25", "-m", "0.1875", "-B", "-f", "1e-6", "-I", "random", "-s", "3", "-D", "0")
    assert r.returncode == 0, r.stdout[-2000:]
    m = re.search(r"^P:\s+([0-9.]+) MFLUP/s\s+d: ([0-9.]+) s\s+iter: 50\s+fnodes: ([0-9.]+) x1e6"
                  r"\s+geo: channel\s+kernel: aa-soa$", r.stdout, flags=re.M)
    assert m and float(m.group(1)) > 0
    assert abs(float(m.group(3)) - 64 * 62 * 64 / 1e6) < 1e-6
    drift = float(re.search(r"mass drift:\s+(\S+)", r.stdout).group(1))
    assert abs(drift) < 1e-13
    for line in ("# Evaluation Stats", "#   runtime:", "#   MEM bandwidth:", "#   performance:"):
        assert line in r.stdout


REF_DRIVER = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "lbmbench-ref-driver-hip-{}")


@pytest.mark.parametrize("precision", ["dp", "sp"])
def test_reference_driver_runs_on_the_gpu(precision):
    """LBM-Bench's own main.c and cli.c, unmodified, compiled against
    include/lbm.h and linked to the GPU library: its -V run (the body force
    of its -DVERIFICATION build switched on by LBM_BODY_FORCE) reproduces
    the norm the reference prints"""
    exe = os.path.abspath(REF_DRIVER.format(precision))
    if not os.path.exists(exe):
        pytest.skip("reference driver not built (needs LBM-Bench's tree at build time)")
    with open(os.path.join(GOLDEN, "verification.json")) as f:
        ver = json.load(f)[precision]
    spread = max(ver.values()) - min(ver.values())
    env = dict(os.environ, LBM_BODY_FORCE="1")
    for kernel in lbmhip.KERNELS:
        r = subprocess.run([exe, "-V", "-k", kernel], capture_output=True, text=True, env=env,
                           timeout=300)
        assert r.returncode == 0, r.stdout[-2000:]
        assert "on the GPU" in r.stdout
        norm = float(re.search(r"deviation from analytical solution: (\S+)", r.stdout).group(1))
        tol = 1e-6 * ver[kernel] if precision == "dp" else 2 * spread
        assert abs(norm - ver[kernel]) <= tol, (kernel, norm, ver[kernel])


@pytest.mark.parametrize("planes,element_bytes,in_place", [(19, 8, True), (19, 8, False),
                                                            (19, 4, True), (19, 2, True),
                                                            (1, 8, False)])
def test_pattern_bandwidth_measures(planes, element_bytes, in_place):
    """the ceiling kernel runs and reports a rate between 1 GB/s and the 8 TB/s
    the part is quoted at (small lattice: launch-bound, so no tighter bound)"""
    lib = lbmhip.Library("dp")
    gbs = lib.pattern_bandwidth(planes, element_bytes, in_place, 1 << 22, 3)
    assert 1.0 < gbs < 8000.0


def test_unstable_run_fails():
    """omega = 20 amplifies the noise of a random start by about 19 per
    step, so 400 steps overflow; the driver reports the non-finite mass and
    exits with EXIT_FAILURE instead of printing NaN and succeeding"""
    r = _run("dp", "-d", "32x32x32", "-g", "channel", "-z", "-k", "aa-soa", "-i", "400", "-o",
             "20", "-I", "random", "-s", "1")
    assert r.returncode == 1, r.stdout[-2000:] + r.stderr[-2000:]
    assert "mass is not finite" in r.stderr


def test_device_allocation_failure_exits():
    """a two-lattice FP64 variant sized from the device's memory so that its
    two arrays (2 x 152 B per cell) exceed it by a tenth: Init compares the
    bytes it needs with hipMemGetInfo before its first hipMalloc and exits
    with EXIT_FAILURE and a message, as the reference's allocate() does
    (src/allocate.c:17-31).  No device memory is allocated and nothing is
    launched; the host still builds the cell flags (4 B per cell)."""
    _, _, mem = lbmhip.Library("dp").device_info()
    n = int(math.ceil((1.1 * mem / 304.0) ** (1.0 / 3.0)))
    r = _run("dp", "-d", f"{n}x{n}x{n}", "-g", "channel", "-z", "-k", "push-soa", "-i", "2")
    assert r.returncode == 1, r.stdout[-2000:] + r.stderr[-2000:]
    assert "insufficient device memory" in r.stderr and "hipMalloc" not in r.stderr
