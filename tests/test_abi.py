"""The C-ABI library loads and exports every symbol include/*.h declares,
and the driver keeps the reference's command-line contract (src/cli.c,
src/main.c:93-104) for the paths that need no GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest
from conftest import ROOT

import lbmhip

HEADERS = [os.path.join(ROOT, "include", h) for h in ("lbm.h", "lbm_hip.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"static inline[^{]*\{[^}]*\}", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name in ("if", "for", "while", "return", "sizeof", "void"):
                continue
            names.add(name)
    return names


def declared_variables():
    names = set()
    for h in HEADERS:
        for m in re.finditer(r"^extern [^;(\n\"]*?\b([A-Za-z_]\w*)\s*(\[[^\]]*\])?;",
                             open(h).read(), flags=re.M):
            names.add(m.group(1))
    return names


@pytest.mark.parametrize("precision", ["dp", "sp"])
def test_library_exports_every_declared_symbol(precision):
    lib = ctypes.CDLL(lbmhip.library_path(precision))
    funcs = declared_functions()
    assert {"kernelInitAaSoA", "kernelPullSoA", "lbmGetNodes", "geometryCreate"} <= funcs
    missing = [n for n in sorted(funcs | declared_variables()) if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("planes,element_bytes,cells,reps",
                         [(3, 8, 1024, 1), (19, 3, 1024, 1), (19, 8, 0, 1), (19, 8, 1024, 0)])
def test_pattern_bandwidth_rejects_unsupported_shapes(planes, element_bytes, cells, reps):
    """argument checks run before any device call"""
    lib = lbmhip.Library("dp")
    with pytest.raises(ValueError):
        lib.pattern_bandwidth(planes, element_bytes, True, cells, reps)


@pytest.mark.parametrize("dims", [(0, 4, 4), (4, -1, 4), (4, 4, 0)])
def test_porous_field_rejects_empty_dimensions(dims):
    """the dimension check runs before any device call and leaves s alone"""
    lib = lbmhip.Library("dp")
    s = np.full(8, 7.0)
    lib.lib.lbmPorousField(lbmhip.I3(*dims), 1, 2.0, 4.0, -2.0,
                           s.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert (s == 7.0).all()


def test_registry_names_and_order():
    lib = lbmhip.Library("dp")
    assert lib.kernel_names() == list(lbmhip.KERNELS)


def test_struct_layout_matches_header():
    """ctypes mirror of KernelDataType / CaseDataType has the C sizes"""
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "lbm.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(KernelDataType), sizeof(CaseDataType),
         sizeof(LatticeDescType), offsetof(KernelDataType, Dev), offsetof(CaseDataType, RhoRef));
  return 0;
}
"""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.check_call(["gcc", "-I" + os.path.join(ROOT, "include"), c, "-o", exe])
        vals = [int(v) for v in subprocess.check_output([exe]).split()]
    lib = lbmhip.Library("dp")
    assert vals == [ctypes.sizeof(lib.KernelData), ctypes.sizeof(lib.CaseData),
                    ctypes.sizeof(lib.LatticeDesc), lib.KernelData.Dev.offset,
                    lib.CaseData.RhoRef.offset]


def _driver(*args):
    return subprocess.run([lbmhip.driver_path("dp"), *args], capture_output=True, text=True,
                          timeout=60)


def test_driver_lists_kernels():
    r = _driver("-l")
    assert r.returncode == 0
    listed = [l.strip() for l in r.stdout.split("Available kernels:")[1].splitlines() if l.strip()]
    assert listed == list(lbmhip.KERNELS)


def test_driver_help_and_errors():
    assert _driver("-h").returncode == 0
    r = _driver("-t", "1")
    assert r.returncode != 0 and "unknown option" in r.stdout
    r = _driver("-d", "0x4x4")
    assert r.returncode != 0 and "X dimension must be > 0" in r.stdout
    r = _driver("-d", "4x4")
    assert r.returncode != 0 and "Z dimension is missing" in r.stdout
    r = _driver("-i", "0")
    assert r.returncode != 0 and "iterations must be > 0" in r.stdout
    r = _driver("-g", "nosuch")
    assert r.returncode != 0 and "unknown geometry type" in r.stdout


def test_driver_unknown_kernel():
    r = _driver("-d", "8x8x8", "-k", "nosuch")
    assert r.returncode == 1 and 'kernel "nosuch" not found' in r.stdout


def test_oracle_is_not_linked_by_the_product():
    for precision in ("dp", "sp"):
        out = subprocess.check_output(["ldd", lbmhip.library_path(precision)], text=True)
        assert "oracle" not in out and "libref" not in out
        out = subprocess.check_output(["ldd", lbmhip.driver_path(precision)], text=True)
        assert "oracle" not in out


# a LatticeDesc built from a numpy flag array (no geometryCreate, which needs
# the device), then Init of the first registry entry in a child process
_HAND_BUILT = (
    "import sys, ctypes; sys.path.insert(0, %r); import numpy as np; import lbmhip; "
    "lib = lbmhip.Library('dp'); L = lib.lib; ld = lib.LatticeDesc(); "
    "flags = np.ones((6, 6, 6), dtype=np.int32); "
    "ld.Dims = lbmhip.I3(6, 6, 6); ld.Lattice = flags.ctypes.data_as(ctypes.POINTER(ctypes.c_int)); "
    "ld.nCells = ld.nFluid = 216; ld.PeriodicX = ld.PeriodicY = 1; ld.PeriodicZ = %d; "
    "cd = lib.CaseData(); L.caseDataDefaults(ctypes.byref(cd)); cd.Omega = 1.0; %s; "
    "kd = ctypes.POINTER(lib.KernelData)(); "
    "lib.registry[0].Init(ctypes.byref(ld), ctypes.byref(kd), ctypes.byref(cd))")


def test_storage_scale_must_be_power_of_two():
    """rejected at Init, before any device call"""
    import sys
    code = _HAND_BUILT % (os.path.dirname(lbmhip.__file__), 1,
                          "cd.Storage = 2; cd.StorageScale = 3.0")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "power of two" in r.stderr


def test_fluid_on_open_face_is_rejected():
    """the kernels wrap every axis; a non-periodic face must be solid
    (rejected at Init, before any device call)"""
    import sys
    code = _HAND_BUILT % (os.path.dirname(lbmhip.__file__), 0, "pass")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "non-periodic boundary" in r.stderr


def test_reference_driver_links_the_gpu_library():
    """LBM-Bench's unmodified main.c/cli.c built against include/lbm.h"""
    exe = os.path.join(ROOT, "oracle", "_ref", "lbmbench-ref-driver-hip-dp")
    if not os.path.exists(exe):
        pytest.skip("reference driver not built")
    out = subprocess.check_output(["ldd", exe], text=True)
    assert "liblbmhip-dp.so" in out
    r = subprocess.run([exe, "-l"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    listed = [l.strip() for l in r.stdout.split("Available kernels:")[1].splitlines() if l.strip()]
    assert listed == list(lbmhip.KERNELS)


def test_driver_rejects_a_missing_device():
    """-D selects the HIP device; an index beyond the devices present (none
    here, one on a one-GPU machine) is refused with a message and
    EXIT_FAILURE before anything runs"""
    r = subprocess.run([lbmhip.driver_path("dp"), "-d", "8x8x8", "-g", "channel", "-z", "-i", "2",
                        "-D", "9"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "no HIP device 9" in r.stderr


@pytest.mark.parametrize("arg", ["abc", "-3", "1.5", "", "2x"])
def test_driver_rejects_a_malformed_device_index(arg):
    """-D takes a non-negative decimal integer; anything else is refused with
    a message and EXIT_FAILURE instead of silently picking a device"""
    r = subprocess.run([lbmhip.driver_path("dp"), "-d", "8x8x8", "-g", "channel", "-z", "-i", "2",
                        "-D", arg], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "device index must be a non-negative integer" in r.stdout
