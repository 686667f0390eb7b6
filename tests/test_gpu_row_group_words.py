"""The AA odd step of 4- and 2-byte storage as one workgroup per row with
whole words per lane (kRowGroupV): 8-byte words (LBM_ROW_SWEEP=3, four
FP16/BF16 or two FP32 cells per lane) or 4-byte words (LBM_ROW_SWEEP=4, two
FP16/BF16 cells), z shifts as funnel shifts across lanes, the neighbouring
wavefront's end word through LDS, per-cell stores near solids.

Cases: one to eight wavefronts per row (nZ = 256 .. 2048 for FP16 in 8-byte
words, 128 .. 1024 in 4-byte words and for FP32), sphere channels with walls
and solids inside the rows, the z wrap, the body force on.  Checked against the FP64 oracle after an even
and an odd step (the bound of the one-step precision test), and against
the per-block kernel (LBM_ROW_SWEEP=0) over 21 steps: both collide the
same stored values with the same rounding seeds, so they differ only where
the two compilations fuse a multiply-add differently."""
import numpy as np
import pytest
from conftest import WEIGHTS

from lbmhip import Lattice

pytestmark = pytest.mark.gpu

U = {"fp32": 2.0 ** -24, "fp16": 2.0 ** -11, "bf16": 2.0 ** -8}
CASES = [
    ("fp16", 256, "3"), ("fp16", 512, "3"), ("fp16", 1024, "3"), ("fp16", 2048, "3"),
    ("bf16", 512, "3"),
    ("fp32", 128, "3"), ("fp32", 512, "3"), ("fp32", 1024, "3"),
    ("fp16", 128, "4"), ("fp16", 512, "4"), ("fp16", 1024, "4"), ("bf16", 256, "4"),
]


def _run(lib, fmt, dims, steps, init, sweep, monkeypatch):
    with monkeypatch.context() as m:
        m.setenv("LBM_ROW_SWEEP", sweep)
        with Lattice(lib, "aa-soa", "spheres-3", dims, (0, 0, 1), omega=1.7, xforce=1e-4,
                     body_force=True, magic=3 / 16, seed=3, storage=fmt) as L:
            L.set_nodes(init)
            start = L.get_nodes()
            L.run(steps)
            return start, L.get_nodes()


@pytest.mark.parametrize("fmt,nz,sweep", CASES)
def test_row_group_words_match_oracle(gpu_lib, oracle, monkeypatch, fmt, nz, sweep):
    dims = (6, 8, nz)
    lat, per, _ = oracle.geometry("spheres-3", dims, (0, 0, 1), 3)
    fluid = lat.transpose(2, 1, 0) != 0
    assert (~fluid).any() and fluid.any()
    init = oracle.init_random(dims, lat, 8, 1e-2)
    start, got = _run(gpu_lib, fmt, dims, 2, init, sweep, monkeypatch)
    exp, _ = oracle.run("aa-soa", dims, lat, per, start, 2, 1.7, 3 / 16, 1e-4, True)
    dev = np.abs((exp - WEIGHTS)[fluid]).max()
    err = np.abs(got - exp)[fluid].max()
    assert err <= (8 * U[fmt] + 64 * 2.0 ** -24) * dev, (fmt, nz, err, dev)


@pytest.mark.parametrize("fmt,nz,sweep", CASES)
def test_row_group_words_match_per_block_kernel(gpu_lib, oracle, monkeypatch, fmt, nz, sweep):
    dims = (6, 8, nz)
    lat, _, _ = oracle.geometry("spheres-3", dims, (0, 0, 1), 3)
    fluid = lat.transpose(2, 1, 0) != 0
    init = oracle.init_random(dims, lat, 5, 1e-2)
    _, words = _run(gpu_lib, fmt, dims, 21, init, sweep, monkeypatch)
    _, block = _run(gpu_lib, fmt, dims, 21, init, "0", monkeypatch)
    assert np.isfinite(words[fluid]).all()
    dev = np.abs((block - WEIGHTS)[fluid]).max()
    diff = np.abs(words - block)[fluid]
    # a differently fused multiply-add moves a stored value by at most a
    # quantum or two, and that difference then spreads over the steps
    assert diff.max() <= 64 * U[fmt] * dev, (fmt, nz, float(diff.max()), dev)


@pytest.mark.parametrize("geo,per,nz", [("fluid", (1, 1, 1), 512), ("spheres-3", (0, 0, 1), 512),
                                        ("spheres-3", (0, 0, 1), 192)])
def test_default_choice(gpu_lib, oracle, monkeypatch, capfd, geo, per, nz):
    """With LBM_ROW_SWEEP unset, 2-byte AA storage takes the row group of
    4-byte words where its shape fits (nZ a multiple of 128 up to 1024),
    whatever the solids, and the row sweep otherwise; Init says when it
    takes the words, and the run equals the run with that kernel named
    explicitly, bit for bit"""
    dims = (6, 8, nz)
    lat, _, _ = oracle.geometry(geo, dims, per, 3)
    init = oracle.init_random(dims, lat, 5, 1e-2)

    def run(sweep):
        with monkeypatch.context() as m:
            if sweep is None:
                m.delenv("LBM_ROW_SWEEP", raising=False)
            else:
                m.setenv("LBM_ROW_SWEEP", sweep)
            with Lattice(gpu_lib, "aa-soa", geo, dims, per, omega=1.7, xforce=1e-4,
                         body_force=True, magic=3 / 16, seed=3, storage="fp16") as L:
                L.set_nodes(init)
                L.run(9)
                return L.get_nodes()

    capfd.readouterr()
    auto = run(None)
    line = [ln for ln in capfd.readouterr().out.splitlines() if ln.startswith("# AA odd step:")]
    words = nz % 128 == 0
    assert line == (["# AA odd step: row group of 4-byte words"] if words else []), line
    named = run("4" if words else "1")
    assert np.array_equal(auto.view(np.uint64), named.view(np.uint64))
