"""Launch shape does not change results.

Every launch knob (block order across XCDs, row order in bands, workgroups
per CU, push-aos in one pass or two, non-temporal hints (also those of
the row group's blocks near a solid), the row sweep
against the per-block kernel, wavefronts per workgroup, plane padding)
picks another way to run the same arithmetic on the same slots, and the
stochastic rounding seeds depend only on cell and step.  So each setting
must give the default run's state bit for bit, for FP64 and for FP16
storage, on a sphere channel with several 64-cell blocks per row (walls,
solids inside the rows, the z wrap) and the body force on.
LBM_ROW_SWEEP (kRowGroup / kRowSweep / kStepRow) and LBM_AOS_PUSH_PASSES (kStepAoS /
kPushCollideAoS) swap kernels, whose compilations fuse multiply-adds of
the collision differently: there the bound is 1e-14 relative (FP64) or a
few FP16 quanta of the stored deviation (FP16)."""
import numpy as np
import pytest
from conftest import WEIGHTS

from lbmhip import Lattice

pytestmark = pytest.mark.gpu

DIMS = (14, 16, 128)
KNOBS = [
    {"LBM_XCD_REMAP": "1"},
    {"LBM_XCD_REMAP": "0"},
    {"LBM_NT": "0"},
    {"LBM_NT": "1"},
    {"LBM_NT": "2"},
    {"LBM_NT": "3"},
    {"LBM_ROW_SWEEP": "0"},
    {"LBM_ROW_SWEEP": "1"},
    {"LBM_ROW_SWEEP": "3"},
    {"LBM_WALL_NT": "0"},
    {"LBM_WALL_NT": "7"},
    {"LBM_WAVES_PER_BLOCK": "8"},
    {"LBM_PLANE_PAD": "4096"},
    {"LBM_GRAPH_STEPS": "4"},
    {"LBM_ROW_BAND": "0"},
    {"LBM_ROW_BAND": "3"},
    {"LBM_ROW_BAND": "16"},
    {"LBM_AOS_WG_PER_CU": "0"},
    {"LBM_AOS_WG_PER_CU": "1"},
    {"LBM_AOS_PUSH_PASSES": "1"},
]


def _state(lib, kernel, storage, steps=7):
    with Lattice(lib, kernel, "spheres-6", DIMS, (0, 0, 1), omega=1.25, xforce=1e-5,
                 body_force=True, magic=3 / 16, storage=storage) as L:
        L.init("random", 9, 1.0, 1e-3)
        L.run(steps)
        return L.get_nodes()


@pytest.mark.parametrize("storage", ["native", "fp16"])
@pytest.mark.parametrize("kernel", ["aa-soa", "push-soa", "pull-soa", "pull-aos", "push-aos",
                                    "blk-push-soa", "blk-pull-soa"])
def test_knobs_leave_results_bit_identical(gpu_lib, monkeypatch, kernel, storage):
    ref = _state(gpu_lib, kernel, storage)
    assert np.isfinite(ref).all()
    for knob in KNOBS:
        with monkeypatch.context() as m:
            for k, v in knob.items():
                m.setenv(k, v)
            got = _state(gpu_lib, kernel, storage)
        if "LBM_ROW_SWEEP" in knob or "LBM_AOS_PUSH_PASSES" in knob:
            fluid = ref != 0
            if storage == "native":
                err = float(np.abs(got - ref)[fluid].max() / np.abs(ref[fluid]).max())
                assert err < 1e-14, (kernel, storage, knob, err)
            else:
                assert float(np.abs(got - ref).max()) < 1e-5, (kernel, storage, knob)
            continue
        same = np.array_equal(got.view(np.uint64), ref.view(np.uint64))
        assert same, (kernel, storage, knob, float(np.abs(got - ref).max()))


def _state_fluid(lib, kernel, storage, steps):
    with Lattice(lib, kernel, "spheres-6", DIMS, (0, 0, 1), omega=1.25, xforce=1e-5,
                 body_force=True, magic=3 / 16, storage=storage) as L:
        L.init("random", 9, 1.0, 1e-3)
        L.run(steps)
        return L.get_nodes(), L.fluid_mask()


@pytest.mark.parametrize("precision,storage", [("dp", "native"), ("dp", "fp32"), ("sp", "native")])
@pytest.mark.parametrize("kernel", ["aa-soa", "push-soa", "blk-push-soa"])
def test_closing_step_equals_force_pass(request, monkeypatch, precision, storage, kernel):
    """The last step of a run adds its own body force at the store
    (LBM_CLOSE_FUSED, default on) instead of a force pass after it: the
    fluid cells' PDFs are the same bit for bit, after runs of one step, an
    even and odd number of steps, and more steps than a step graph holds;
    the row group, row sweep and per-block kernels each close a run"""
    lib = request.getfixturevalue("gpu_lib" if precision == "dp" else "gpu_lib_sp")
    for steps in (1, 2, 7, 40):
        for sweep in ("3", "1", "0"):
            with monkeypatch.context() as m:
                m.setenv("LBM_ROW_SWEEP", sweep)
                fused, fluid = _state_fluid(lib, kernel, storage, steps)
                m.setenv("LBM_CLOSE_FUSED", "0")
                ref, _ = _state_fluid(lib, kernel, storage, steps)
            assert np.isfinite(ref[fluid]).all()
            w = np.uint64 if ref.dtype == np.float64 else np.uint32
            same = np.array_equal(fused[fluid].view(w), ref[fluid].view(w))
            assert same, (kernel, storage, precision, steps, sweep,
                          float(np.abs(fused - ref)[fluid].max()))


@pytest.mark.parametrize("storage,u", [("fp16", 2.0 ** -11), ("bf16", 2.0 ** -8)])
@pytest.mark.parametrize("kernel", ["aa-soa", "push-soa", "blk-push-soa"])
def test_closing_step_rounds_once(request, monkeypatch, storage, u, kernel):
    """With stochastically rounded storage the closing step rounds the
    collided value plus the force once, where the force pass rounds the
    collided value and then the sum: the fluid PDFs differ by the last
    rounding only, after runs of one, an even and odd number of steps and
    more than a step graph, for the row group, row sweep and per-block
    kernels"""
    lib = request.getfixturevalue("gpu_lib")
    for steps in (1, 2, 7, 40):
        for sweep in ("3", "1", "0"):
            with monkeypatch.context() as m:
                m.setenv("LBM_ROW_SWEEP", sweep)
                fused, fluid = _state_fluid(lib, kernel, storage, steps)
                m.setenv("LBM_CLOSE_FUSED", "0")
                ref, _ = _state_fluid(lib, kernel, storage, steps)
            assert np.isfinite(fused[fluid]).all()
            dev = np.abs(ref - WEIGHTS)[fluid].max()
            diff = float(np.abs(fused - ref)[fluid].max())
            assert diff <= 4 * u * dev, (kernel, storage, steps, sweep, diff, dev)

