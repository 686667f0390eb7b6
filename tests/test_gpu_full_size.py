"""The bench workload at its full size: a 512^3 channel (walls at y,
periodic x and z) with 2000 seeded spheres r ~ U[4,12], aa-soa, omega
1.25, Lambda 3/16, body force 1e-6, f = w(1+eps) from seed 0.  The oracle
does not finish this size in seconds, so the checks are properties that
hold at any size:
  - mass is conserved: the body force and half-way bounce-back are both
    mass-neutral, so the FP64 mass drifts by round-off only;
  - from rest one pull step adds exactly g of x momentum per fluid cell:
    rest populations stream, bounce back and collide unchanged, and the
    force adds 3 w_i c_ix g, which sums to g over the 19 directions (the
    AA variant runs steps in pairs, src/kernel_aa_soa.c:36, so its first
    run from rest already meets the walls);
  - an AA run of 1 step runs 2, as the reference's loop does; steps split
    into eager launches and graph replays (1 -> 2, then 32) give the
    per-cell moments of one run of 34 steps bit for bit;
  - FP16 storage stays within the precision form's velocity error bound of
    1e-3 against the FP64 run;
  - over 5000 steps, FP32, FP16 and BF16 storage stay within stated
    velocity-error and mass-drift bounds of the FP64 run.
Device memory: 20.4 GB of FP64 PDFs (40.8 GB for the two-array pull step)
and 5.1 GB of FP16 at most at once; host memory: two 4.3 GB arrays of
per-cell moments."""
import numpy as np
import pytest

from lbmhip import Lattice

pytestmark = pytest.mark.gpu

DIMS = (512, 512, 512)
GEOMETRY = "spheres-2000"
G = 1e-6
KW = dict(omega=1.25, xforce=G, body_force=True, magic=3 / 16, seed=0)
STEPS = 34


def _lattice(lib, storage="native", kernel="aa-soa"):
    return Lattice(lib, kernel, GEOMETRY, DIMS, (0, 0, 1), storage=storage, **KW)


def test_first_step_from_rest_adds_g_per_fluid_cell(gpu_lib):
    with _lattice(gpu_lib, kernel="pull-soa") as L:
        L.init("rest", 0, 1.0, 0.0)
        m0 = L.mass_momentum()
        L.run(1)
        m1 = L.mass_momentum()
        nf = L.n_fluid
    print(f"{nf} fluid cells; jx after one step {m1[1]:.15e}, g * fluid {G * nf:.15e}")
    assert nf == 128_407_050
    assert abs(m1[1] / (G * nf) - 1.0) < 1e-9
    assert abs(m1[2]) < 1e-9 * G * nf and abs(m1[3]) < 1e-9 * G * nf
    assert abs(m1[0] - m0[0]) / m0[0] < 1e-14


@pytest.fixture(scope="module")
def split_and_whole(gpu_lib):
    """per-cell moments after runs of 1 (that is, 2) and 32 steps and after
    one run of 34 steps, the mass before and after, and the FP64 lattice of
    the whole run (kept open for the FP16 comparison)"""
    with _lattice(gpu_lib) as A:
        A.init("random", 0, 1.0, 1e-3)
        m0 = A.mass_momentum()
        A.run(1)
        A.run(STEPS - 2)
        ma = A.mass_momentum()
        split = A.cell_moments()
    B = _lattice(gpu_lib)
    B.init("random", 0, 1.0, 1e-3)
    B.run(STEPS)
    whole = B.cell_moments()
    yield m0, ma, split, whole, B
    B.close()


def test_mass_conserved(split_and_whole):
    m0, ma, _, _, _ = split_and_whole
    drift = abs(ma[0] - m0[0]) / m0[0]
    print(f"mass {m0[0]:.17e} -> {ma[0]:.17e}, relative drift {drift:.2e}")
    assert drift < 1e-13


def test_split_run_equals_one_run(split_and_whole):
    _, _, split, whole, _ = split_and_whole
    assert np.isfinite(whole).all()
    assert np.array_equal(split.view(np.uint64), whole.view(np.uint64))


def test_fp16_storage_velocity_error(gpu_lib, split_and_whole):
    ref = split_and_whole[4]
    with _lattice(gpu_lib, "fp16") as L:
        L.init("random", 0, 1.0, 1e-3)
        L.run(STEPS)
        cmp = L.compare_momentum(ref)
    l2 = (cmp[0] / cmp[1]) ** 0.5
    print(f"fp16 velocity L2 error after {STEPS} steps: {l2:.3e}")
    assert l2 <= 1e-3


# bounds over 5000 steps (measured on MI355X: velocity L2 error FP32 7.2e-5,
# FP16 2.4e-4, BF16 1.4e-3; mass drift FP64 -3.5e-13, FP32 -6.7e-10, FP16
# 4.5e-9, BF16 -2.3e-9).  FP16's 1e-3 is the precision form's target.
HORIZON = 5000
FORM_BOUNDS = {"fp32": (5e-4, 1e-8), "fp16": (1e-3, 1e-7), "bf16": (5e-3, 1e-7)}


@pytest.mark.timeout(900)
def test_storage_forms_over_5000_steps(gpu_lib):
    """The precision form's workload at full size: the bench lattice run
    5000 steps in FP64 and in FP32, FP16 and BF16 shifted-deviation storage
    (FP32 collision, stochastic rounding for the 16-bit forms); each form's
    velocity L2 error against the FP64 run and its mass drift are bounded,
    and the FP64 mass drifts by round-off only"""
    with _lattice(gpu_lib) as ref:
        ref.init("random", 0, 1.0, 1e-3)
        m0 = ref.mass_momentum()
        for k in range(HORIZON // 1000):
            ref.run(1000)
            print(f"fp64: {1000 * (k + 1)} steps", flush=True)
        m1 = ref.mass_momentum()
        drift64 = abs(m1[0] - m0[0]) / m0[0]
        print(f"fp64 mass drift after {HORIZON} steps {drift64:.2e}", flush=True)
        assert np.isfinite(m1).all() and drift64 < 1e-11
        for form, (u_bound, m_bound) in FORM_BOUNDS.items():
            with _lattice(gpu_lib, form) as L:
                L.init("random", 0, 1.0, 1e-3)
                mf0 = L.mass_momentum()
                L.run(HORIZON)
                mf = L.mass_momentum()
                cmp = L.compare_momentum(ref)
            l2 = (cmp[0] / cmp[1]) ** 0.5
            drift = abs(mf[0] - mf0[0]) / mf0[0]
            print(f"{form}: velocity L2 error {l2:.3e}, mass drift {drift:.2e} after {HORIZON} "
                  f"steps", flush=True)
            assert l2 <= u_bound, form
            assert drift <= m_bound, form


def test_device_flags_equal_oracle_at_full_size(gpu_lib, oracle):
    """geometryCreate builds the flags on the device (walls, 2000 spheres
    wrapped through the periodic faces); they equal the oracle's host
    restatement cell for cell, and the fluid count matches"""
    with _lattice(gpu_lib) as L:
        got = L.lattice_flags()
        nf = L.n_fluid
    exp, per, counts = oracle.geometry(GEOMETRY, DIMS, (0, 0, 1), 0)
    assert per == (1, 0, 1)
    assert counts[0] == nf == 128_407_050
    assert np.array_equal(got, exp)


def test_verify_reads_one_plane(gpu_lib):
    """kernelVerify at 512^3 reduces the x = nX/2 plane on the device: the
    norm is finite and no host mirror of the lattice (20.4 GB) is made, so
    kd->PdfsActive stays NULL"""
    with _lattice(gpu_lib) as L:
        L.init("random", 0, 1.0, 1e-3)
        L.run(2)
        norm = L.verify()
        assert not L.kd.contents.PdfsActive
    assert np.isfinite(norm) and norm > 0


def test_placement_probe_keeps_results(gpu_lib, monkeypatch):
    """The 20 GB PDF array is chosen among probed candidate allocations
    (LBM_PLACEMENT_PROBE): the probe reports what it did, and the run's
    momentum and density fields equal those of a lattice allocated without
    the probe exactly"""
    with _lattice(gpu_lib) as A:
        n, gbs = A.placement()
        print(f"placement: {n} candidate(s), kept one at {gbs:.0f} GB/s", flush=True)
        assert n >= 1 and gbs > 0.0
        with monkeypatch.context() as m:
            m.setenv("LBM_PLACEMENT_PROBE", "0")
            B = _lattice(gpu_lib)
        try:
            assert B.placement() == (1, 0.0)
            for L in (A, B):
                L.init("random", 0, 1.0, 1e-3)
                L.run(4)
            cmp = A.compare_momentum(B)
            assert cmp[0] == 0.0 and cmp[2] == 0.0 and cmp[3] == 0.0
            assert np.array_equal(A.mass_momentum(), B.mass_momentum())
        finally:
            B.close()
