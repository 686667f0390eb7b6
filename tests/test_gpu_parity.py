"""Parity of the GPU kernels (through the lbm.h C-ABI) with the reference's
golden fixtures and with the oracle, plus the configurations' checks:
  64^3 channel, pull-soa, tau 0.8, Lambda 3/16, g 1e-6, 1000 steps;
  256^3 channel, all seven variants, f_i = w_i (1 + eps), 500 steps;
  every variant 1000 steps on two sphere channels (nZ = 128 and 100).
FP64 tolerance: 1e-12 relative on every fluid PDF."""
import json
import os

import numpy as np
import pytest
from conftest import CX, GOLDEN, WEIGHTS, mark_inlet_outlet, moments, rel_err

from lbmhip import KERNELS, Lattice

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _flags_fluid(L):
    return L.fluid_mask()


@pytest.mark.parametrize("case", [0, 1, 2])
@pytest.mark.parametrize("kernel", KERNELS)
def test_matches_reference_fixtures(gpu_lib, golden_runs, manifest, case, kernel):
    g, dims, per = manifest["run_cases"][case]
    dims = tuple(dims)
    init = golden_runs[f"r{case}_init"]
    with Lattice(gpu_lib, kernel, g, dims, per, omega=manifest["omega"],
                 xforce=manifest["xforce"], body_force=True, magic=manifest["magic"]) as L:
        assert (L.lattice_flags() == golden_runs[f"r{case}_flags"]).all()
        fluid = _flags_fluid(L)
        for n in manifest["steps"]:
            L.set_nodes(init)
            L.run(n)
            done = n + (n % 2 if kernel == "aa-soa" else 0)
            key = f"r{case}_n{done}"
            if key not in golden_runs:
                continue
            err = rel_err(L.get_nodes(), golden_runs[key], fluid)
            assert err < TOL, (kernel, n, err)


CASES = [
    ("channel", (12, 13, 14), (0, 0, 1)),
    ("fluid", (10, 11, 12), (0, 0, 0)),
    ("blocks-4", (12, 12, 12), (0, 0, 0)),
    ("box", (10, 11, 12), (1, 0, 1)),
    ("pipe", (12, 12, 12), (0, 0, 0)),
    ("spheres-6", (20, 22, 70), (0, 0, 1)),
    # nZ a multiple of 64 with several 64-blocks per row: the row-sweep
    # push and AA odd kernels (aligned scatter, carries between blocks and
    # the z wrap), with solids inside the rows, walls, and no solids at all
    ("spheres-6", (16, 18, 128), (0, 0, 1)),
    ("box", (10, 11, 192), (1, 0, 1)),
    ("fluid", (6, 7, 192), (1, 1, 1)),
    # the row group at its largest workgroup (eight 64-blocks per row), and
    # nZ = 576, past it, where the row sweep runs instead
    ("spheres-3", (6, 8, 512), (0, 0, 1)),
    ("spheres-3", (4, 6, 576), (0, 0, 1)),
]


@pytest.mark.parametrize("geo,dims,per", CASES)
@pytest.mark.parametrize("kernel", KERNELS)
def test_matches_oracle(gpu_lib, oracle, geo, dims, per, kernel):
    """Lambda 3/16, with and without the force, an odd step count; the
    first spheres case has nZ = 70 (a partial 64-wide segment) and wraps
    spheres through the periodic faces"""
    lat, per_o, _ = oracle.geometry(geo, dims, per, 3)
    init = oracle.init_random(dims, lat, 17, 1e-2)
    fluid = lat.transpose(2, 1, 0) != 0
    for force in (False, True):
        with Lattice(gpu_lib, kernel, geo, dims, per, omega=1.7, xforce=1e-4,
                     body_force=force, magic=3 / 16, seed=3) as L:
            assert (L.lattice_flags() == lat).all()
            L.set_nodes(init)
            L.run(7)
            exp, done = oracle.run(kernel, dims, lat, per_o, init, 7, 1.7, 3 / 16, 1e-4, force)
            assert rel_err(L.get_nodes(), exp, fluid) < TOL


# degenerate shapes: an axis of one or two cells, whose neighbours wrap onto
# the cell itself or its one partner, a single fluid row between two walls,
# and the row sweep (nZ = 64, 128) on rows that are their own x neighbours
SMALL = [
    ("fluid", (1, 4, 5), (1, 1, 1)),
    ("fluid", (2, 1, 7), (1, 1, 1)),
    ("fluid", (3, 5, 1), (1, 1, 1)),
    ("fluid", (2, 2, 2), (1, 1, 1)),
    ("fluid", (1, 1, 1), (1, 1, 1)),
    ("channel", (2, 3, 64), (0, 0, 1)),
    ("channel", (1, 5, 128), (0, 0, 1)),
]


@pytest.mark.parametrize("geo,dims,per", SMALL)
@pytest.mark.parametrize("kernel", KERNELS)
def test_degenerate_shapes_match_oracle(gpu_lib, oracle, geo, dims, per, kernel):
    lat, per_o, _ = oracle.geometry(geo, dims, per, 3)
    init = oracle.init_random(dims, lat, 21, 1e-2)
    fluid = lat.transpose(2, 1, 0) != 0
    assert fluid.any()
    with Lattice(gpu_lib, kernel, geo, dims, per, omega=1.4, xforce=1e-4, body_force=True,
                 magic=3 / 16, seed=3) as L:
        L.set_nodes(init)
        L.run(5)
        exp, _ = oracle.run(kernel, dims, lat, per_o, init, 5, 1.4, 3 / 16, 1e-4, True)
        assert rel_err(L.get_nodes(), exp, fluid) < TOL


@pytest.mark.parametrize("kernel", KERNELS)
def test_split_runs_equal_one_run(gpu_lib, kernel):
    dims = (12, 14, 16)
    with Lattice(gpu_lib, kernel, "blocks-4", dims, omega=1.3, xforce=1e-4,
                 body_force=True, magic=3 / 16) as A, \
         Lattice(gpu_lib, kernel, "blocks-4", dims, omega=1.3, xforce=1e-4,
                 body_force=True, magic=3 / 16) as B:
        A.init("random", 5, 1.0, 1e-2)
        B.init("random", 5, 1.0, 1e-2)
        A.run(4)
        A.run(6)
        B.run(10)
        assert np.array_equal(A.get_nodes(), B.get_nodes())


def test_device_init_equals_oracle_init(gpu_lib, oracle):
    dims = (16, 18, 20)
    lat, _, _ = oracle.geometry("channel", dims, (0, 0, 1))
    exp = oracle.init_random(dims, lat, 1234, 1e-3)
    for kernel in KERNELS:
        with Lattice(gpu_lib, kernel, "channel", dims, (0, 0, 1)) as L:
            L.init("random", 1234, 1.0, 1e-3)
            assert np.array_equal(L.get_nodes(), exp), kernel
            L.init("rest", 0, 1.0, 0.0)
            got = L.get_nodes()[lat.transpose(2, 1, 0) != 0]
            assert np.array_equal(got, np.broadcast_to(WEIGHTS, got.shape))


def test_node_accessors_round_trip(gpu_lib):
    dims = (8, 9, 10)
    for kernel in KERNELS:
        with Lattice(gpu_lib, kernel, "channel", dims, (0, 0, 1)) as L:
            L.init("random", 3, 1.0, 1e-2)
            v = np.linspace(0.1, 1.9, 19)
            L.set_node(4, 5, 6, v)
            assert np.array_equal(L.get_node(4, 5, 6), v)
            full = L.get_nodes()
            assert np.array_equal(full[4, 5, 6], v)
            L.set_nodes(full)
            assert np.array_equal(L.get_nodes(), full)


def test_reductions_match_numpy(gpu_lib, oracle):
    dims = (20, 24, 28)
    lat, per, _ = oracle.geometry("blocks-4", dims)
    init = oracle.init_random(dims, lat, 9, 1e-2)
    fluid = lat.transpose(2, 1, 0) != 0
    for kernel in ("pull-soa", "pull-aos", "aa-soa"):
        with Lattice(gpu_lib, kernel, "blocks-4", dims, omega=1.1, xforce=1e-4,
                     body_force=True) as L:
            L.set_nodes(init)
            L.run(5)
            state = L.get_nodes()
            rho, j = moments(state)
            m = L.mass_momentum()
            exp = np.array([rho[fluid].sum(), *j[fluid].sum(0)])
            assert abs(m[0] - exp[0]) < 1e-12 * exp[0]
            assert np.abs(m[1:] - exp[1:]).max() < 1e-12 * exp[0]
            assert abs(L.density() - exp[0] / fluid.sum()) < 1e-14


def test_add_body_force(gpu_lib, oracle):
    dims = (8, 10, 12)
    g = 1e-3
    for kernel in ("pull-soa", "push-soa", "aa-soa"):
        with Lattice(gpu_lib, kernel, "channel", dims, (0, 0, 1), xforce=g) as L:
            L.init("random", 1, 1.0, 1e-2)
            before = L.get_nodes()
            L.add_body_force()
            after = L.get_nodes()
            fluid = L.fluid_mask()
            F = 3.0 * WEIGHTS * CX * g
            F[[2, 4, 5, 10, 17]] = [3.0 * (1 / 18) * g] + [3.0 * (1 / 36) * g] * 4
            F[[3, 6, 7, 11, 16]] = -F[[2, 4, 5, 10, 17]]
            assert np.array_equal(after[fluid], before[fluid] + F)


def test_channel_64_pull_soa(gpu_lib, oracle):
    """64^3 channel (walls at y), pull-soa, tau 0.8, Lambda 3/16, g 1e-6,
    rest start, 1000 steps: every PDF against the oracle, mass conservation,
    and the start-up Poiseuille profile"""
    dims, omega, g, steps = (64, 64, 64), 1.25, 1e-6, 1000
    lat, per, counts = oracle.geometry("channel", dims, (0, 0, 1))
    init = oracle.init_random(dims, lat, 0, 0.0)
    exp, _ = oracle.run("pull-soa", dims, lat, per, init, steps, omega, 3 / 16, g, True)
    fluid = lat.transpose(2, 1, 0) != 0
    with Lattice(gpu_lib, "pull-soa", "channel", dims, (0, 0, 1), omega=omega, xforce=g,
                 body_force=True, magic=3 / 16) as L:
        L.init("rest", 0, 1.0, 0.0)
        m0 = L.mass_momentum()
        L.run(steps)
        got = L.get_nodes()
        m1 = L.mass_momentum()
    assert rel_err(got, exp, fluid) < TOL
    assert abs(m1[0] - m0[0]) / m0[0] < 1e-13
    # momentum gained = g per fluid cell per step minus the wall friction,
    # which the start-up series gives: compare the mean profile over x, z
    nu = (1 / omega - 0.5) / 3
    H = dims[1] - 2
    y = np.arange(dims[1])
    eta = y - 0.5
    series = g / (2 * nu) * eta * (H - eta)
    for n in range(1, 400, 2):
        series -= (4 * g * H * H / (nu * (n * np.pi) ** 3) * np.sin(n * np.pi * eta / H)
                   * np.exp(-nu * (n * np.pi / H) ** 2 * steps))
    ux = (got * CX).sum(-1).mean(axis=(0, 2)) - g / 2
    fl = slice(1, dims[1] - 1)
    assert np.abs(ux[fl] - series[fl]).max() / series[fl].max() < 2e-2


@pytest.fixture(scope="module")
def channel256(gpu_lib, oracle):
    dims = (256, 256, 256)
    lat, per, counts = oracle.geometry("channel", dims, (0, 0, 1))
    return dims, lat, per


@pytest.mark.timeout(900)
def test_channel_256_all_variants(gpu_lib, oracle, channel256):
    """256^3 channel (walls at y, periodic x and z), the seven variants from
    the same seeded start f = w (1 + eps), eps ~ U(-1e-3, 1e-3), omega 1.25,
    Lambda 3/16, body force 1e-6, 500 steps: every fluid PDF of every
    variant within 1e-12 of the oracle's 500-step state, and mass conserved.
    The oracle runs in a thread (ctypes releases the GIL) while the GPU
    runs the variants."""
    import threading
    import time
    dims, lat, per = channel256
    fluid = lat.transpose(2, 1, 0) != 0
    kw = dict(omega=1.25, xforce=1e-6, body_force=True, magic=3 / 16)
    init = oracle.init_random(dims, lat, 0, 1e-3)
    box = {}

    def run_oracle():
        t0 = time.perf_counter()
        box["exp"], box["done"] = oracle.run("pull-soa", dims, lat, per, init, 500, 1.25, 3 / 16,
                                             1e-6, True)
        box["seconds"] = time.perf_counter() - t0

    th = threading.Thread(target=run_oracle)
    th.start()
    states = {}
    try:
        for kernel in KERNELS:
            with Lattice(gpu_lib, kernel, "channel", dims, (0, 0, 1), **kw) as L:
                L.init("random", 0, 1.0, 1e-3)
                m0 = L.mass_momentum()
                L.run(500)
                m1 = L.mass_momentum()
                assert abs(m1[0] - m0[0]) / m0[0] < 1e-13, kernel
                states[kernel] = L.get_nodes()
            print(f"{kernel}: 500 steps done", flush=True)
    finally:
        th.join()
    assert box["done"] == 500
    print(f"oracle: 256^3 x 500 steps in {box['seconds']:.1f} s", flush=True)
    for kernel, state in states.items():
        err = rel_err(state, box["exp"], fluid)
        print(f"{kernel}: max rel err vs oracle after 500 steps {err:.3e}")
        assert err < TOL, kernel


@pytest.mark.parametrize("kernel", ["pull-soa", "aa-soa", "push-aos"])
def test_plane_velocity_matches_nodes(gpu_lib, oracle, kernel):
    """lbmPlaneVelocity (kernelVerify's profile, reduced on the device)
    equals the x momentum of the logical PDFs summed in the reference's
    order, bit for bit, after an even and an odd number of steps"""
    dims = (12, 14, 70)
    for steps in (3, 4):
        with Lattice(gpu_lib, kernel, "spheres-6", dims, (0, 0, 1), omega=1.3, xforce=1e-4,
                     body_force=True, magic=3 / 16, seed=3) as L:
            L.init("random", 2, 1.0, 1e-2)
            L.run(steps)
            f = L.get_nodes()
            for x in (0, 5, 11):
                got = L.plane_velocity(x)
                e = f[x]
                exp = (e[..., 2] + e[..., 4] + e[..., 5] + e[..., 10] + e[..., 17] - e[..., 3]
                       - e[..., 6] - e[..., 7] - e[..., 11] - e[..., 16])
                assert np.array_equal(got, exp), (kernel, steps, x)


def test_verification_case(gpu_lib):
    """LBM-Bench's -V case through the library: kernelSetInitial*, 1000
    steps with the force, kernelVerify; the norm the reference prints"""
    with open(os.path.join(GOLDEN, "verification.json")) as f:
        ver = json.load(f)
    for kernel in KERNELS:
        with Lattice(gpu_lib, kernel, "box", (16, 16, 16), (1, 1, 0), omega=1.0, xforce=1e-5,
                     body_force=True) as L:
            L.init_reference()
            L.run(1000)
            norm = L.verify()
            assert abs(norm - ver["dp"][kernel]) < 1e-9 * ver["dp"][kernel], kernel


def test_single_precision_build(gpu_lib_sp):
    """PdfType float build against the reference's PRECISION=sp fixture"""
    sp = np.load(os.path.join(GOLDEN, "runs_sp.npz"))
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        man = json.load(f)
    with open(os.path.join(GOLDEN, "verification.json")) as f:
        ver = json.load(f)
    for kernel in KERNELS:
        with Lattice(gpu_lib_sp, kernel, "channel", (6, 8, 10), (0, 0, 1), omega=man["omega"],
                     xforce=man["xforce"], body_force=True) as L:
            fluid = L.fluid_mask()
            L.set_nodes(sp["init"].astype(np.float32).astype(np.float64))
            L.run(10)
            exp = sp[f"n10_{kernel}"]
            assert rel_err(L.get_nodes(), exp, fluid) < 2e-5, kernel
        with Lattice(gpu_lib_sp, kernel, "box", (16, 16, 16), (1, 1, 0), omega=1.0, xforce=1e-5,
                     body_force=True) as L:
            L.init_reference()
            L.run(1000)
            # in float the norm carries rounding noise: the reference's own sp
            # variants spread by 0.7 % (0.04340 ... 0.04371)
            spread = max(ver["sp"].values()) - min(ver["sp"].values())
            assert abs(L.verify() - ver["sp"][kernel]) < 2 * spread, kernel


@pytest.mark.parametrize("kernel", KERNELS)
def test_zou_he_matches_oracle(gpu_lib, oracle, kernel):
    """kernelComputeBoundaryConditions (src/lbm.c:676-785) through the
    library's boundary accessors after 0 and 5 steps, against the oracle's
    restatement (pinned to the compiled reference in
    test_oracle_vs_reference.py)"""
    dims, per = (12, 11, 10), (1, 0, 1)
    lat, per_o, _ = oracle.geometry("channel", dims, per)
    init = oracle.init_random(dims, lat, 21, 1e-2)
    fluid = lat.transpose(2, 1, 0) != 0
    marked = mark_inlet_outlet(lat)
    for steps in (0, 5):
        with Lattice(gpu_lib, kernel, "channel", dims, per, omega=1.25, xforce=1e-4,
                     body_force=True) as L:
            L.set_nodes(init)
            L.run(steps)
            L.set_flags(marked)
            L.boundary_conditions(1.01, 0.99)
            got = L.get_nodes()
        run, _ = oracle.run(kernel, dims, lat, per_o, init, steps, 1.25, 1 / 12, 1e-4, True)
        exp = oracle.zou_he(dims, marked, run, 1.01, 0.99)
        assert rel_err(got, exp, fluid) < TOL, (kernel, steps)


@pytest.mark.parametrize("force", [False, True])
@pytest.mark.parametrize("kernel", KERNELS)
def test_pressure_boundaries_every_step_match_oracle(gpu_lib, oracle, kernel, force):
    """CaseDataType.PressureBoundaries: the Zou-He inlet / outlet
    (src/lbm.c:676-785) applied on the device after every step, inside the
    captured step graphs, 20 steps (one graph of 16 plus eager steps) and 7
    more in a second run, against the oracle's time loop with the boundary
    pass after every step; every fluid PDF within 1e-12"""
    dims, per = (14, 11, 10), (1, 0, 1)
    lat, per_o, _ = oracle.geometry("channel", dims, per)
    init = oracle.init_random(dims, lat, 21, 1e-2)
    fluid = lat.transpose(2, 1, 0) != 0
    marked = mark_inlet_outlet(lat)
    kw = dict(omega=1.3, xforce=1e-4, body_force=force, magic=3 / 16, pressure=(1.01, 0.99))
    with Lattice(gpu_lib, kernel, "channel", dims, per, **kw) as L:
        L.set_flags(marked)
        L.set_nodes(init)
        L.run(20)
        got20 = L.get_nodes()
        L.run(8)
        got28 = L.get_nodes()
    exp20, _ = oracle.run_bc(kernel, dims, marked, per_o, init, 20, 1.3, 3 / 16, 1e-4, force,
                             1.01, 0.99)
    exp28, _ = oracle.run_bc(kernel, dims, marked, per_o, exp20, 8, 1.3, 3 / 16, 1e-4, force,
                             1.01, 0.99)
    assert rel_err(got20, exp20, fluid) < TOL
    assert rel_err(got28, exp28, fluid) < TOL
    plain, _ = oracle.run(kernel, dims, lat, per_o, init, 20, 1.3, 3 / 16, 1e-4, force)
    assert rel_err(got20, plain, fluid) > 1e-6


# nZ a multiple of 64 (two 64-blocks per row: the row sweep's carries and
# z wrap) and nZ = 100 (a partial last segment: kStepRow / kStepAoS lanes
# past nZ), both with spheres cut by the walls and the periodic faces
LONG_CASES = [("spheres-12", (32, 40, 128), 9), ("spheres-8", (20, 26, 100), 4)]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("geo,dims,seed", LONG_CASES)
@pytest.mark.parametrize("kernel", KERNELS)
def test_thousand_steps_match_oracle(gpu_lib, oracle, kernel, geo, dims, seed):
    """1000 steps of every variant (spheres and walls, the force on,
    Lambda 3/16) against the oracle, every fluid PDF within 1e-12
    (src/kernel_*.c time loops)"""
    steps = 1000
    lat, per, _ = oracle.geometry(geo, dims, (0, 0, 1), seed)
    init = oracle.init_random(dims, lat, 5, 1e-2)
    exp, _ = oracle.run(kernel, dims, lat, per, init, steps, 1.25, 3 / 16, 1e-5, True)
    fluid = lat.transpose(2, 1, 0) != 0
    with Lattice(gpu_lib, kernel, geo, dims, (0, 0, 1), omega=1.25, xforce=1e-5,
                 body_force=True, magic=3 / 16, seed=seed) as L:
        assert (L.lattice_flags() == lat).all()
        L.set_nodes(init)
        L.run(steps)
        got = L.get_nodes()
    err = rel_err(got, exp, fluid)
    print(f"{kernel} {geo} {dims}: max rel err after {steps} steps {err:.3e}")
    assert err < TOL


def test_push_aos_post_planes_counted(gpu_lib, monkeypatch):
    """push-aos in two passes holds Q x cells extra elements of
    post-collision planes, and lbmDeviceBytes counts them"""
    dims = (6, 7, 70)
    sizes = {}
    for passes in ("1", "2"):
        monkeypatch.setenv("LBM_AOS_PUSH_PASSES", passes)
        with Lattice(gpu_lib, "push-aos", "channel", dims, (0, 0, 1)) as L:
            sizes[passes] = L.device_bytes()
    assert sizes["2"] - sizes["1"] == 19 * 6 * 7 * 70 * 8
