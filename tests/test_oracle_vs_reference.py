"""The oracle against LBM-Bench's own sources compiled into oracle/_ref/
(only where that build exists; the golden fixtures carry the same checks
to machines without it)."""
import numpy as np
import pytest
from conftest import mark_inlet_outlet, rel_err

from oracle import KERNELS, Reference

pytestmark = pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built")

CASES = [
    ("channel", (12, 12, 12), (0, 0, 1)),
    ("fluid", (10, 11, 12), (0, 0, 0)),
    ("blocks-4", (12, 12, 12), (0, 0, 0)),
    ("box", (10, 11, 12), (1, 0, 1)),
    ("pipe", (12, 12, 12), (0, 0, 0)),
    ("box", (7, 9, 8), (0, 1, 0)),
]


@pytest.fixture(scope="module")
def ref():
    return Reference()


@pytest.mark.parametrize("geo,dims,per", CASES)
def test_geometry_matches_reference(oracle, ref, geo, dims, per):
    a, pa, ca = oracle.geometry(geo, dims, per)
    b, pb, cb = ref.geometry(geo, dims, per)
    assert (a == b).all() and pa == pb and ca == cb


@pytest.mark.parametrize("geo,dims,per", CASES)
@pytest.mark.parametrize("kernel", KERNELS)
def test_run_matches_reference(oracle, ref, geo, dims, per, kernel):
    lat, per_o, _ = oracle.geometry(geo, dims, per)
    init = oracle.init_random(dims, lat, 11, 1e-2)
    fluid = lat.transpose(2, 1, 0) != 0
    for steps in (1, 2, 5, 20):
        got, _ = oracle.run(kernel, dims, lat, per_o, init, steps, 1.25, 1 / 12, 1e-4, True)
        exp, stats = ref.run(kernel, geo, dims, per, init, steps, 1.25, 1e-4)
        assert rel_err(got, exp, fluid) < 1e-14, (kernel, steps)
        assert abs(oracle.density(dims, lat, got) - stats[0]) < 1e-14


@pytest.mark.parametrize("kernel", KERNELS)
def test_verification_norm_matches_reference(oracle, ref, kernel):
    """the -V case (cli.c:84-99): reference init, 1000 steps, kernelVerify"""
    dims = (16, 16, 16)
    lat, per, _ = oracle.geometry("box", dims, (1, 1, 0))
    exp, stats = ref.run(kernel, "box", dims, (1, 1, 0), None, 1000, 1.0, 1e-5, ref_init=True)
    # kernelSetInitialDensity/Velocity at omega 1, rho 1, u 0: f_i = w_i
    w = np.array([1 / 18] * 4 + [1 / 36] * 4 + [1 / 18] + [1 / 36] * 4 + [1 / 18] + [1 / 36] * 4 + [1 / 3])
    init = np.broadcast_to(w, dims + (19,)).copy()
    got, _ = oracle.run(kernel, dims, lat, per, init, 1000, 1.0, 1 / 12, 1e-5, True)
    fluid = lat.transpose(2, 1, 0) != 0
    assert rel_err(got, exp, fluid) < 1e-13
    norm = oracle.verify(dims, lat, got, 1.0, 1e-5, True)
    assert abs(norm - stats[1]) < 1e-10 * stats[1]



@pytest.mark.parametrize("geo,dims,per", [("channel", (12, 11, 10), (1, 0, 1)),
                                          ("box", (9, 10, 11), (1, 0, 0))])
@pytest.mark.parametrize("kernel", KERNELS)
def test_zou_he_matches_reference(oracle, ref, geo, dims, per, kernel):
    """kernelComputeBoundaryConditions (src/lbm.c:676-785) after runs of
    0, 1, 2 and 5 steps (both AA parities; each variant reads and writes the
    boundary populations through its own accessors, src/lbm.c:296-396)"""
    lat, per_o, _ = oracle.geometry(geo, dims, per)
    init = oracle.init_random(dims, lat, 21, 1e-2)
    fluid = lat.transpose(2, 1, 0) != 0
    marked = mark_inlet_outlet(lat)
    for steps in (0, 1, 2, 5):
        run, _ = oracle.run(kernel, dims, lat, per_o, init, steps, 1.25, 1 / 12, 1e-4, True)
        got = oracle.zou_he(dims, marked, run, 1.01, 0.99)
        exp = ref.zou_he(kernel, geo, dims, per, init, steps, 1.25, 1e-4, 1.01, 0.99)
        assert rel_err(got, exp, fluid) < 1e-14, (kernel, steps)
        assert not np.array_equal(got, run)  # the boundary pass changed something


TWO_LATTICE = [k for k in KERNELS if k != "aa-soa"]


@pytest.mark.parametrize("geo,dims,per", [("channel", (12, 11, 10), (1, 0, 1)),
                                          ("box", (9, 10, 11), (1, 0, 0))])
@pytest.mark.parametrize("kernel", TWO_LATTICE)
def test_per_step_zou_he_matches_reference(oracle, ref, geo, dims, per, kernel):
    """oracle_run_bc (Zou-He after every step) against the reference's own
    Kernel called one step at a time, each followed by its
    kernelComputeBoundaryConditions (src/lbm.c:676-785), 20 steps with the
    body force"""
    lat, per_o, _ = oracle.geometry(geo, dims, per)
    init = oracle.init_random(dims, lat, 21, 1e-2)
    fluid = lat.transpose(2, 1, 0) != 0
    marked = mark_inlet_outlet(lat)
    got, done = oracle.run_bc(kernel, dims, marked, per_o, init, 20, 1.25, 1 / 12, 1e-4, True,
                              1.01, 0.99)
    exp = ref.run_bc(kernel, geo, dims, per, init, 20, 1.25, 1e-4, 1.01, 0.99)
    assert done == 20
    assert rel_err(got, exp, fluid) < 1e-14
    plain, _ = oracle.run(kernel, dims, lat, per_o, init, 20, 1.25, 1 / 12, 1e-4, True)
    assert rel_err(got, plain, fluid) > 1e-6  # the boundaries changed the flow
