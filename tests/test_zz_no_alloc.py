"""lbmCompareMomentumNoAlloc (no device allocation, partial sums in the
library's static scratch buffer) against lbmCompareMomentum, through the
C-ABI: two calls in a row, the second on smaller lattices, to 1e-12
relative on the sums and exactly on the maximum; refusals of lattices of
other dimensions and of more cells than the scratch buffer covers."""
import numpy as np
import pytest

from lbmhip import Lattice

pytestmark = pytest.mark.gpu

KW = dict(omega=1.3, xforce=1e-5, body_force=True, magic=3 / 16, seed=4)


def _pair(lib, kernel, geo, dims, steps, form):
    A = Lattice(lib, kernel, geo, dims, (0, 0, 1), storage=form, **KW)
    B = Lattice(lib, "aa-soa", geo, dims, (0, 0, 1), **KW)
    for L in (A, B):
        L.init("random", 7, 1.0, 1e-3)
        L.run(steps)
    return A, B


def _agree(got, exp):
    for k in (0, 1, 3):
        assert abs(got[k] - exp[k]) <= 1e-12 * abs(exp[k]), (k, got[k], exp[k])
    assert got[2] == exp[2]


@pytest.mark.parametrize("kernel,form", [("pull-soa", "fp16"), ("push-aos", "native"),
                                         ("aa-soa", "fp32")])
def test_no_alloc_matches_allocating_entry_point(gpu_lib, kernel, form):
    # even step counts: aa-soa runs its steps in pairs (src/kernel_aa_soa.c:36)
    for dims, steps in (((48, 40, 128), 10), ((20, 18, 70), 4)):   # the second call is smaller
        A, B = _pair(gpu_lib, kernel, "spheres-12", dims, steps, form)
        try:
            exp = A.compare_momentum(B)
            got = A.compare_momentum_noalloc(B)
            print(f"{kernel} {form} {dims}: {got}")
            assert np.isfinite(got).all() and got[1] > 0
            if form == "native":
                assert got[0] < 1e-24 * got[1]        # the same flow, two layouts
            else:
                assert 0 < got[0] < 1e-4 * got[1]
            # rho of the deviation forms includes RhoRef: the density
            # differences are those of the flows, not of the reference density
            assert got[3] < 1e-6 * A.n_fluid
            _agree(got, exp)
        finally:
            A.close()
            B.close()


def test_no_alloc_refusals(gpu_lib):
    with Lattice(gpu_lib, "aa-soa", "channel", (8, 10, 64), (0, 0, 1), **KW) as A, \
         Lattice(gpu_lib, "aa-soa", "channel", (8, 10, 70), (0, 0, 1), **KW) as B:
        with pytest.raises(ValueError):
            A.compare_momentum_noalloc(B)              # other dimensions
    # 2049 * 2048 * 64 cells, one more row than the 2^28 the buffer covers
    # (FP16: 10.2 GB of PDFs); refused on the host before any launch
    with Lattice(gpu_lib, "aa-soa", "fluid", (2049, 2048, 64), (1, 1, 1), storage="fp16") as L:
        with pytest.raises(ValueError):
            L.compare_momentum_noalloc(L)
