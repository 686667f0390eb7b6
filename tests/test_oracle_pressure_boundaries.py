"""The oracle's pressure boundaries inside the time loop (oracle_run_bc:
Zou-He, src/lbm.c:676-785, after every step), pinned by what the
construction fixes:
  - the five populations Zou-He writes make the density of every inlet
    cell exactly rhoIn and of every outlet cell rhoOut (the known 14 plus
    the five new ones sum to rho by construction), to round-off;
  - every variant gives the same logical state (aa-soa reads and writes
    the boundary populations through its parity-dependent accessors,
    src/lbm.c:262-294, the others through theirs), which also pins the AA
    path the reference cannot call one step at a time;
  - rhoIn = rhoOut = 1 on a rest state leaves it at rest."""
import numpy as np
import pytest
from conftest import mark_inlet_outlet, rel_err

from oracle import KERNELS

DIMS, PER = (14, 11, 10), (1, 0, 1)


@pytest.fixture(scope="module")
def case(oracle):
    lat, per, _ = oracle.geometry("channel", DIMS, PER)
    return lat, per, mark_inlet_outlet(lat), oracle.init_random(DIMS, lat, 3, 1e-2)


@pytest.mark.parametrize("kernel", KERNELS)
def test_boundary_densities(oracle, case, kernel):
    lat, per, marked, init = case
    out, done = oracle.run_bc(kernel, DIMS, marked, per, init, 10, 1.3, 3 / 16, 0.0, False,
                              1.02, 0.98)
    assert done == 10
    rho = out.sum(-1)
    inlet = marked.transpose(2, 1, 0) == 2
    outlet = marked.transpose(2, 1, 0) == 4
    assert inlet.sum() == 9 * 8 and outlet.sum() == 9 * 8
    assert np.abs(rho[inlet] - 1.02).max() < 1e-14
    assert np.abs(rho[outlet] - 0.98).max() < 1e-14


def test_variants_agree(oracle, case):
    lat, per, marked, init = case
    fluid = lat.transpose(2, 1, 0) != 0
    base, _ = oracle.run_bc("pull-soa", DIMS, marked, per, init, 12, 1.3, 3 / 16, 1e-4, True,
                            1.01, 0.99)
    for kernel in KERNELS:
        got, _ = oracle.run_bc(kernel, DIMS, marked, per, init, 12, 1.3, 3 / 16, 1e-4, True,
                               1.01, 0.99)
        assert rel_err(got, base, fluid) < 1e-15, kernel


def test_rest_state_stays_at_rest(oracle, case):
    lat, per, marked, _ = case
    w = np.array([1 / 18] * 4 + [1 / 36] * 4 + [1 / 18] + [1 / 36] * 4 + [1 / 18] + [1 / 36] * 4
                 + [1 / 3])
    fluid = lat.transpose(2, 1, 0) != 0
    rest = np.where(fluid[..., None], np.broadcast_to(w, DIMS + (19,)), 0.0)
    for kernel in ("pull-soa", "aa-soa"):
        got, _ = oracle.run_bc(kernel, DIMS, marked, per, rest, 6, 1.3, 3 / 16, 0.0, False, 1.0,
                               1.0)
        assert np.abs(got - rest)[fluid].max() < 1e-15, kernel
