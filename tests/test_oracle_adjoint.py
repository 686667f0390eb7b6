"""Pins of the oracle's differentiation form (no reference counterpart):
  - at s = 0 the Brinkman model is the reference's push step (the oracle's
    halo-and-list restatement, itself pinned to LBM-Bench);
  - the adjoint gradient equals central finite differences of J;
  - the porous field equals scipy's periodic Gaussian filter path."""
import numpy as np
import pytest


def _case(oracle, dims=(6, 8, 7), seed=3):
    lat, per, _ = oracle.geometry("channel", dims, (0, 0, 1))
    init = oracle.init_random(dims, lat, seed, 2e-2)
    return lat, per, init


def test_brinkman_at_zero_is_push_step(oracle):
    dims = (6, 8, 7)
    lat, per, init = _case(oracle, dims)
    s = np.zeros(dims)
    for T in (1, 5):
        _, final, _ = oracle.brinkman(dims, lat, s, init, T, 1.3, 3 / 16, 1e-4, dims[0] - 1)
        exp, _ = oracle.run("push-soa", dims, lat, per, init, T, 1.3, 3 / 16, 1e-4, True)
        fluid = lat.transpose(2, 1, 0) != 0
        assert np.abs(final - exp)[fluid].max() < 1e-15


def test_brinkman_drag_slows_the_flow(oracle):
    dims = (6, 8, 7)
    lat, per, init = _case(oracle, dims)
    J0, _, _ = oracle.brinkman(dims, lat, np.zeros(dims), init, 30, 1.3, 3 / 16, 1e-3, 5)
    J1, _, _ = oracle.brinkman(dims, lat, np.full(dims, 0.5), init, 30, 1.3, 3 / 16, 1e-3, 5)
    assert J1 < J0


@pytest.mark.parametrize("omega,magic", [(1.3, 3 / 16), (0.9, 1 / 12)])
def test_adjoint_matches_finite_differences(oracle, omega, magic):
    dims = (6, 8, 7)
    lat, per, init = _case(oracle, dims)
    s = oracle.porous_field(dims, 5, 1.0) * 0.6
    T, outlet, g = 12, 4, 1e-3
    J, _, grad = oracle.brinkman(dims, lat, s, init, T, omega, magic, g, outlet, gradient=True)
    fluid = lat.transpose(2, 1, 0) != 0
    rng = np.random.default_rng(0)
    cells = [tuple(c) for c in np.argwhere(fluid)[rng.choice(fluid.sum(), 12, replace=False)]]
    h = 1e-6
    for c in cells:
        sp, sm = s.copy(), s.copy()
        sp[c] += h
        sm[c] -= h
        Jp, _, _ = oracle.brinkman(dims, lat, sp, init, T, omega, magic, g, outlet)
        Jm, _, _ = oracle.brinkman(dims, lat, sm, init, T, omega, magic, g, outlet)
        fd = (Jp - Jm) / (2 * h)
        assert abs(fd - grad[c]) <= 1e-6 * np.abs(grad).max(), (c, fd, grad[c])
    assert np.abs(grad[~fluid]).max() == 0.0


def test_porous_field_matches_scipy(oracle):
    from scipy.ndimage import gaussian_filter
    dims, seed, sigma = (20, 24, 16), 7, 3.0
    u = np.array([(oracle.hash(seed, i, 0) >> 11) / 2.0 ** 53
                  for i in range(int(np.prod(dims)))]).reshape(dims)
    sm = gaussian_filter(u, sigma, mode="wrap", truncate=3.0)
    z = (sm - sm.mean()) / sm.std()
    exp = 1.0 / (1.0 + np.exp(-(4.0 * z - 2.0)))
    got = oracle.porous_field(dims, seed, sigma)
    assert np.abs(got - exp).max() < 1e-12
