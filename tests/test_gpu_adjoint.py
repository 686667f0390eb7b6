"""The differentiation form on the GPU (lbmAdjoint* through the C-ABI):
  - the Brinkman forward step equals the oracle (and at s = 0 the
    reference's push step) to 1e-12;
  - the adjoint gradient equals the oracle's FP64 adjoint (dual-number
    Jacobians) for every checkpoint layout;
  - 128^3 porous channel, T = 1024, sqrt(T) checkpoints: dJ/ds matches
    central finite differences (h = 1e-6) to 1e-4 on every one of 64 seeded
    cells where |dJ/ds| stands above the FP64 round-off of J (and within a
    stated round-off floor on 64 cells from anywhere);
  - a torch.autograd.Function around the C-ABI passes gradcheck;
  - an adjoint step costs at most 2.5 forward steps."""
import numpy as np
import pytest

from lbmhip import Adjoint, porous_field

pytestmark = pytest.mark.gpu

KW = dict(omega=1.25, magic=3 / 16)


def _field(oracle, dims, seed, sigma=1.0, scale=0.6):
    return oracle.porous_field(dims, seed, sigma) * scale


def test_porous_field_matches_oracle(gpu_lib, oracle):
    for dims, sigma in (((16, 12, 20), 3.0), ((9, 7, 5), 0.0)):
        got = porous_field(gpu_lib, dims, 11, sigma)
        assert np.abs(got - oracle.porous_field(dims, 11, sigma)).max() < 1e-13


@pytest.mark.parametrize("dims", [(8, 10, 12), (6, 9, 70)])
def test_forward_matches_oracle(gpu_lib, oracle, dims):
    lat, per, _ = oracle.geometry("channel", dims, (0, 0, 1))
    init = oracle.init_random(dims, lat, 5, 2e-2)
    fluid = lat.transpose(2, 1, 0) != 0
    for s in (np.zeros(dims), _field(oracle, dims, 2)):
        with Adjoint(gpu_lib, "channel", dims, (0, 0, 1), xforce=1e-4, **KW) as A:
            A.set_field(s)
            A.set_state(init)
            J, _ = A.forward(9)
            got = A.state()
        Jo, exp, _ = oracle.brinkman(dims, lat, s, init, 9, 1.25, 3 / 16, 1e-4, dims[0] - 1)
        assert np.abs(got - exp)[fluid].max() < 1e-12 * np.abs(exp[fluid]).max()
        assert abs(J - Jo) <= 1e-12 * abs(Jo)
        if not s.any():
            push, _ = oracle.run("push-soa", dims, lat, per, init, 9, 1.25, 3 / 16, 1e-4, True)
            assert np.abs(got - push)[fluid].max() < 1e-12 * np.abs(push[fluid]).max()


@pytest.mark.parametrize("intervals", [[], [5], [7, 2], [24], [1]])
def test_gradient_matches_oracle(gpu_lib, oracle, intervals):
    dims, T = (8, 10, 70), 24
    lat, per, _ = oracle.geometry("channel", dims, (0, 0, 1))
    init = oracle.init_random(dims, lat, 9, 2e-2)
    s = _field(oracle, dims, 4)
    Jo, _, go = oracle.brinkman(dims, lat, s, init, T, 1.25, 3 / 16, 1e-3, 5, gradient=True)
    with Adjoint(gpu_lib, "channel", dims, (0, 0, 1), xforce=1e-3, outlet_x=5, **KW) as A:
        A.set_field(s)
        A.set_state(init)
        J, g, stats = A.gradient(T, intervals)
    assert abs(J - Jo) <= 1e-12 * abs(Jo)
    assert np.abs(g - go).max() <= 1e-10 * np.abs(go).max()


@pytest.mark.parametrize("storage,bound", [("fp32", 1e-3), ("fp16", 2e-2)])
def test_low_precision_gradient(gpu_lib, oracle, storage, bound):
    """FP32 / FP16 forward with an FP32 adjoint against the FP64 oracle"""
    dims, T = (8, 10, 64), 24
    lat, per, _ = oracle.geometry("channel", dims, (0, 0, 1))
    s = _field(oracle, dims, 4)
    init = oracle.init_random(dims, lat, 9, 2e-2)
    _, _, go = oracle.brinkman(dims, lat, s, init, T, 1.25, 3 / 16, 1e-3, 5, gradient=True)
    with Adjoint(gpu_lib, "channel", dims, (0, 0, 1), xforce=1e-3, outlet_x=5, storage=storage,
                 **KW) as A:
        A.set_field(s)
        A.set_state(init)
        _, g, _ = A.gradient(T, [5])
    err = np.linalg.norm(g - go) / np.linalg.norm(go)
    print(f"{storage}: gradient L2 error vs FP64 {err:.3e}")
    assert err <= bound


def test_porous_channel_128_finite_differences(gpu_lib):
    """128^3 porous channel (walls at y, periodic x and z), s from U[0,1]
    noise smoothed with sigma = 3 and a logistic, body force 1e-5, rest
    start, T = 1024, J at the outlet plane x = 127; sqrt(T) checkpoints.

    A central difference at h = 1e-6 carries the FP64 round-off of two
    1024-step runs: measured on MI355X, |FD(1e-6) - FD(1e-4)| is 1.5e-12
    (median) and 5.4e-12 (max) over 64 cells, about 7e-15 J / h, whatever
    the cell.  Most cells of the lattice have |dJ/ds| below 1e-8 (the
    gradient concentrates upstream of the outlet plane), where that noise
    alone exceeds 1e-4 relative.  So:
      - 64 seeded cells drawn from those with |g| >= 5 % of max |g|
        (above 2e-7, 40x the largest noise over the tolerance): at
        h = 1e-6, every cell |FD - g| <= 1e-4 |g|, no floor;
      - 64 seeded cells drawn from all fluid cells: every cell
        |FD - g| <= 1e-4 |g| + floor, with floor 1e-11 at h = 1e-6 (twice
        the largest measured noise) and 1e-13 at h = 1e-4 (noise 100x
        smaller; the O(h^2) truncation is below 1e-6 relative)."""
    dims, T = (128, 128, 128), 1024
    s = porous_field(gpu_lib, dims, 0, 3.0)
    with Adjoint(gpu_lib, "channel", dims, (0, 0, 1), xforce=1e-5, **KW) as A:
        fluid = A.fluid_mask()
        A.set_field(s)
        A.init("rest", 0, 1.0, 0.0)
        J, g, stats = A.gradient(T, [32])
        J0, _ = A.forward(T)
        assert abs(J0 - J) <= 1e-13 * abs(J)
        rng = np.random.default_rng(64)
        anywhere = np.argwhere(fluid)[rng.choice(int(fluid.sum()), 64, replace=False)]
        strong_set = np.argwhere(fluid & (np.abs(g) >= 0.05 * np.abs(g).max()))
        strong = strong_set[rng.choice(len(strong_set), 64, replace=False)]

        def fd(c, h):
            keep = s[c]
            s[c] = keep + h
            A.set_field(s)
            Jp, _ = A.forward(T)
            s[c] = keep - h
            A.set_field(s)
            Jm, _ = A.forward(T)
            s[c] = keep
            return (Jp - Jm) / (2 * h)

        checks = []
        for c in map(tuple, strong):
            checks.append(("strong", 1e-6, 0.0, c, fd(c, 1e-6)))
        for c in map(tuple, anywhere):
            checks.append(("anywhere", 1e-6, 1e-11, c, fd(c, 1e-6)))
            checks.append(("anywhere", 1e-4, 1e-13, c, fd(c, 1e-4)))
        A.set_field(s)
    print(f"J = {J:.6e}, max |g| {np.abs(g).max():.3e}, {len(strong_set)} cells with |g| >= 5 % of "
          f"max; forward {stats['forward_s']:.3f} s, recompute {stats['recompute_s']:.3f} s, adjoint "
          f"{stats['adjoint_s']:.3f} s, device {stats['device_bytes'] / 2**30:.1f} GiB")
    for group, h, floor in (("strong", 1e-6, 0.0), ("anywhere", 1e-6, 1e-11),
                            ("anywhere", 1e-4, 1e-13)):
        rows = [(abs(v - g[c]), abs(g[c])) for gr, hh, _, c, v in checks if gr == group and hh == h]
        rel = [e / a for e, a in rows]
        over = [e / (1e-4 * a + floor) for e, a in rows]
        print(f"{group} cells, h = {h:g}: relative error max {max(rel):.2e}, median "
              f"{np.median(rel):.2e}; largest error / bound {max(over):.2f}")
        assert max(over) <= 1.0, (group, h)


def test_autograd_function_gradcheck(gpu_lib):
    import torch

    dims, T = (6, 8, 6), 10
    A = Adjoint(gpu_lib, "channel", dims, (0, 0, 1), xforce=1e-3, outlet_x=3, **KW)
    A.init("random", 1, 1.0, 2e-2)

    class OutletEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, s):
            ctx.save_for_backward(s)
            A.set_field(s.detach().numpy())
            J, _ = A.forward(T)
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            (s,) = ctx.saved_tensors
            A.set_field(s.detach().numpy())
            _, g, _ = A.gradient(T, [3])
            return gJ * torch.from_numpy(g)

    try:
        s = torch.from_numpy(np.full(dims, 0.3)).requires_grad_(True)
        assert torch.autograd.gradcheck(OutletEnergy.apply, (s,), eps=1e-6, atol=1e-11, rtol=1e-5)
    finally:
        A.close()


def test_adjoint_step_cost(gpu_lib):
    dims = (256, 256, 256)
    with Adjoint(gpu_lib, "channel", dims, (0, 0, 1), xforce=1e-5, **KW) as A:
        A.init("rest", 0, 1.0, 0.0)
        A.set_field(np.full(dims, 0.1))
        A.step_times(3)
        t = A.step_times(20)
    ratio = t["adjoint_s"] / t["forward_s"]
    print(f"forward step {t['forward_s'] * 1e3:.3f} ms, adjoint step {t['adjoint_s'] * 1e3:.3f} ms, "
          f"ratio {ratio:.2f}")
    assert ratio <= 2.5


# measured on MI355X: J 1.0e-5 / 9.6e-7, gradient 5.3e-5 / 9.1e-6 (FP16 / FP32)
@pytest.mark.parametrize("storage,j_bound,g_bound", [("fp16", 5e-5, 2e-4), ("fp32", 5e-6, 5e-5)])
def test_long_horizon_gradient_against_fp64(gpu_lib, storage, j_bound, g_bound):
    """The large differentiation case checked on a 128^3 box: s from U[0,1]
    noise smoothed with sigma = 5, T = 4096, two checkpoint levels (every
    64 steps, then every 8); the gradient from FP16 (or FP32) forward
    storage with an FP32 adjoint against the FP64 gradient"""
    dims, T, levels = (128, 128, 128), 4096, [64, 8]
    s = porous_field(gpu_lib, dims, 0, 5.0)
    out = {}
    for form in ("native", storage):
        with Adjoint(gpu_lib, "channel", dims, (0, 0, 1), xforce=1e-5, storage=form, **KW) as A:
            A.set_field(s)
            A.init("rest", 0, 1.0, 0.0)
            J, g, stats = A.gradient(T, levels)
            out[form] = (J, g)
    J64, g64 = out["native"]
    J, g = out[storage]
    err = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    print(f"{storage}: T {T}, J rel diff {abs(J - J64) / abs(J64):.3e}, gradient L2 error {err:.3e}")
    assert abs(J - J64) <= j_bound * abs(J64)
    assert err <= g_bound


@pytest.mark.timeout(900)
def test_combined_at_scale_384(gpu_lib):
    """The combined case at its full size: a 384^3 porous channel (s from
    U[0,1] noise smoothed with sigma = 5 and a logistic), body force 1e-5,
    rest start, T = 4096, FP16 forward storage with an FP32 adjoint and two
    checkpoint levels (every 64 steps, then every 8).  The peak device
    footprint fits the device, J of the FP16 forward lies within 5e-5 of an
    FP64 forward run of the same case (measured 4.5e-6), and the gradient is
    finite and non-zero on the fluid cells.  Its accuracy against an FP64
    gradient is checked on a 128^3 box (test_long_horizon_gradient_against_fp64):
    the FP64 gradient at 384^3 would need 77 FP64 states of 8.6 GB."""
    dims, T, levels = (384, 384, 384), 4096, [64, 8]
    _, _, device_total = gpu_lib.device_info()
    s = porous_field(gpu_lib, dims, 0, 5.0)
    with Adjoint(gpu_lib, "channel", dims, (0, 0, 1), xforce=1e-5, **KW) as A:
        A.set_field(s)
        A.init("rest", 0, 1.0, 0.0)
        J64, t64 = A.forward(T)
    print(f"fp64 forward: J {J64:.9e}, {t64:.2f} s", flush=True)
    with Adjoint(gpu_lib, "channel", dims, (0, 0, 1), xforce=1e-5, storage="fp16", **KW) as A:
        fluid = A.fluid_mask()
        A.set_field(s)
        A.init("rest", 0, 1.0, 0.0)
        J, g, stats = A.gradient(T, levels)
    nf = int(fluid.sum())
    total = stats["forward_s"] + stats["recompute_s"] + stats["adjoint_s"]
    print(f"fp16/fp32: J {J:.9e} (rel diff vs fp64 {abs(J - J64) / abs(J64):.2e}), peak device "
          f"{stats['device_bytes'] / 2**30:.1f} GiB of {device_total / 2**30:.0f}, gradient "
          f"{total:.1f} s = {nf * T / total / 1e6:.0f} MLUPS (forward {stats['forward_s']:.1f} s, "
          f"recompute {stats['recompute_s']:.1f} s, adjoint {stats['adjoint_s']:.1f} s); FP64 "
          f"forward {nf * T / t64 / 1e6:.0f} MLUPS", flush=True)
    assert 0 < stats["device_bytes"] < device_total
    assert abs(J - J64) <= 5e-5 * abs(J64)
    assert np.isfinite(g).all()
    assert np.abs(g[fluid]).max() > 0
    assert (g[~fluid] == 0).all()
