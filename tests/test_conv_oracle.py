"""Pins of the FP64 convolution oracle (oracle/conv_oracle.py): the backward
formulas against central finite differences of the forward, delta kernels as
periodic shifts, and BF16 rounding against torch's round-to-nearest-even."""
import numpy as np
import torch

from oracle import conv_oracle as co


def test_delta_kernel_is_periodic_shift():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((1, 5, 6, 2))
    for t in range(9):
        w = np.zeros((9, 2, 2))
        w[t] = np.eye(2)
        y = co.conv3x3(x, w)
        dr, ds = t // 3 - 1, t % 3 - 1
        for h in range(5):
            for ww in range(6):
                assert np.allclose(y[0, h, ww], x[0, (h + dr) % 5, (ww + ds) % 6])


def test_backward_formulas_match_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 4, 5, 3))
    w = rng.standard_normal((9, 2, 3))
    g = rng.standard_normal((2, 4, 5, 2))  # upstream gradient

    def loss(xx, ww):
        return float((co.conv3x3(xx, ww) * g).sum())

    dx = co.conv3x3_bwd_data(g, w)
    dw = co.conv3x3_bwd_weight(x, g)
    eps = 1e-6
    for idx in [(0, 0, 0, 0), (1, 3, 4, 2), (0, 2, 1, 1), (1, 0, 4, 0)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += eps
        xm[idx] -= eps
        fd = (loss(xp, w) - loss(xm, w)) / (2 * eps)
        assert abs(fd - dx[idx]) < 1e-7
    for idx in [(0, 0, 0), (4, 1, 2), (8, 0, 1), (3, 1, 0)]:
        wp, wm = w.copy(), w.copy()
        wp[idx] += eps
        wm[idx] -= eps
        fd = (loss(x, wp) - loss(x, wm)) / (2 * eps)
        assert abs(fd - dw[idx]) < 1e-7


def test_silu_grad_and_pool_adjoints():
    v = np.linspace(-6, 6, 101)
    eps = 1e-6
    fd = (co.silu(v + eps) - co.silu(v - eps)) / (2 * eps)
    assert np.allclose(fd, co.silu_grad(v), atol=1e-8)
    rng = np.random.default_rng(2)
    a = rng.standard_normal((1, 4, 6, 3))
    b = rng.standard_normal((1, 2, 3, 3))
    # <pool_avg(a), b> = <a, upsample(b)/4>;  <pool_sum(a), b> = <a, upsample(b)>
    assert np.isclose((co.pool2(a) * b).sum(), (a * co.upsample2(b) / 4).sum())
    assert np.isclose((co.pool2(a, "sum") * b).sum(), (a * co.upsample2(b)).sum())


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    a = (rng.standard_normal(10000) * 10.0 ** rng.integers(-5, 5, 10000)).astype(np.float32)
    a[:4] = [1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -(1.0 + 2 ** -8), 65504.0]  # ties
    want = torch.from_numpy(a).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(co.bf16_round(a), want)
    bits = co.to_bf16_bits(a)
    assert np.array_equal(co.from_bf16_bits(bits), want)


def test_unet_roofline_counts():
    """The per-kernel cost table behind the surrogate's combined roofline:
    its FLOPs add up to the network's FLOP count, the 32-channel 256^2 layer
    moves one BF16 input and one BF16 output, and with unlimited bandwidth the
    bound is the FLOP count over the MFMA peak."""
    import math

    from symphas_amd.nn import UNet

    net = UNet(seed=1)
    costs = {name: (f, b) for name, f, b in net.layer_costs(256, 256)}
    assert sum(f for f, _ in costs.values()) == net.flops_per_sample(256, 256)
    assert costs["enc0b"] == (2 * 9 * 32 * 32 * 256 * 256, 2 * (2 * 256 * 256 * 32))
    # upsample-add at the finest level: 64-channel 128^2 input, 32-channel
    # 256^2 skip and output
    assert costs["up0"][1] == 2 * 128 * 128 * 64 + 2 * 2 * 256 * 256 * 32
    t_flop = net.roofline_seconds(256, 256, 64, peak_bytes=math.inf)
    assert abs(t_flop / (net.flops_per_sample(256, 256) * 64 / 2.5e15) - 1) < 1e-12
    t = net.roofline_seconds(256, 256, 64)
    assert t > t_flop
    assert t >= sum(b for _, b in costs.values()) * 64 / 8e12
