"""BF16 MFMA convolution kernels against the FP64 oracle (oracle/conv_oracle.py)
on BF16-rounded inputs.  BF16 outputs must be within one BF16 ulp of the FP64
result (the kernels accumulate in FP32 and round once); FP32 weight gradients
within the FP32 accumulation tolerance."""
import numpy as np
import pytest
import torch

from oracle import conv_oracle as co
from symphas_amd import nn

pytestmark = pytest.mark.gpu
DEV = "cuda"


def bf(a):
    return torch.from_numpy(co.bf16_round(a).astype(np.float32)).to(DEV).to(torch.bfloat16)


def np64(t):
    return t.float().cpu().numpy().astype(np.float64)


def within_ulp(got, want, ulps=1.0, floor=1e-30):
    ulp = np.maximum(np.abs(want), floor) * 2.0 ** -7
    return np.abs(got - want) <= ulps * ulp + 1e-6


def rnd(shape, seed, scale=1.0):
    return np.random.default_rng(seed).standard_normal(shape) * scale


@pytest.mark.parametrize("N,H,W,C,K", [(2, 16, 16, 32, 64), (2, 16, 32, 32, 32),
                                       (1, 8, 32, 64, 32), (2, 16, 32, 64, 32),
                                       (4, 32, 32, 128, 256), (1, 16, 64, 256, 128)])
def test_conv_forward_plain(gpu_lib, N, H, W, C, K):
    print(f"[case] conv fwd {N}x{H}x{W} {C}->{K}", flush=True)
    x = co.bf16_round(rnd((N, H, W, C), 1))
    w = co.bf16_round(rnd((9, K, C), 2, 1.0 / np.sqrt(9 * C)))
    b = rnd(K, 3, 0.1).astype(np.float32)
    xt, wt = bf(x), bf(w)
    bt = torch.from_numpy(b).to(DEV)
    y, pre = nn.conv3x3_raw(xt, wt, bt, nn.ACT_SILU, want_pre=True)
    torch.cuda.synchronize()
    ref = co.conv3x3(x, w, b.astype(np.float64))
    assert within_ulp(np64(pre), ref).all()
    assert within_ulp(np64(y), co.silu(co.bf16_round(ref)), ulps=2).mean() > 0.999
    y0, _ = nn.conv3x3_raw(xt, wt, None, nn.ACT_NONE)
    assert within_ulp(np64(y0), co.conv3x3(x, w)).all()


@pytest.mark.parametrize("H,C", [(8, 64), (16, 64), (16, 128)])
def test_conv_forward_upsample_add(gpu_lib, H, C):
    print(f"[case] conv fwd upsample-add H={H} C={C}", flush=True)
    N, W, K = 2, 16, 32
    x = co.bf16_round(rnd((N, H, W, C), 4))
    w = co.bf16_round(rnd((9, K, C), 5, 1.0 / np.sqrt(9 * C)))
    b = rnd(K, 6, 0.1).astype(np.float32)
    skip = co.bf16_round(rnd((N, 2 * H, 2 * W, K), 7))
    y, pre = nn.conv3x3_raw(bf(x), bf(w), torch.from_numpy(b).to(DEV), nn.ACT_SILU,
                            nn.EPI_UPADD, bf(skip), want_pre=True)
    # the kernel adds the FP32 accumulator to the skip before rounding
    ref = co.upsample2(co.conv3x3(x, w, b.astype(np.float64))) + skip
    assert within_ulp(np64(pre), ref, ulps=1.01).all()


@pytest.mark.parametrize("N,H,W,C,K", [(2, 16, 32, 32, 64), (1, 32, 32, 128, 64)])
def test_conv_backward_data(gpu_lib, N, H, W, C, K):
    print(f"[case] conv bwd data {C}->{K}", flush=True)
    w = co.bf16_round(rnd((9, K, C), 8, 0.2))
    dy = co.bf16_round(rnd((N, H, W, K), 9))
    wb = nn.flip_weights(torch.from_numpy(w).float()).to(DEV).to(torch.bfloat16)
    dx, _ = nn.conv3x3_raw(bf(dy), wb, None, nn.ACT_NONE)
    assert within_ulp(np64(dx), co.conv3x3_bwd_data(dy, w)).all()


@pytest.mark.parametrize("N,H,W,C,K", [(2, 8, 32, 32, 64), (1, 16, 64, 64, 32),
                                       (3, 4, 32, 64, 64), (1, 8, 32, 32, 32),
                                       (2, 8, 8, 64, 64), (1, 4, 16, 32, 64)])
def test_conv_backward_weight(gpu_lib, N, H, W, C, K):
    print(f"[case] conv wgrad {C}->{K}", flush=True)
    x = co.bf16_round(rnd((N, H, W, C), 10))
    dy = co.bf16_round(rnd((N, H, W, K), 11))
    dw = nn.conv3x3_wgrad_raw(bf(x), bf(dy), K)
    ref = co.conv3x3_bwd_weight(x, dy)
    got = dw.cpu().numpy().astype(np.float64)
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max() * np.sqrt(N * H * W)


@pytest.mark.parametrize("N,H,W,K", [(2, 16, 16, 32), (3, 12, 20, 64), (1, 9, 7, 8),
                                     (2, 32, 48, 64)])
def test_thin_layers(gpu_lib, N, H, W, K):
    # 720 and 63 pixels leave a partial 256-pixel chunk in the weight gradients
    print(f"[case] thin layers {N}x{H}x{W} K={K}", flush=True)
    lib = nn._lib()
    x = rnd((N, H, W), 12).astype(np.float32)
    w = rnd((9, K), 13, 0.3).astype(np.float32)
    b = rnd(K, 14, 0.1).astype(np.float32)
    xt, wt, bt = (torch.from_numpy(a).to(DEV) for a in (x, w, b))
    y = torch.empty((N, H, W, K), device=DEV, dtype=torch.bfloat16)
    pre = torch.empty_like(y)
    nn._check(lib.sps_nn_conv3x3_in1(nn._ptr(xt), N, H, W, nn._ptr(wt), nn._ptr(bt), K, 1,
                                     nn._ptr(y), nn._ptr(pre), nn._stream()))
    ref = co.conv3x3(x[..., None].astype(np.float64), w[:, :, None].astype(np.float64), b)
    assert within_ulp(np64(pre), ref).all()
    dpre = co.bf16_round(rnd((N, H, W, K), 15))
    dw = torch.zeros((9, K), device=DEV)
    ws = nn.wgrad_workspace(nn.WGRAD_IN1, N, H, W, 1, K, DEV)
    nn._check(lib.sps_nn_conv3x3_in1_wgrad(nn._ptr(xt), nn._ptr(bf(dpre)), N, H, W, K,
                                           nn._ptr(dw), nn._ptr(ws), ws.numel(), nn._stream()))
    refw = co.conv3x3_bwd_weight(x[..., None].astype(np.float64), dpre)[:, :, 0]
    assert np.allclose(dw.cpu().numpy(), refw, rtol=1e-4, atol=1e-4)
    # last layer
    h = co.bf16_round(rnd((N, H, W, K), 16))
    wo = rnd((9, K), 17, 0.2).astype(np.float32)
    bo = np.array([0.3], np.float32)
    out = torch.empty((N, H, W), device=DEV)
    wot, bot = torch.from_numpy(wo).to(DEV), torch.from_numpy(bo).to(DEV)
    nn._check(lib.sps_nn_conv3x3_out1(nn._ptr(bf(h)), N, H, W, K, nn._ptr(wot), nn._ptr(bot),
                                      nn._ptr(out), nn._stream()))
    refo = co.conv3x3(h, wo[:, None, :].astype(np.float64), bo.astype(np.float64))[..., 0]
    assert np.allclose(out.cpu().numpy(), refo, rtol=1e-5, atol=1e-4)
    dy = rnd((N, H, W), 18).astype(np.float32)
    dx = torch.empty((N, H, W, K), device=DEV, dtype=torch.bfloat16)
    dwo = torch.zeros((9, K), device=DEV)
    wso = nn.wgrad_workspace(nn.WGRAD_OUT1, N, H, W, K, 1, DEV)
    nn._check(lib.sps_nn_conv3x3_out1_bwd(nn._ptr(bf(h)), nn._ptr(torch.from_numpy(dy).to(DEV)),
                                          N, H, W, K, nn._ptr(wot), nn._ptr(dx), nn._ptr(dwo),
                                          nn._ptr(wso), wso.numel(), nn._stream()))
    refdx = co.conv3x3_bwd_data(dy[..., None].astype(np.float64), wo[:, None, :].astype(np.float64))
    assert within_ulp(np64(dx), refdx, ulps=1.01).all()
    refdw = co.conv3x3_bwd_weight(h, dy[..., None].astype(np.float64))[:, 0, :]
    assert np.allclose(dwo.cpu().numpy(), refdw, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("C", [32, 12])
def test_elementwise(gpu_lib, C):
    # 32 channels take the 8-channel kernels, 12 the one-value kernels
    print(f"[case] elementwise C={C}", flush=True)
    a = co.bf16_round(rnd((2, 8, 16, C), 19, 3.0))
    d = co.bf16_round(rnd((2, 8, 16, C), 20))
    dx = nn.silu_bwd_raw(bf(d), bf(a))
    assert within_ulp(np64(dx), d * co.silu_grad(a), ulps=2).all()
    p = nn.pool2_raw(bf(a), 0)
    assert within_ulp(np64(p), co.pool2(a), ulps=1.01).all()
    s = nn.pool2_raw(bf(a), 1)
    assert within_ulp(np64(s), co.pool2(a, "sum"), ulps=1.01).all()
    u = nn.unpool2_raw(bf(a), 0.25)
    assert np.array_equal(np64(u), co.upsample2(a) * 0.25)


def _unet_cpu_fp64(net, psi):
    """FP64 CPU restatement of UNet.forward with torch's conv2d (circular
    padding) and autograd: same weights rounded to BF16, activations not
    rounded.  Returns (output, parameters in the order of net.parameters())."""
    import torch.nn.functional as F

    def cw(w):  # [9][K][C] -> [K][C][3][3]
        w = w.detach().cpu().to(torch.bfloat16).double()
        return w.reshape(3, 3, w.shape[1], w.shape[2]).permute(2, 3, 0, 1).contiguous()

    def conv(x, w, b):
        return F.conv2d(F.pad(x, (1, 1, 1, 1), mode="circular"), w, b)

    params = {}
    for name, p in net.named_parameters():
        if name in ("w_in", "w_out"):
            t = p.detach().cpu().double()
        elif "_w." in name:
            t = cw(p)
        else:
            t = p.detach().cpu().double()
        params[name] = t.requires_grad_(True)
    L = len(net.ch)
    x = psi.cpu().double()[:, None]
    wi = params["w_in"].t().reshape(-1, 1, 3, 3)  # [9][K] -> [K][1][3][3]
    x = F.silu(conv(x, wi, params["b_in"]))
    skips = []
    for l in range(L):
        if l > 0:
            x = F.avg_pool2d(x, 2)
        for j in (0, 1):
            x = F.silu(conv(x, params[f"enc_w.{2 * l + j}"], params[f"enc_b.{2 * l + j}"]))
        skips.append(x)
    for i, l in enumerate(reversed(range(L - 1))):
        y = conv(x, params[f"dec_w.{2 * i}"], params[f"dec_b.{2 * i}"])
        x = F.silu(F.interpolate(y, scale_factor=2, mode="nearest") + skips[l])
        x = F.silu(conv(x, params[f"dec_w.{2 * i + 1}"], params[f"dec_b.{2 * i + 1}"]))
    wo = params["w_out"].t().reshape(1, -1, 3, 3)
    d = conv(x, wo, None)[:, 0]
    d = d - d.mean(dim=(1, 2), keepdim=True)
    return psi.cpu().double() + net.out_scale * d, params


def test_unet_forward_and_gradient_vs_cpu(gpu_lib):
    """Whole network: forward and the gradient of a loss against an FP64 CPU
    autograd restatement (torch conv2d, circular padding).  The GPU keeps
    activations and activation gradients in BF16, so agreement is at the
    percent level."""
    print("[case] unet forward/gradient", flush=True)
    torch.manual_seed(0)
    net = nn.UNet(seed=1).to(DEV)
    with torch.no_grad():
        net.w_out.copy_(torch.randn_like(net.w_out) * 0.1)
    psi = torch.from_numpy(rnd((2, 64, 64), 21, 0.5).astype(np.float32)).to(DEV)
    target = torch.from_numpy(rnd((2, 64, 64), 22, 0.5).astype(np.float32)).to(DEV)
    out = net(psi)
    loss = ((out - target) ** 2).mean()
    net.zero_grad()
    loss.backward()
    ref_out, params = _unet_cpu_fp64(net, psi)
    ref_loss = ((ref_out - target.cpu().double()) ** 2).mean()
    ref_loss.backward()
    inc = (out - psi).detach().cpu().double()
    ref_inc = (ref_out - psi.cpu().double()).detach()
    err = float((inc - ref_inc).norm() / ref_inc.norm())
    print(f"unet increment rel L2 vs FP64 CPU: {err:.2e}")
    assert err < 2e-2
    for name, p in net.named_parameters():
        g = p.grad.detach().cpu().double()
        r = params[name].grad
        if "_w." in name:  # back to [9][K][C]
            r = r.permute(2, 3, 0, 1).reshape(9, r.shape[0], r.shape[1])
        elif name in ("w_in", "w_out"):
            r = r.reshape(g.shape)
        rel = float((g - r).norm() / max(float(r.norm()), 1e-30))
        print(f"grad {name:10s} rel L2 {rel:.2e}")
        assert rel < 5e-2, name


def test_center_removes_each_samples_mean(gpu_lib):
    d = torch.from_numpy(rnd((3, 8, 16), 30).astype(np.float32)).to(DEV)
    out = nn.center_raw(d).cpu().numpy().astype(np.float64)
    ref = d.cpu().numpy().astype(np.float64)
    ref = ref - ref.mean(axis=(1, 2), keepdims=True)
    assert np.abs(out - ref).max() < 1e-6
    assert np.abs(out.sum(axis=(1, 2))).max() < 1e-4


def _center_fixed_order(x):
    """The centring's stated order: thread t of 256 sums x[t], x[t + 256], ...
    in FP64 one after another, each wave of 64 combines its threads with an
    xor butterfly, the four wave sums are added as (w0 + w1) + (w2 + w3), and
    the FP32 mean is subtracted in FP32."""
    out = np.empty_like(x)
    for k in range(x.shape[0]):
        xs = x[k].reshape(-1)
        part = np.zeros(256)
        for t in range(256):
            acc = 0.0
            for v in xs[t::256].tolist():
                acc += v
            part[t] = acc
        lanes = np.arange(64)
        sums = []
        for w in range(4):
            v = part[64 * w: 64 * w + 64].copy()
            for o in (32, 16, 8, 4, 2, 1):
                v = v + v[lanes ^ o]
            sums.append(v[0])
        mean = np.float32(((sums[0] + sums[1]) + (sums[2] + sums[3])) / xs.size)
        out[k] = x[k] - mean
    return out


@pytest.mark.parametrize("shape", [(3, 8, 16), (2, 67, 73), (2, 256, 256)])
def test_center_fixed_order_bitwise(gpu_lib, shape):
    # 67 x 73 has 4891 values per sample: one batch of 16 loads per thread
    # and a ragged tail; 256 x 256 is the surrogate's field
    print(f"[case] center {shape}", flush=True)
    x = (rnd(shape, 31) + 0.3).astype(np.float32)
    out = nn.center_raw(torch.from_numpy(x).to(DEV)).cpu().numpy()
    assert np.array_equal(out, _center_fixed_order(x))


def test_graphed_forward_replays_match_eager(gpu_lib):
    """The captured forward is replayed several times with different inputs,
    with eager work and synchronisations between the replays; every replay
    must equal the eager forward of its input bit for bit."""
    print("[case] graphed U-Net forward", flush=True)
    torch.manual_seed(0)
    net = nn.UNet(seed=3).to(DEV)
    fwd = nn.GraphedForward(net)
    inputs_ = [torch.from_numpy(rnd((4, 64, 64), 40 + i, 0.1).astype(np.float32)).to(DEV)
               for i in range(3)]
    for rep in range(2):
        for x in inputs_:
            with torch.no_grad():
                want = net(x)
            _ = torch.randn(1000, device=DEV).sum()  # eager work between replays
            torch.cuda.synchronize()
            got = fwd(x)
            torch.cuda.synchronize()
            assert torch.equal(got, want), (rep, (got - want).abs().max().item())


def test_weight_gradients_are_deterministic(gpu_lib):
    """Per-workgroup partial sums added in a fixed order: two calls on the same
    inputs give the same bits, for the MFMA gradient and both thin layers."""
    print("[case] deterministic weight gradients", flush=True)
    N, H, W, C, K = 4, 32, 64, 64, 64
    x = bf(rnd((N, H, W, C), 50))
    dy = bf(rnd((N, H, W, K), 51))
    a, b = nn.conv3x3_wgrad_raw(x, dy, K), nn.conv3x3_wgrad_raw(x, dy, K)
    assert torch.equal(a, b)
    lib = gpu_lib
    xin = torch.from_numpy(rnd((N, H, W), 52).astype(np.float32)).to(DEV)
    outs = []
    for _ in range(2):
        dw = torch.zeros((9, K), device=DEV)
        ws = nn.wgrad_workspace(nn.WGRAD_IN1, N, H, W, 1, K, DEV)
        nn._check(lib.sps_nn_conv3x3_in1_wgrad(nn._ptr(xin), nn._ptr(dy), N, H, W, K, nn._ptr(dw),
                                               nn._ptr(ws), ws.numel(), nn._stream()))
        dwo = torch.zeros((9, C), device=DEV)
        wso = nn.wgrad_workspace(nn.WGRAD_OUT1, N, H, W, C, 1, DEV)
        wo = torch.zeros((9, C), device=DEV)
        nn._check(lib.sps_nn_conv3x3_out1_bwd(nn._ptr(x), nn._ptr(xin), N, H, W, C, nn._ptr(wo),
                                              None, nn._ptr(dwo), nn._ptr(wso), wso.numel(),
                                              nn._stream()))
        outs.append((dw.clone(), dwo.clone()))
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
