"""BF16 convolutional U-Net whose convolutions run on the project's own MFMA
kernels (include/sps_nn.h).

torch provides device memory, streams, autograd bookkeeping and the optimiser;
every convolution forward, backward-data and backward-weight pass, the
activation backward and the pooling passes are HIP kernels of
symphas_amd/lib/libsps_fd.so called through ctypes.  Activations are NHWC
BF16; master weights are FP32 [9][out][in] and are rounded to BF16 (round to
nearest even) for each use.

The network is the 4-level U-Net of the surrogate: 32/64/128/256 channels,
two 3x3 circular convolutions per level, average-pool down, and on the way up
a 3x3 convolution at the coarse level whose output is upsampled 2x and added
to the skip of the finer level (fused into the convolution's epilogue).  It
predicts the increment of a fixed number of solver steps; the increment's
mean is removed so that the surrogate conserves mass exactly, as Model B does.
"""
from __future__ import annotations

import ctypes
import math
from ctypes import c_void_p

import torch

from . import abi

ACT_NONE, ACT_SILU = 0, 1
EPI_PLAIN, EPI_UPADD = 0, 1


def _lib():
    return abi.load()


def _ptr(t):
    return c_void_p(t.data_ptr()) if t is not None else c_void_p(0)


def _stream():
    return c_void_p(torch.cuda.current_stream().cuda_stream)


def _check(rc):
    abi.check(rc)


def conv3x3_raw(x, w_bf16, bias, act, epi=EPI_PLAIN, skip=None, want_pre=False):
    """x bf16 [N,H,W,C]; w_bf16 [9,K,C]; returns (y, pre)."""
    N, H, W, C = x.shape
    K = w_bf16.shape[1]
    oh, ow = (H, W) if epi == EPI_PLAIN else (2 * H, 2 * W)
    y = torch.empty((N, oh, ow, K), device=x.device, dtype=torch.bfloat16)
    pre = torch.empty_like(y) if want_pre else None
    _check(_lib().sps_nn_conv3x3(_ptr(x), N, H, W, C, _ptr(w_bf16), _ptr(bias), K, act, epi,
                                 _ptr(skip), _ptr(y), _ptr(pre), _stream()))
    return y, pre


WGRAD_CONV, WGRAD_IN1, WGRAD_OUT1 = 0, 1, 2


def wgrad_workspace(kind, N, H, W, C, K, device):
    """FP32 scratch for a deterministic weight gradient (per-workgroup partial
    sums, added in a fixed order)."""
    n = _lib().sps_nn_wgrad_workspace(kind, N, H, W, C, K)
    return torch.empty(max(int(n), 1), device=device, dtype=torch.float32)


def conv3x3_wgrad_raw(x, dpre, K):
    N, H, W, C = x.shape
    dw = torch.zeros((9, K, C), device=x.device, dtype=torch.float32)
    ws = wgrad_workspace(WGRAD_CONV, N, H, W, C, K, x.device)
    _check(_lib().sps_nn_conv3x3_wgrad(_ptr(x), _ptr(dpre), N, H, W, C, K, _ptr(dw), _ptr(ws),
                                       ws.numel(), _stream()))
    return dw


def silu_bwd_raw(dy, pre):
    dx = torch.empty_like(dy)
    _check(_lib().sps_nn_silu_bwd(_ptr(dy), _ptr(pre), _ptr(dx), dy.numel(), _stream()))
    return dx


def pool2_raw(x, mode):
    N, H, W, C = x.shape
    y = torch.empty((N, H // 2, W // 2, C), device=x.device, dtype=torch.bfloat16)
    _check(_lib().sps_nn_pool2(_ptr(x), N, H, W, C, mode, _ptr(y), _stream()))
    return y


def unpool2_raw(x, scale):
    N, H, W, C = x.shape
    y = torch.empty((N, 2 * H, 2 * W, C), device=x.device, dtype=torch.bfloat16)
    _check(_lib().sps_nn_unpool2(_ptr(x), N, H, W, C, ctypes.c_float(scale), _ptr(y), _stream()))
    return y


def flip_weights(w):
    """[9][K][C] -> [9][C][K] with the taps reversed: the weights of the
    backward-data convolution (a correlation with the flipped kernel)."""
    return w.flip(0).transpose(1, 2).contiguous()


class _Conv(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w, b, act, epi, skip, need_grad):
        wbf = w.detach().to(torch.bfloat16).contiguous()
        y, pre = conv3x3_raw(x.contiguous(), wbf, b.detach(), act, epi,
                             skip.contiguous() if skip is not None else None,
                             want_pre=need_grad and act == ACT_SILU)
        ctx.act, ctx.epi = act, epi
        ctx.save_for_backward(x, w, pre)
        return y

    @staticmethod
    def backward(ctx, dy):
        x, w, pre = ctx.saved_tensors
        dy = dy.contiguous().to(torch.bfloat16)
        dpre = silu_bwd_raw(dy, pre) if ctx.act == ACT_SILU else dy
        dskip = None
        if ctx.epi == EPI_UPADD:
            dskip = dpre
            dlow = pool2_raw(dpre, 1)
        else:
            dlow = dpre
        K, C = w.shape[1], w.shape[2]
        wb = flip_weights(w.detach()).to(torch.bfloat16)
        dx, _ = conv3x3_raw(dlow, wb, None, ACT_NONE)
        dw = conv3x3_wgrad_raw(x.contiguous(), dlow, K)
        # FP32 accumulation straight from the BF16 gradient (no FP32 copy)
        db = torch.sum(dlow, dim=(0, 1, 2), dtype=torch.float32)
        return dx, dw, db, None, None, dskip, None


class _In1(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w, b, act):
        N, H, W = x.shape
        K = w.shape[1]
        y = torch.empty((N, H, W, K), device=x.device, dtype=torch.bfloat16)
        # the pre-activation is kept only when a backward pass will need it
        pre = torch.empty_like(y) if act == ACT_SILU and any(ctx.needs_input_grad) else None
        _check(_lib().sps_nn_conv3x3_in1(_ptr(x.contiguous()), N, H, W, _ptr(w.detach().contiguous()),
                                         _ptr(b.detach()), K, act, _ptr(y), _ptr(pre), _stream()))
        ctx.act = act
        ctx.save_for_backward(x, w, pre)
        return y

    @staticmethod
    def backward(ctx, dy):
        x, w, pre = ctx.saved_tensors
        dy = dy.contiguous().to(torch.bfloat16)
        dpre = silu_bwd_raw(dy, pre) if ctx.act == ACT_SILU else dy
        N, H, W = x.shape
        K = w.shape[1]
        dw = torch.zeros((9, K), device=x.device, dtype=torch.float32)
        ws = wgrad_workspace(WGRAD_IN1, N, H, W, 1, K, x.device)
        _check(_lib().sps_nn_conv3x3_in1_wgrad(_ptr(x.contiguous()), _ptr(dpre), N, H, W, K,
                                               _ptr(dw), _ptr(ws), ws.numel(), _stream()))
        db = torch.sum(dpre, dim=(0, 1, 2), dtype=torch.float32)
        return None, dw, db, None


class _Out1(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w):
        N, H, W, C = x.shape
        y = torch.empty((N, H, W), device=x.device, dtype=torch.float32)
        _check(_lib().sps_nn_conv3x3_out1(_ptr(x.contiguous()), N, H, W, C,
                                          _ptr(w.detach().contiguous()), c_void_p(0),
                                          _ptr(y), _stream()))
        ctx.save_for_backward(x, w)
        return y

    @staticmethod
    def backward(ctx, dy):
        x, w = ctx.saved_tensors
        N, H, W, C = x.shape
        dy = dy.contiguous().float()
        dx = torch.empty_like(x)
        dw = torch.zeros((9, C), device=x.device, dtype=torch.float32)
        ws = wgrad_workspace(WGRAD_OUT1, N, H, W, C, 1, x.device)
        _check(_lib().sps_nn_conv3x3_out1_bwd(_ptr(x.contiguous()), _ptr(dy), N, H, W, C,
                                              _ptr(w.detach().contiguous()), _ptr(dx), _ptr(dw),
                                              _ptr(ws), ws.numel(), _stream()))
        return dx, dw


class _Pool(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return pool2_raw(x.contiguous(), 0)

    @staticmethod
    def backward(ctx, dy):
        return unpool2_raw(dy.contiguous().to(torch.bfloat16), 0.25)


def center_raw(d):
    """d [N, ...] FP32 minus each sample's mean (FP64 sum, fixed order)."""
    out = torch.empty_like(d)
    _check(_lib().sps_nn_center(_ptr(d), _ptr(out), d.shape[0], d[0].numel(), _stream()))
    return out


class _Center(torch.autograd.Function):
    # the projection I - 11^T/n is symmetric: its backward is itself
    @staticmethod
    def forward(ctx, d):
        return center_raw(d.contiguous())

    @staticmethod
    def backward(ctx, g):
        return center_raw(g.contiguous())


class UNet(torch.nn.Module):
    """4-level U-Net on periodic fields, predicting psi(t + k steps)."""

    def __init__(self, base=32, levels=4, out_scale=0.1, seed=0):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        ch = [base * 2 ** i for i in range(levels)]
        self.ch = ch
        self.out_scale = out_scale

        def conv(cin, cout):
            w = torch.randn(9, cout, cin, generator=g) * math.sqrt(2.0 / (9 * cin))
            return torch.nn.Parameter(w), torch.nn.Parameter(torch.zeros(cout))

        self.w_in, self.b_in = conv(1, ch[0])
        self.w_in = torch.nn.Parameter(self.w_in.detach()[:, :, 0].contiguous())  # [9][K]
        enc = []
        for l in range(levels):
            cin = ch[0] if l == 0 else ch[l - 1]
            a = conv(cin if l > 0 else ch[0], ch[l])
            b = conv(ch[l], ch[l])
            enc += [a, b]
        self.enc_w = torch.nn.ParameterList([p[0] for p in enc])
        self.enc_b = torch.nn.ParameterList([p[1] for p in enc])
        dec = []
        for l in reversed(range(levels - 1)):
            up = conv(ch[l + 1], ch[l])
            same = conv(ch[l], ch[l])
            dec += [up, same]
        self.dec_w = torch.nn.ParameterList([p[0] for p in dec])
        self.dec_b = torch.nn.ParameterList([p[1] for p in dec])
        # no output bias: the increment's mean is removed (mass conservation),
        # which would cancel it
        self.w_out = torch.nn.Parameter(torch.zeros(9, ch[0]))

    def increment(self, psi):
        """psi: FP32 [N,H,W] -> FP32 increment [N,H,W] (mean removed)."""
        L = len(self.ch)
        g = torch.is_grad_enabled()
        x = _In1.apply(psi, self.w_in, self.b_in, ACT_SILU)
        skips = []
        for l in range(L):
            if l > 0:
                x = _Pool.apply(x)
            x = _Conv.apply(x, self.enc_w[2 * l], self.enc_b[2 * l], ACT_SILU, EPI_PLAIN, None, g)
            x = _Conv.apply(x, self.enc_w[2 * l + 1], self.enc_b[2 * l + 1], ACT_SILU, EPI_PLAIN, None, g)
            skips.append(x)
        for i, l in enumerate(reversed(range(L - 1))):
            x = _Conv.apply(x, self.dec_w[2 * i], self.dec_b[2 * i], ACT_SILU, EPI_UPADD, skips[l], g)
            x = _Conv.apply(x, self.dec_w[2 * i + 1], self.dec_b[2 * i + 1], ACT_SILU, EPI_PLAIN, None, g)
        d = _Out1.apply(x, self.w_out)
        return _Center.apply(d)

    def forward(self, psi):
        return psi + self.out_scale * self.increment(psi)

    def flops_per_sample(self, H, W):
        """Multiply-add FLOPs (2 per MAC) of one forward pass."""
        L = len(self.ch)
        f = 2 * 9 * 1 * self.ch[0] * H * W
        for l in range(L):
            hw = (H >> l) * (W >> l)
            cin = self.ch[0] if l == 0 else self.ch[l - 1]
            f += 2 * 9 * cin * self.ch[l] * hw + 2 * 9 * self.ch[l] * self.ch[l] * hw
        for l in reversed(range(L - 1)):
            hw_c = (H >> (l + 1)) * (W >> (l + 1))
            hw = (H >> l) * (W >> l)
            f += 2 * 9 * self.ch[l + 1] * self.ch[l] * hw_c + 2 * 9 * self.ch[l] * self.ch[l] * hw
        f += 2 * 9 * self.ch[0] * H * W
        return f

    def layer_costs(self, H, W):
        """(name, FLOPs, HBM bytes) per sample for every kernel of one
        inference pass: each convolution reads its input once and writes its
        output once (BF16 activations; FP32 network input and increment; the
        upsample-add layers also read the skip), each pooling reads its input
        and writes a quarter of it, the centring reads and writes the
        increment, and psi + scale * increment reads two fields and writes
        one."""
        L, ch = len(self.ch), self.ch
        out = [("in1", 2 * 9 * ch[0] * H * W, 4 * H * W + 2 * ch[0] * H * W)]
        for l in range(L):
            hw = (H >> l) * (W >> l)
            if l > 0:
                phw = (H >> (l - 1)) * (W >> (l - 1))
                out.append((f"pool{l}", 0, 2 * ch[l - 1] * (phw + phw // 4)))
            cin = ch[0] if l == 0 else ch[l - 1]
            out.append((f"enc{l}a", 2 * 9 * cin * ch[l] * hw, 2 * (cin + ch[l]) * hw))
            out.append((f"enc{l}b", 2 * 9 * ch[l] * ch[l] * hw, 4 * ch[l] * hw))
        for l in reversed(range(L - 1)):
            hw_c = (H >> (l + 1)) * (W >> (l + 1))
            hw = (H >> l) * (W >> l)
            out.append((f"up{l}", 2 * 9 * ch[l + 1] * ch[l] * hw_c,
                        2 * ch[l + 1] * hw_c + 2 * 2 * ch[l] * hw))
            out.append((f"dec{l}", 2 * 9 * ch[l] * ch[l] * hw, 4 * ch[l] * hw))
        out.append(("out1", 2 * 9 * ch[0] * H * W, 2 * ch[0] * H * W + 4 * H * W))
        out.append(("center", 0, 8 * H * W))
        out.append(("update", 0, 12 * H * W))
        return out

    def roofline_seconds(self, H, W, batch, peak_flops=2.5e15, peak_bytes=8e12):
        """Lower bound on one inference pass of batch samples: the sum over its
        kernels of max(FLOPs / MFMA peak, bytes / HBM peak)."""
        return sum(max(f * batch / peak_flops, b * batch / peak_bytes)
                   for _, f, b in self.layer_costs(H, W))


class GraphedForward:
    """Inference of a network replayed from captured HIP graphs, one per input
    shape.  Every kernel of the forward pass runs on the capturing stream
    (the convolutions take torch's current stream) and nothing in it clears
    memory with a memset, so a replay repeats the captured launches exactly;
    it removes the per-layer host work of the eager forward.  The returned
    tensor is a copy: later replays do not change it."""

    def __init__(self, net):
        self.net = net
        self.graphs = {}

    def __call__(self, x):
        key = tuple(x.shape)
        if key not in self.graphs:
            static_in = x.detach().clone()
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side), torch.no_grad():
                self.net(static_in)  # warm: allocations and lazy set-up
            torch.cuda.current_stream().wait_stream(side)
            g = torch.cuda.CUDAGraph()
            with torch.no_grad(), torch.cuda.graph(g):
                static_out = self.net(static_in)
            self.graphs[key] = (g, static_in, static_out)
        g, static_in, static_out = self.graphs[key]
        static_in.copy_(x)
        g.replay()
        return static_out.clone()
