"""The configurations at their stated sizes: Model B 3D 512^3 FP32 (config
3), all 256 instances of the batched workload (config 4) and the surrogate's
256^2 fields in batches of 64 (config 5).  Each test prints its case before
any GPU work."""
import sys

import numpy as np
import pytest
import torch

from symphas_amd import abi, inputs, models
from symphas_amd import surrogate as sg
from symphas_amd.models import Stencil3d2h
from symphas_amd.solver import FDBatch, FDPlan

pytestmark = pytest.mark.gpu

F32, F64 = abi.F32, abi.F64


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


# --------------------------------------------------------------- config 3
N3 = 512


def _cfg3(dtype, path=abi.PATH_FUSED):
    return models.model_b((N3, N3, N3), h=1.0, dt=0.005, stencil=Stencil3d2h(7, 10, 25),
                          dtype=dtype, path=path)


@pytest.fixture(scope="module")
def cfg3_field():
    return inputs.uniform((N3, N3, N3), seed=3, lo=-0.05, hi=0.05, dtype=np.float32)


def test_cfg3_full_size_against_oracle(gpu_lib, orc, cfg3_field):
    """512^3 FP32 at the autotuner's shapes (2 fused steps) and at one step
    per launch, 2 steps, against the FP64 oracle (OpenMP on the host)."""
    psi = cfg3_field
    say("cfg3 512^3 FP32 2 steps, shapes (2,12,74) and (1,8,64)")
    outs = []
    for k, rows, planes in ((2, 12, 74), (1, 8, 64)):
        with FDPlan(_cfg3(F32)) as p:
            p.tune(k, rows, planes, 0)
            p.upload(psi)
            p.step(2)
            outs.append(p.download())
    # the per-site arithmetic is the same for every launch shape
    assert np.array_equal(outs[0], outs[1])
    say("cfg3 oracle 2 steps")
    d64 = _cfg3(F64, abi.PATH_REFERENCE)
    g = orc.pad(d64, psi.astype(np.float64))
    rc = orc.load().orc_run(orc._d(d64), orc._p(g), 2, 16)
    assert rc == 0
    want = orc.unpad(d64, g)
    del g
    err = rel_l2(outs[0].astype(np.float64), want)
    print(f"512^3 FP32 vs FP64 oracle after 2 steps: rel L2 {err:.3e}")
    assert err <= 2e-6


def test_cfg3_full_size_fp64_and_mass(gpu_lib, cfg3_field):
    """512^3: FP32 against the FP64 fused path after 200 steps (the tests pin
    that path to the oracle within 1e-12), mass drift within 1e-6."""
    psi = cfg3_field
    say("cfg3 512^3 FP32 vs FP64, 200 steps")
    with FDPlan(_cfg3(F32)) as p:
        p.upload(psi)
        m0, _ = p.reduce()
        p.step(200)
        m1, _ = p.reduce()
        got = p.download().astype(np.float64)
    with FDPlan(_cfg3(F64)) as p:
        p.upload(psi.astype(np.float64))
        p.step(200)
        want = p.download()
    err = rel_l2(got, want)
    drift = abs(m1 - m0) / float(np.abs(psi.astype(np.float64)).sum())
    print(f"512^3 FP32 vs FP64 after 200 steps: rel L2 {err:.3e}, mass drift {drift:.3e}")
    assert err <= 1e-5
    assert drift <= 1e-6


# --------------------------------------------------------------- config 4
def test_cfg4_all_256_instances(gpu_lib, orc):
    """All 256 instances of the batched workload in one batch: bit-identical
    to 256 single-field fused runs after 8 steps (K = 4: two persistent
    launches over every item), and a sample of instances against the FP64
    oracle."""
    inst = models.batched_instances(256, seed=7)
    descs = [d for d, _ in inst]
    fields = [inputs.uniform((d.n[1], d.n[0]), seed=s, dtype=np.float32) for d, s in inst]
    sites = sum(d.n[0] * d.n[1] for d in descs)
    say(f"cfg4 256 instances, {sites} sites, 8 steps")
    with FDBatch(descs) as b:
        b.tune(4, 48)
        for i, f in enumerate(fields):
            b.upload(i, f)
        b.step(8)
        got = [b.download(i) for i in range(len(descs))]
    for i, (d, f) in enumerate(zip(descs, fields)):
        with FDPlan(d) as p:
            p.upload(f)
            p.step(8)
            assert np.array_equal(got[i], p.download()), i
    say("cfg4 oracle sample")
    for i in (0, 77, 128, 200, 255):
        d64 = type(descs[i]).from_buffer_copy(descs[i])
        d64.dtype = F64
        d64.path = abi.PATH_REFERENCE
        g = orc.run(d64, fields[i].astype(np.float64), 8)
        want = orc.unpad(d64, g)
        err = rel_l2(got[i].astype(np.float64), want)
        assert err <= 2e-6, (i, err)


# --------------------------------------------------------------- config 5
class _Exact(torch.nn.Module):
    """A 'surrogate' that is the plain solver itself."""

    def __init__(self, batch, n, k):
        super().__init__()
        self.s, self.k = sg.PlainSolver(batch, n), k

    def forward(self, psi):
        return self.s.advance(psi, self.k)


def _cfg5_fields(batch, n, seed0):
    return torch.stack([torch.from_numpy(sg.initial_field(n, seed0 + i))
                        for i in range(batch)]).cuda()


def test_cfg5_exact_surrogate_never_falls_back(gpu_lib):
    n, k, batch = 256, 50, 64
    say("cfg5 256^2 x 64 exact surrogate")
    psi0 = _cfg5_fields(batch, n, 3000)
    ex = _Exact(batch, n, k)
    ctl = sg.SurrogateSolver(ex, batch, n, k=k, tol=1e-3, check_every=4)
    out, rep = ctl.advance(psi0, 8)
    ctl.close()
    ex.s.close()
    assert rep.fallbacks == 0 and rep.checks == 2 and rep.max_check_err == 0.0
    plain, _ = sg.time_plain(psi0, 8 * k, n)
    assert torch.equal(out, plain)


def test_cfg5_fallback_returns_plain_trajectory(gpu_lib):
    from symphas_amd.nn import UNet
    n, k, batch = 256, 50, 64
    say("cfg5 256^2 x 64 unfit surrogate")
    net = UNet(seed=4).cuda()
    with torch.no_grad():
        net.w_out.copy_(torch.randn_like(net.w_out))
    psi0 = _cfg5_fields(batch, n, 4000)
    ctl = sg.SurrogateSolver(net, batch, n, k=k, tol=1e-3, check_every=4)
    out, rep = ctl.advance(psi0, 8)
    ctl.close()
    assert rep.fallbacks == batch and rep.surrogate_calls == 1
    plain, _ = sg.time_plain(psi0, 8 * k, n)
    assert torch.equal(out, plain)
