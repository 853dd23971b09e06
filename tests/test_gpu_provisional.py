"""Provisional variables (PROVISIONAL_DEF) on the GPU against the oracle.

MBB (examples/models/modeldefinitions.h:150-155) and BB (tests/testdefs.h:
29-37) evaluate var(1) into a grid of their own before the right-hand side
reads it (sol/inc/modelspecialized.h:430-434).  The reference-order path runs
that as a pass of its own and must match the FP64 oracle bit for bit; the
fused path substitutes var(1) into the canonical form.
"""
import numpy as np
import pytest

from symphas_amd import abi, inputs, models
from symphas_amd.models import Stencil2d2h, Stencil3d2h, TermSpec
from symphas_amd.solver import FDPlan, FDBatch

pytestmark = pytest.mark.gpu

F32, F64 = abi.F32, abi.F64
REF, FUSED, AUTO = abi.PATH_REFERENCE, abi.PATH_FUSED, abi.PATH_AUTO


def say(*a):
    print("[case]", *a, flush=True)


def field(desc, seed, lo=-0.5, hi=0.5):
    shape = tuple(int(desc.n[a]) for a in reversed(range(desc.dim)))
    return inputs.uniform(shape, seed=seed, lo=lo, hi=hi)


def gpu_run(desc, psi, nsteps, K=0, TY=0):
    with FDPlan(desc) as p:
        if K or TY:
            p.tune(K, TY)
        p.upload(psi.astype(p.dtype))
        p.step(nsteps)
        return p.download()


def oracle_run(orc, desc, psi, nsteps):
    return orc.unpad(desc, orc.run(desc, psi.astype(np.float64), nsteps, 8))


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


CASES = [("mbb", Stencil2d2h(5, 6, 13), (40, 28)),
         ("bb", Stencil2d2h(5, 6, 13), (40, 28)),
         ("mbb", Stencil2d2h(9, 6, 17), (33, 21)),
         ("bb", Stencil2d2h(9, 12, 21), (33, 21)),
         ("mbb", Stencil3d2h(7, 10, 25), (18, 14, 11)),
         ("bb", Stencil3d2h(15, 10, 41), (18, 14, 11))]


def build(model, n, stencil, dtype, path, dt=0.002):
    if model == "mbb":
        return models.model_mbb(n, h=0.9, dt=dt, stencil=stencil, dtype=dtype, path=path,
                                c1=1.1, c2=0.4)
    return models.model_bb(n, h=0.9, dt=dt, stencil=stencil, dtype=dtype, path=path,
                           a=1.1, b=0.4)


@pytest.mark.parametrize("model,stencil,n", CASES)
def test_reference_path_bitwise(gpu_lib, orc, model, stencil, n):
    desc = build(model, n, stencil, F64, REF)
    psi = field(desc, seed=41)
    say("provisional reference", model, stencil, n)
    assert np.array_equal(gpu_run(desc, psi, 25), oracle_run(orc, desc, psi, 25))


@pytest.mark.parametrize("model,stencil,n", [c for c in CASES if c[1].L in (5, 7)])
def test_fused_path(gpu_lib, orc, model, stencil, n):
    say("provisional fused", model, stencil, n)
    d64 = build(model, n, stencil, F64, AUTO)
    with FDPlan(d64) as p:
        assert p.query()[0] == FUSED
    psi = field(d64, seed=42)
    want = oracle_run(orc, d64, psi, 60)
    assert rel_l2(gpu_run(d64, psi, 60), want) <= 1e-12
    d32 = build(model, n, stencil, F32, FUSED)
    psi32 = psi.astype(np.float32)
    base = gpu_run(d32, psi32, 24, K=1)
    for K in (2, 3, 4) if stencil.dim == 2 else (2,):
        assert np.array_equal(gpu_run(d32, psi32, 24, K=K, TY=0 if stencil.dim == 2 else 4), base), K
    want32 = oracle_run(orc, d64, psi32.astype(np.float64), 24)
    assert rel_l2(base.astype(np.float64), want32) <= 2e-6


def test_non_5point_stencils_take_the_reference_path(gpu_lib):
    for model, stencil, n in CASES:
        if stencil.L in (5, 7):
            continue
        with FDPlan(build(model, n, stencil, F64, AUTO)) as p:
            assert p.query()[0] == REF, (model, stencil)


def test_nested_provisional_reference_bitwise(gpu_lib, orc):
    """var(1) = psi^3 - psi, var(2) = lap(var(1)) + 0.3*var(1),
    dpsi = -lap(var(2)) + 0.2*psi: a variable read through another one keeps
    the reference-order path."""
    P = models.PSI
    prov = [[TermSpec(abi.OP_ID, 1.0, (0.0, -1.0, 0.0, 1.0))],
            [TermSpec(abi.OP_LAP, 1.0, P, operand=1), TermSpec(abi.OP_ID, 0.3, P, operand=1)]]
    evo = [TermSpec(abi.OP_LAP, -1.0, P, operand=2), TermSpec(abi.OP_ID, 0.2, P)]
    desc = models.make_desc((24, 20), Stencil2d2h(5, 6, 13), 1.0, 0.002, evo, F64, AUTO, prov)
    psi = field(desc, seed=43)
    say("nested provisionals")
    with FDPlan(desc) as p:
        assert p.query()[0] == REF
    assert np.array_equal(gpu_run(desc, psi, 20), oracle_run(orc, desc, psi, 20))


def test_equation_includes_provisionals(gpu_lib, orc):
    desc = build("bb", (30, 22), Stencil2d2h(5, 6, 13), F64, FUSED)
    psi = field(desc, seed=44)
    say("equation with provisional")
    with FDPlan(desc) as p:
        p.upload(psi)
        got = p.equation()
    g = orc.refresh(desc, orc.pad(desc, psi))
    assert np.array_equal(got, orc.unpad(desc, orc.equation(desc, g)))


def test_batched_mbb_matches_single_plans(gpu_lib):
    descs = [build("mbb", (128, 128), Stencil2d2h(5, 6, 13), F32, FUSED, dt=0.01),
             build("bb", (192, 192), Stencil2d2h(5, 6, 13), F32, FUSED, dt=0.01)]
    fields = [field(d, seed=45 + i).astype(np.float32) for i, d in enumerate(descs)]
    say("batched provisionals")
    with FDBatch(descs) as b:
        for i, f in enumerate(fields):
            b.upload(i, f)
        b.step(12)
        got = [b.download(i) for i in range(len(descs))]
    for d, f, g in zip(descs, fields, got):
        assert np.array_equal(g, gpu_run(d, f, 12))


def test_invalid_provisional_descriptors(gpu_lib):
    d = build("bb", (16, 16), Stencil2d2h(5, 6, 13), F64, AUTO)
    d.terms[0].operand = 2
    with pytest.raises(abi.SpsError) as ei:
        FDPlan(d)
    assert ei.value.status == abi.E_INVALID
    d = build("bb", (16, 16), Stencil2d2h(5, 6, 13), F64, AUTO)
    d.prov[0][0].operand = 1
    with pytest.raises(abi.SpsError):
        FDPlan(d)
    d = build("bb", (16, 16), Stencil2d2h(5, 6, 13), F64, AUTO)
    d.nprov = 3
    with pytest.raises(abi.SpsError):
        FDPlan(d)
