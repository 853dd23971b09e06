"""Coupled-field models (Model C, Gray-Scott, Turing) on the GPU against the
oracle: the reference-order path bit for bit in FP64, the fused two-field
path within 1e-12 (FP64) and bitwise invariant across launch shapes."""
import numpy as np
import pytest

from symphas_amd import abi, inputs, models
from symphas_amd.models import Stencil2d2h, Stencil3d2h
from symphas_amd.solver import MultiPlan

pytestmark = pytest.mark.gpu

F32, F64 = abi.F32, abi.F64
REF, FUSED, AUTO = abi.PATH_REFERENCE, abi.PATH_FUSED, abi.PATH_AUTO


def say(*a):
    print("[case]", *a, flush=True)


def fields_for(d, seed, lo=-0.5, hi=0.5, offset=(0.0, 0.0)):
    shape = tuple(int(d.n[a]) for a in reversed(range(d.dim)))
    return [inputs.uniform(shape, seed=seed + f, lo=lo, hi=hi) + offset[f]
            for f in range(d.nfields)]


def gpu_run(d, fields, nsteps, K=0, TY=0, vec=0):
    with MultiPlan(d) as p:
        if K or TY or vec:
            p.tune(K, TY, vec)
        for f, a in enumerate(fields):
            p.upload(f, a.astype(p.dtype))
        p.step(nsteps)
        return [p.download(f) for f in range(d.nfields)]


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


CASES = [
    ("model_c", Stencil2d2h(5, 6, 13), (40, 28), 0.002, (1.1, 0.8, 0.6, 1.3, 0.4)),
    ("model_c_swapped", Stencil2d2h(5, 6, 13), (36, 30), 0.002, (1.1, 0.8, 0.6, 1.3, 0.4)),
    ("model_c", Stencil2d2h(9, 6, 17), (21, 19), 0.002, (1.0, 1.0, 0.5, 0.9, 0.3)),
    ("model_c", Stencil3d2h(7, 10, 25), (14, 11, 9), 0.002, (1.0, 1.0, 0.5, 0.9, 0.3)),
    ("gray_scott", Stencil2d2h(5, 6, 13), (48, 36), 0.5, (0.16, 0.08, 0.035, 0.065)),
    ("turing", Stencil2d2h(5, 6, 13), (32, 30), 0.01, (1.0, 0.7, 1.2, 0.5, -0.4, 0.9)),
    ("turing", Stencil3d2h(15, 10, 41), (9, 8, 7), 0.005, (1.0, 0.7, 1.2, 0.5, -0.4, 0.9)),
    # both fields with a bilaplacian and mixed monomials: the general
    # radius-two kernel with the triangular and dense polynomial classes
    ("mixed", Stencil2d2h(5, 6, 13), (44, 30), 0.002, (0.6, 0.8)),
    ("mixed", Stencil3d2h(7, 10, 25), (12, 10, 9), 0.002, (0.6, 0.8)),
]


def build(name, stencil, n, dt, c, dtype, path):
    if name == "model_c_swapped":
        # Model C with the fields in the other order: the field without a
        # bilaplacian (rho, B = rho) comes first
        eqs = [[models.MTermSpec(t.op, t.coeff, tuple(t.pow[::-1]), t.axis, t.component)
                for t in eq] for eq in reversed(models.model_c_terms(c))]
        return models.make_mdesc(n, stencil, 1.0, dt, eqs, dtype, path)
    return getattr(models, name)(n, h=1.0, dt=dt, stencil=stencil, dtype=dtype, path=path, c=c)


@pytest.mark.parametrize("name,stencil,n,dt,c", CASES)
def test_reference_path_bitwise(gpu_lib, orc, name, stencil, n, dt, c):
    say("coupled reference", name, stencil, n)
    d = build(name, stencil, n, dt, c, F64, REF)
    fields = fields_for(d, 71)
    got = gpu_run(d, fields, 15)
    want = orc.mrun(d, fields, 15, 8)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


@pytest.mark.parametrize("name,stencil,n,dt,c", [c for c in CASES if c[1].dim == 2 and c[1].L == 5])
def test_fused_path(gpu_lib, orc, name, stencil, n, dt, c):
    say("coupled fused", name, n)
    d64 = build(name, stencil, n, dt, c, F64, AUTO)
    with MultiPlan(d64) as p:
        assert p.query()[0] == FUSED
    fields = fields_for(d64, 72)
    want = orc.mrun(d64, fields, 40, 8)
    for g, w in zip(gpu_run(d64, fields, 40), want):
        assert rel_l2(g, w) <= 1e-12
    d32 = build(name, stencil, n, dt, c, F32, FUSED)
    f32 = [a.astype(np.float32) for a in fields]
    base = gpu_run(d32, f32, 24, K=1, TY=8, vec=2)
    for vec in (2, 4):
        for K in (2, 3, 4):
            for TY in (4, 16, 23, 64):
                got = gpu_run(d32, f32, 24, K=K, TY=TY, vec=vec)
                for g, b in zip(got, base):
                    assert np.array_equal(g, b), (K, TY, vec)
    want32 = orc.mrun(d64, [a.astype(np.float64) for a in f32], 24, 8)
    for g, w in zip(base, want32):
        assert rel_l2(g.astype(np.float64), w) <= 2e-6


def test_gray_scott_pattern_full_size(gpu_lib, orc):
    """2048^2 Gray-Scott seeded with a square of the second species: the
    fused FP32 run stays within FP32 rounding of the FP64 reference path."""
    n = (2048, 2048)
    say("gray-scott 2048^2")
    u = np.ones((2048, 2048))
    v = np.zeros((2048, 2048))
    u[960:1088, 960:1088] = 0.5
    v[960:1088, 960:1088] = 0.25
    d32 = models.gray_scott(n, dtype=F32, path=FUSED)
    d64 = models.gray_scott(n, dtype=F64, path=REF)
    got = gpu_run(d32, [u, v], 200)
    want = gpu_run(d64, [u, v], 200)
    for g, w in zip(got, want):
        assert rel_l2(g.astype(np.float64), w) <= 1e-5


def test_paths_and_layouts(gpu_lib):
    d = models.model_c((12, 10), stencil=Stencil2d2h(5, 6, 13), dtype=F64, path=AUTO)
    with MultiPlan(d) as p:
        assert p.query()[0] == FUSED
        a = inputs.uniform((10, 12), seed=73)
        p.upload(1, np.pad(a, 3, mode="wrap"), abi.LAYOUT_PADDED)
        p.upload(0, a)
        assert np.array_equal(p.download(1), a)
        assert np.array_equal(p.download(0, abi.LAYOUT_PADDED), np.pad(a, 3, mode="wrap"))
    with MultiPlan(models.model_c((12, 10, 6), stencil=Stencil3d2h(7, 10, 25),
                                  path=AUTO)) as p:
        assert p.query()[0] == FUSED
    for d in (models.model_c((12, 10, 6), stencil=Stencil3d2h(15, 10, 25), path=AUTO),
              models.turing((12, 10), stencil=Stencil2d2h(9, 6, 13), path=AUTO)):
        with MultiPlan(d) as p:
            assert p.query()[0] == REF
    g = models.make_mdesc((12, 10), Stencil2d2h(5, 6, 13), 1.0, 0.01,
                          [[models.MTermSpec(abi.OP_GRADLAP, 1.0, (1, 0), axis=0)],
                           [models.MTermSpec(abi.OP_LAP, 1.0, (0, 1))]], F64, AUTO)
    with MultiPlan(g) as p:
        assert p.query()[0] == REF


def test_invalid_descriptors(gpu_lib):
    d = models.gray_scott((16, 16))
    d.terms[0][1].pow[2] = 1
    with pytest.raises(abi.SpsError) as ei:
        MultiPlan(d)
    assert ei.value.status == abi.E_INVALID
    d = models.gray_scott((16, 16))
    d.nfields = 0
    with pytest.raises(abi.SpsError):
        MultiPlan(d)
    d = models.model_c((16, 16, 8), stencil=Stencil3d2h(15, 10, 25), path=FUSED)
    with pytest.raises(abi.SpsError) as ei:
        MultiPlan(d)
    assert ei.value.status == abi.E_UNSUPPORTED
    with MultiPlan(models.model_c((16, 16, 8), stencil=Stencil3d2h(7, 10, 25),
                                  path=FUSED)) as p:
        for K, TY in ((2, 8), (1, 6), (1, 16)):
            with pytest.raises(abi.SpsError):
                p.tune(K, TY)


def test_autotune_keeps_fields(gpu_lib):
    """The coupled autotuner times launch shapes on copies of the fields: the
    state after autotune + steps equals the state after the same steps alone."""
    d = models.gray_scott((256, 192), dtype=F32, path=FUSED)
    u = np.ones((192, 256), np.float32)
    v = np.zeros((192, 256), np.float32)
    u[80:112, 100:140] = 0.5
    v[80:112, 100:140] = 0.25
    say("coupled autotune 256x192")
    want = gpu_run(d, [u, v], 12, K=1, TY=8)
    with MultiPlan(d) as p:
        p.upload(0, u)
        p.upload(1, v)
        k, rows, vec = p.autotune()
        assert 1 <= k <= 4 and 8 <= rows <= 512 and vec in (2, 4)
        p.step(12)
        for f in range(2):
            assert np.array_equal(p.download(f), want[f])


def test_coupled_field_files(gpu_lib, orc):
    """Each field of a coupled plan written as the reference's checkpoint,
    plot and XDR files (one file per field), byte-identical to the oracle's
    restatements; reading the checkpoints back into a fresh plan resumes the
    run bit for bit."""
    d = models.model_c((40, 28), h=0.7, dt=0.002, stencil=Stencil2d2h(5, 6, 13), dtype=F32,
                       path=FUSED, c=(1.1, 0.8, 0.6, 1.3, 0.4))
    fields = [a.astype(np.float32) for a in fields_for(d, 81)]
    single = models.make_desc((40, 28), Stencil2d2h(5, 6, 13), 0.7, 0.002, [])
    say("coupled checkpoint files")
    with MultiPlan(d) as p:
        for f, a in enumerate(fields):
            p.upload(f, a)
        p.step(10)
        now = [p.download(f) for f in range(2)]
        texts = [p.checkpoint_text(f, index=10, origin=[0.5, -1.0]) for f in range(2)]
        xdrs = [p.checkpoint_xdr(f, index=10) for f in range(2)]
        plots = [p.plot_text(f, abi.PLOT_GNU) for f in range(2)]
        csvs = [p.plot_text(f, abi.PLOT_CSV) for f in range(2)]
        p.step(7)
        want = [p.download(f) for f in range(2)]
    for f in range(2):
        g = orc.pad(single, now[f].astype(np.float64))
        assert texts[f] == orc.checkpoint_text(single, g, 10, origin=[0.5, -1.0])
        assert xdrs[f] == orc.xdr_write(single, g, 10)
        assert plots[f] == orc.plot_text(single, g, abi.PLOT_GNU)
        assert csvs[f] == orc.plot_text(single, g, abi.PLOT_CSV)
    with MultiPlan(d) as q:
        for f in range(2):
            assert q.read_checkpoint(f, texts[f]) == 10
        q.step(7)
        assert all(np.array_equal(q.download(f), want[f]) for f in range(2))
    with MultiPlan(d) as q:
        for f in range(2):
            assert q.read_checkpoint_xdr(f, xdrs[f]) == 10
        assert all(np.array_equal(q.download(f), now[f]) for f in range(2))
        with pytest.raises(abi.SpsError):
            q.checkpoint_text(2)


CASES3 = [
    ("model_c", (40, 18, 12), 0.002, (1.1, 0.8, 0.6, 1.3, 0.4)),
    ("model_c_swapped", (34, 16, 10), 0.002, (1.1, 0.8, 0.6, 1.3, 0.4)),
    ("gray_scott", (36, 20, 14), 0.5, (0.16, 0.08, 0.035, 0.065)),
    ("turing", (30, 22, 9), 0.01, (1.0, 0.7, 1.2, 0.5, -0.4, 0.9)),
    ("mixed", (38, 20, 12), 0.002, (0.6, 0.8)),
]


@pytest.mark.parametrize("name,n,dt,c", CASES3)
def test_fused_3d_coupled(gpu_lib, orc, name, n, dt, c):
    """3D two-field fused kernel: FP64 within 1e-12 of the oracle, FP32
    bitwise across steps per launch and rows per wave, and within FP32
    rounding of the FP64 oracle."""
    st = Stencil3d2h(7, 10, 25)
    say("coupled fused 3D", name, n)
    d64 = build(name, st, n, dt, c, F64, AUTO)
    with MultiPlan(d64) as p:
        assert p.query()[0] == FUSED
    fields = fields_for(d64, 74)
    want = orc.mrun(d64, fields, 12, 8)
    for g, w in zip(gpu_run(d64, fields, 12), want):
        assert rel_l2(g, w) <= 1e-12
    d32 = build(name, st, n, dt, c, F32, FUSED)
    f32 = [a.astype(np.float32) for a in fields]
    base = gpu_run(d32, f32, 12, K=1, TY=4)
    radius_one = name not in ("model_c", "model_c_swapped", "mixed")
    for K, TY in ((1, 8),) + (((2, 4), (2, 8)) if radius_one else ()):
        got = gpu_run(d32, f32, 12, K=K, TY=TY)
        for g, b in zip(got, base):
            assert np.array_equal(g, b), (K, TY)
    want32 = orc.mrun(d64, [a.astype(np.float64) for a in f32], 12, 8)
    for g, w in zip(base, want32):
        assert rel_l2(g.astype(np.float64), w) <= 2e-6
