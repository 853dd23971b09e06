"""Regional-grid plans (SolverSystemFDwSD) on the GPU against the oracle.

The FP64 plan runs the reference-order kernel on the box, so its field and
its box (origin, dims, empty, cutoff, time) must equal the oracle's bit for
bit, through resizes in both modes.  A box that spans the domain must also
equal the plain plan.  Argument checks run before any HIP call and are
tested without a GPU.
"""
import ctypes

import numpy as np
import pytest

from oracle import pyoracle as po
from symphas_amd import abi, inputs, models
from symphas_amd.solver import FDPlan, RegionalPlan, interior_shape


def _blob(shape_np, lo, size, value=0.5, noise_seed=None):
    psi = np.zeros(shape_np)
    sl = tuple(slice(l, l + s) for l, s in zip(reversed(lo), reversed(size)))
    psi[sl] = value
    if noise_seed is not None:
        psi[sl] += 0.05 * inputs.uniform(psi[sl].shape, seed=noise_seed, lo=0.0, hi=1.0)
    return psi


def _params(fixed=1, factor=2.0, every=3.0, eps=1e-6, rel=0.05, spl=0):
    return abi.Regional(eps, factor, every, fixed, spl, rel)


def _oparams(p):
    return po.Regional(p.cutoff_eps, p.resize_factor, p.resize_time, p.fixed_resize, 0,
                       p.dims_relative_eps)


def _run_gpu(desc, params, psi, steps):
    with RegionalPlan(desc, params) as rp:
        rp.upload(psi)
        rp.step(steps)
        return rp.download(), rp.region()


def _same_region(info, st, dim):
    assert list(info.origin)[:dim] == list(st.origin)[:dim]
    assert list(info.dims)[:dim] == list(st.dims)[:dim]
    assert info.empty == st.empty and info.cutoff == st.cutoff and info.time == st.time


def test_regional_rejects_bad_arguments_before_any_hip_call():
    lib = abi.load()
    h = ctypes.c_void_p()
    d = models.model_bb((16, 16), dtype=abi.F64)
    assert lib.sps_fdr_create(ctypes.byref(d), None, ctypes.byref(h)) == abi.E_UNSUPPORTED
    d = models.model_b((16, 16), dtype=abi.F64)
    for bad in (_params(factor=0.5), _params(every=0.0), _params(fixed=2),
                _params(rel=-1.0), _params(eps=float("nan")), _params(spl=5), _params(spl=-1)):
        assert lib.sps_fdr_create(ctypes.byref(d), ctypes.byref(bad), ctypes.byref(h)) == abi.E_INVALID
        assert b"regional parameters" in lib.sps_fd_last_error()


REF = abi.PATH_REFERENCE
CASES = [
    # name, desc factory (path: the reference order), field factory, params, steps
    ("mb2_fixed", lambda path=REF: models.model_b((64, 48), dt=0.01, dtype=abi.F64, path=path),
     lambda d: _blob(interior_shape(d), (20, 14), (12, 8), noise_seed=1), _params(every=0.04), 60),
    ("mb2_resize", lambda path=REF: models.model_b((64, 48), dt=0.01, dtype=abi.F64, path=path),
     lambda d: _blob(interior_shape(d), (50, 30), (10, 12), noise_seed=2),
     _params(fixed=0, every=0.05), 80),
    ("ac2_wrapping", lambda path=REF: models.allen_cahn((40, 36), h=0.5, dt=0.01, dtype=abi.F64,
                                                        path=path),
     lambda d: _blob(interior_shape(d), (34, 30), (9, 9), value=0.8, noise_seed=3) - 0.8,
     _params(fixed=0, every=0.02, factor=1.5), 50),
    ("mb3_resize", lambda path=REF: models.model_b((28, 24, 20), dt=0.005, dtype=abi.F64,
                                                   stencil=models.Stencil3d2h(7, 10, 25),
                                                   path=path),
     lambda d: _blob(interior_shape(d), (6, 8, 5), (7, 6, 5), noise_seed=4),
     _params(fixed=0, every=0.02), 24),
    ("mb2_lap9_bilap17", lambda path=REF: models.model_b((40, 40), dt=0.005, dtype=abi.F64,
                                                         stencil=models.Stencil2d2h(9, 6, 17),
                                                         path=path),
     lambda d: _blob(interior_shape(d), (10, 12), (8, 6), noise_seed=5), _params(every=0.03), 40),
]
FUSABLE = [c for c in CASES if c[0] != "mb2_lap9_bilap17"]


@pytest.mark.gpu
@pytest.mark.parametrize("name,mk,field,params,steps", CASES, ids=[c[0] for c in CASES])
def test_regional_fp64_bitwise_vs_oracle(gpu_lib, name, mk, field, params, steps):
    d = mk()
    psi = field(d)
    print(f"regional {name}: n {[d.n[a] for a in range(d.dim)]} steps {steps}", flush=True)
    got, info = _run_gpu(d, params, psi, steps)
    want, st = po.regional_run(d, _oparams(params), psi, steps)
    _same_region(info, st, d.dim)
    assert info.index == steps
    assert np.array_equal(got, want), np.abs(got - want).max()


@pytest.mark.gpu
@pytest.mark.parametrize("name,mk,field,params,steps", FUSABLE, ids=[c[0] for c in FUSABLE])
def test_regional_fused_fp64_vs_oracle(gpu_lib, name, mk, field, params, steps):
    """The fused box step (the canonical form) rounds differently from the
    reference order: within 1e-12 relative L2 of the oracle, same box."""
    d = mk(abi.PATH_FUSED)
    psi = field(d)
    print(f"regional fused {name}: steps {steps}", flush=True)
    got, info = _run_gpu(d, params, psi, steps)
    want, st = po.regional_run(d, _oparams(params), psi, steps)
    _same_region(info, st, d.dim)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["mb2", "ac2", "mb3"])
def test_regional_fused_spanning_box_equals_one_step_launches(gpu_lib, kind):
    """A box that spans the domain runs the streamed kernels' arithmetic: FP32
    results equal the plain plan at one step per launch bit for bit."""
    if kind == "mb2":
        d = models.model_b((64, 48), dt=0.01, dtype=abi.F32, path=abi.PATH_FUSED)
    elif kind == "ac2":
        d = models.allen_cahn((64, 40), h=0.5, dt=0.01, dtype=abi.F32, path=abi.PATH_FUSED)
    else:
        d = models.model_b((32, 24, 20), dt=0.005, dtype=abi.F32, path=abi.PATH_FUSED,
                           stencil=models.Stencil3d2h(7, 10, 25))
    psi = inputs.uniform(interior_shape(d), seed=23).astype(np.float32)
    print(f"regional fused spanning {kind}", flush=True)
    got, info = _run_gpu(d, _params(factor=60.0, every=0.05), psi, 60)
    assert [info.dims[a] for a in range(d.dim)] == [d.n[a] + 6 for a in range(d.dim)]
    with FDPlan(d) as p:
        if d.dim == 2:
            p.tune(1, 16, 0, 2)
        else:
            p.tune(1, 4, 16, 0)
        p.upload(psi)
        p.step(60)
        want = p.download()
    assert np.array_equal(got, want), np.abs(got - want).max()


MULTI = [
    # name, desc factory, field factory, params (resizes cut the runs of steps)
    ("mb2_free", lambda dt: models.model_b((96, 80), dt=0.01, dtype=dt, path=abi.PATH_FUSED),
     lambda d: _blob(interior_shape(d), (30, 24), (14, 10), noise_seed=11),
     dict(fixed=0, every=0.13)),
    ("mb2_fixed", lambda dt: models.model_b((64, 48), dt=0.01, dtype=dt, path=abi.PATH_FUSED),
     lambda d: _blob(interior_shape(d), (20, 14), (12, 8), noise_seed=1), dict(every=0.05)),
    ("ac2_wrapping", lambda dt: models.allen_cahn((72, 64), h=0.5, dt=0.01, dtype=dt,
                                                  path=abi.PATH_FUSED),
     lambda d: _blob(interior_shape(d), (60, 50), (16, 16), value=0.8, noise_seed=3) - 0.8,
     dict(fixed=0, every=0.07, factor=1.5)),
    # boxes wider than one strip: 152..196 columns (even: two values per lane
    # in FP32) and 79..85 columns (odd: one value per lane)
    ("mb2_wide", lambda dt: models.model_b((256, 160), dt=0.01, dtype=dt, path=abi.PATH_FUSED),
     lambda d: _blob(interior_shape(d), (60, 40), (70, 30), noise_seed=5),
     dict(fixed=0, every=0.11)),
    ("ac2_wide_odd", lambda dt: models.allen_cahn((200, 120), h=0.5, dt=0.01, dtype=dt,
                                                  path=abi.PATH_FUSED),
     lambda d: _blob(interior_shape(d), (40, 30), (47, 25), value=0.8, noise_seed=5) - 0.8,
     dict(fixed=0, every=0.09, factor=1.5)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [abi.F32, abi.F64], ids=["f32", "f64"])
@pytest.mark.parametrize("name,mk,field,kw", MULTI, ids=[c[0] for c in MULTI])
def test_regional_multi_step_equals_one_step_launches(gpu_lib, name, mk, field, kw, dtype):
    """A box whose layer is frozen runs up to 4 fused steps per launch
    between updates: the same field, box and time as one step per launch,
    bit for bit, whatever the run lengths (steps per call, resizes)."""
    d = mk(dtype)
    psi = field(d)
    if dtype == abi.F32:
        psi = psi.astype(np.float32)
    calls = [7, 13, 1, 4, 22, 3, 10]
    results = {}
    for spl in (1, 4, 0):
        print(f"regional multi-step {name} dtype {dtype} steps_per_launch {spl}", flush=True)
        with RegionalPlan(d, _params(spl=spl, **kw)) as rp:
            rp.upload(psi)
            for n in calls:
                rp.step(n)
            info = rp.region()
            results[spl] = (rp.download(), list(info.origin)[:2], list(info.dims)[:2],
                            info.time, info.index)
    one = results[1]
    assert 0 < one[2][0] - 6 < d.n[0] and 0 < one[2][1] - 6 < d.n[1]  # frozen layer
    for spl in (4, 0):
        got = results[spl]
        assert got[1:] == one[1:]
        assert np.array_equal(got[0], one[0]), np.abs(got[0] - one[0]).max()


@pytest.mark.gpu
def test_regional_spanning_box_equals_plain_plan(gpu_lib):
    d = models.model_b((48, 40), dt=0.01, dtype=abi.F64, path=abi.PATH_REFERENCE)
    psi = inputs.uniform(interior_shape(d), seed=21)
    got, info = _run_gpu(d, _params(factor=40.0, every=0.05), psi, 100)
    assert [info.dims[a] for a in range(2)] == [54, 46]
    with FDPlan(d) as p:
        p.upload(psi)
        p.step(100)
        want = p.download()
    assert np.array_equal(got, want)


@pytest.mark.gpu
def test_regional_fp32_against_fp64_oracle(gpu_lib):
    d64 = models.model_b((64, 48), dt=0.01, dtype=abi.F64)
    d32 = models.model_b((64, 48), dt=0.01, dtype=abi.F32)
    psi = _blob(interior_shape(d64), (20, 14), (12, 8), noise_seed=1).astype(np.float32)
    params = _params(every=0.1)
    got, info = _run_gpu(d32, params, psi, 50)
    want, st = po.regional_run(d64, _oparams(params), psi.astype(np.float64), 50)
    assert list(info.dims)[:2] == list(st.dims)[:2]
    assert list(info.origin)[:2] == list(st.origin)[:2]
    err = np.linalg.norm(got.astype(np.float64) - want) / np.linalg.norm(want)
    assert err < 1e-6, err


@pytest.mark.gpu
def test_regional_padded_download_and_no_active_site(gpu_lib):
    d = models.model_b((32, 24), dt=0.01, dtype=abi.F64)
    psi = _blob(interior_shape(d), (4, 6), (8, 5), noise_seed=7)
    with RegionalPlan(d) as rp:
        rp.upload(psi)
        assert np.array_equal(rp.download(), psi)  # before the first update
        rp.step(10)
        inner = rp.download()
        padded = rp.download(abi.LAYOUT_PADDED)
    assert np.array_equal(padded, np.pad(inner, 3, mode="wrap"))
    flat = np.full(interior_shape(d), 0.25)
    with RegionalPlan(d) as rp:
        rp.upload(flat)
        rp.step(5)
        info = rp.region()
        out = rp.download()
    assert list(info.dims) == [0, 0, 0]
    assert np.all(out == 0.25)
