"""Parity of the HIP step (through the C ABI) with the CPU oracle.

Bars:
  * reference-order path, FP64: bit-identical to the oracle (same groupings,
    same term order, no fused multiply-add);
  * fused path, FP64: relative L2 <= 1e-12 after the stated steps (the bar
    the FP64 GPU path must meet against the CPU SolverFT after 1000 steps);
  * fused path, FP32: relative L2 against the FP64 oracle within the stated
    tolerance (FP32 storage and arithmetic);
  * temporal blocking and graph replay change the schedule only: bitwise
    equal to single-step launches.
"""
import sys

import numpy as np
import pytest

from symphas_amd import abi, inputs, models
from symphas_amd.models import Stencil2d2h, Stencil3d2h, TermSpec
from symphas_amd.solver import FDPlan

pytestmark = pytest.mark.gpu

F32, F64 = abi.F32, abi.F64
REF, FUSED, AUTO = abi.PATH_REFERENCE, abi.PATH_FUSED, abi.PATH_AUTO


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def gpu_run(desc, psi, nsteps, K=0, TY=0, chunks=None, TZ=0, VEC=0):
    with FDPlan(desc) as p:
        if K or TY or TZ or VEC:
            p.tune(K, TY, TZ, VEC)
        p.upload(psi.astype(p.dtype))
        if chunks:
            for c in chunks:
                p.step(c)
        else:
            p.step(nsteps)
        return p.download()


def oracle_run(orc, desc, psi, nsteps, nthreads=8):
    return orc.unpad(desc, orc.run(desc, psi.astype(np.float64), nsteps, nthreads))


def field(desc, seed, lo=-0.1, hi=0.1):
    shape = tuple(int(desc.n[a]) for a in reversed(range(desc.dim)))
    return inputs.uniform(shape, seed=seed, lo=lo, hi=hi)


# --------------------------------------------------------------- config 1
def test_cfg1_allen_cahn_reference_bitwise(gpu_lib, orc):
    desc = models.allen_cahn((64, 64), h=0.5, dt=0.01, dtype=F64, path=REF)
    psi = field(desc, seed=1)
    say("cfg1 AC 64^2 FP64 reference path, 1000 steps")
    got = gpu_run(desc, psi, 1000)
    want = oracle_run(orc, desc, psi, 1000)
    assert rel_l2(got, want) <= 1e-12
    assert np.array_equal(got, want), float(np.abs(got - want).max())


def test_cfg1_allen_cahn_fused_fp64(gpu_lib, orc):
    desc = models.allen_cahn((64, 64), h=0.5, dt=0.01, dtype=F64, path=FUSED)
    psi = field(desc, seed=1)
    say("cfg1 AC 64^2 FP64 fused path, 1000 steps")
    with FDPlan(desc) as p:
        assert p.query()[0] == FUSED
    got = gpu_run(desc, psi, 1000)
    want = oracle_run(orc, desc, psi, 1000)
    assert rel_l2(got, want) <= 1e-12


@pytest.mark.parametrize("stencil", [Stencil2d2h(5, 6, 13), Stencil2d2h(9, 6, 13),
                                     Stencil2d2h(9, 8, 17), Stencil2d2h(5, 16, 21)])
def test_model_b_reference_bitwise(gpu_lib, orc, stencil):
    desc = models.model_b((64, 48), h=1.0, dt=0.01, stencil=stencil, dtype=F64, path=REF)
    psi = field(desc, seed=2)
    say("Model B 64x48 FP64 reference", stencil)
    got = gpu_run(desc, psi, 100)
    want = oracle_run(orc, desc, psi, 100)
    assert np.array_equal(got, want), float(np.abs(got - want).max())


def test_model_b_fused_fp64_1000_steps(gpu_lib, orc):
    desc = models.model_b((64, 64), h=1.0, dt=0.01, dtype=F64, path=FUSED)
    psi = field(desc, seed=3)
    say("Model B 64^2 FP64 fused 1000 steps")
    got = gpu_run(desc, psi, 1000)
    want = oracle_run(orc, desc, psi, 1000)
    assert rel_l2(got, want) <= 1e-12


# --------------------------------------------------------------- stencils
def _apply_cases():
    for L in (5, 9):
        yield Stencil2d2h(L, 6, 13), abi.OP_LAP, 0, 0
    for B in (13, 17, 21):
        yield Stencil2d2h(5, 6, B), abi.OP_BILAP, 0, 0
    for G in (6, 8, 12, 16):
        for ax in (0, 1):
            for comp in ((0, 1) if G == 6 else (0,)):
                yield Stencil2d2h(5, G, 13), abi.OP_GRADLAP, ax, comp
    for L in (7, 15, 19, 21, 27):
        yield Stencil3d2h(L, 10, 25), abi.OP_LAP, 0, 0
    for B in (21, 25, 41, 52, 57):
        yield Stencil3d2h(7, 10, B), abi.OP_BILAP, 0, 0
    for G in (10, 12, 28, 36, 40):
        for ax in (0, 1, 2):
            yield Stencil3d2h(7, G, 25), abi.OP_GRADLAP, ax, 0


@pytest.mark.parametrize("stencil,op,axis,comp", list(_apply_cases()))
def test_stencil_apply_bitwise(gpu_lib, orc, stencil, op, axis, comp):
    shape = (13, 10) if stencil.dim == 2 else (9, 7, 8)
    desc = models.make_desc(shape, stencil, 0.7, 0.01, [], dtype=F64, path=REF)
    psi = field(desc, seed=4, lo=-1, hi=1)
    say("apply", stencil, op, axis, comp)
    with FDPlan(desc) as p:
        p.upload(psi)
        got = p.apply(op, axis, comp)
    g = orc.refresh(desc, orc.pad(desc, psi))
    want = orc.unpad(desc, orc.apply(desc, op, axis, comp, g))
    assert np.array_equal(got, want), float(np.abs(got - want).max())


def test_equation_general_terms_bitwise(gpu_lib, orc):
    terms = [TermSpec(abi.OP_BILAP, -0.5, (0, 0, 0, 1.0)),       # bilap(psi^3)
             TermSpec(abi.OP_LAP, 1.3, (0.2, -1.0, 0.5)),          # lap(P2)
             TermSpec(abi.OP_GRADLAP, 0.7, (0, 0, 1.0), axis=1),   # gradlap_y(psi^2)
             TermSpec(abi.OP_ID, 1.0, (0.1, 1.0, 0, -1.0, 0, 0, 0.01))]
    for stencil, shape in ((Stencil2d2h(9, 12, 17), (17, 11)), (Stencil3d2h(27, 36, 41), (8, 9, 7))):
        desc = models.make_desc(shape, stencil, 0.9, 0.005, terms, dtype=F64, path=REF)
        psi = field(desc, seed=5, lo=-0.9, hi=0.9)
        say("equation general", stencil)
        with FDPlan(desc) as p:
            p.upload(psi)
            got = p.equation()
            gotp = p.equation(abi.LAYOUT_PADDED)
        g = orc.refresh(desc, orc.pad(desc, psi))
        wantp = orc.equation(desc, g)
        want = orc.unpad(desc, wantp)
        assert np.array_equal(got, want)
        assert np.array_equal(gotp, wantp)  # ghosts of dframe stay zero
        got = gpu_run(desc, psi, 5)
        assert np.array_equal(got, oracle_run(orc, desc, psi, 5))


# --------------------------------------------------------------- 3D
@pytest.mark.parametrize("stencil", [Stencil3d2h(7, 10, 25), Stencil3d2h(15, 10, 25)])
def test_model_b_3d_reference_bitwise(gpu_lib, orc, stencil):
    desc = models.model_b((12, 10, 8), h=1.0, dt=0.005, stencil=stencil, dtype=F64, path=REF)
    psi = field(desc, seed=6, lo=-0.05, hi=0.05)
    say("Model B 3D FP64 reference", stencil)
    got = gpu_run(desc, psi, 20)
    assert np.array_equal(got, oracle_run(orc, desc, psi, 20))


def test_model_b_3d_fused_fp64(gpu_lib, orc):
    desc = models.model_b((40, 36, 20), h=1.0, dt=0.005, stencil=Stencil3d2h(7, 10, 25),
                          dtype=F64, path=FUSED)
    psi = field(desc, seed=7, lo=-0.05, hi=0.05)
    say("Model B 3D FP64 fused 200 steps")
    got = gpu_run(desc, psi, 200)
    assert rel_l2(got, oracle_run(orc, desc, psi, 200)) <= 1e-12


def test_allen_cahn_3d_fused(gpu_lib, orc):
    desc = models.allen_cahn((24, 20, 70), h=1.0, dt=0.01, stencil=Stencil3d2h(7, 10, 25),
                             dtype=F64, path=FUSED)
    psi = field(desc, seed=8)
    say("AC 3D FP64 fused 100 steps")
    got = gpu_run(desc, psi, 100)
    assert rel_l2(got, oracle_run(orc, desc, psi, 100)) <= 1e-12


# --------------------------------------------------------------- FP32
@pytest.mark.parametrize("model,shape,steps,tol", [
    ("mb", (128, 128), 500, 2e-5),
    ("ac", (128, 96), 500, 2e-6),
    ("mb", (256, 64), 100, 1e-5),
])
def test_fused_fp32_vs_fp64_oracle(gpu_lib, orc, model, shape, steps, tol):
    build = models.model_b if model == "mb" else models.allen_cahn
    h = 1.0 if model == "mb" else 0.5
    desc32 = build(shape, h=h, dt=0.01, dtype=F32, path=FUSED)
    desc64 = build(shape, h=h, dt=0.01, dtype=F64, path=REF)
    psi = field(desc64, seed=9)
    say("FP32 fused", model, shape, steps)
    got = gpu_run(desc32, psi.astype(np.float32), steps).astype(np.float64)
    want = oracle_run(orc, desc64, psi.astype(np.float32).astype(np.float64), steps)
    err = rel_l2(got, want)
    print(f"FP32 {model} {shape} {steps} steps: rel L2 {err:.3e}, max {np.abs(got-want).max():.3e}")
    assert err <= tol


def test_fused_fp32_3d_vs_oracle(gpu_lib, orc):
    st = Stencil3d2h(7, 10, 25)
    desc32 = models.model_b((32, 32, 32), h=1.0, dt=0.005, stencil=st, dtype=F32, path=FUSED)
    desc64 = models.model_b((32, 32, 32), h=1.0, dt=0.005, stencil=st, dtype=F64, path=REF)
    psi = field(desc64, seed=10, lo=-0.05, hi=0.05).astype(np.float32)
    say("FP32 3D fused 32^3 100 steps")
    got = gpu_run(desc32, psi, 100).astype(np.float64)
    want = oracle_run(orc, desc64, psi.astype(np.float64), 100)
    assert rel_l2(got, want) <= 1e-5


# --------------------------------------------------------------- schedule
@pytest.mark.parametrize("dtype", [F32, F64])
@pytest.mark.parametrize("model", ["mb", "ac"])
def test_temporal_blocking_is_bitwise(gpu_lib, dtype, model):
    build = models.model_b if model == "mb" else models.allen_cahn
    desc = build((264, 200), h=1.0, dt=0.01, dtype=dtype, path=FUSED)
    psi = field(desc, seed=11)
    say("temporal blocking", model, dtype)
    base = gpu_run(desc, psi, 24, K=1, TY=32)
    wide = 4 if dtype == F32 else 2
    for vec in (wide, wide // 2):
        for K in (1, 2, 3, 4):
            for TY in (4, 7, 12, 50, 64, 200, 512):
                got = gpu_run(desc, psi, 24, K=K, TY=TY, VEC=vec)
                assert np.array_equal(got, base), (vec, K, TY)


def test_narrow_lanes_take_widths_of_two(gpu_lib, orc):
    """x extent 130 (not a multiple of 4): the fused FP32 path runs with two
    values per lane; every steps-per-launch gives the same bits and the result
    follows the FP64 oracle."""
    desc = models.model_b((130, 48), h=1.0, dt=0.01, dtype=F32, path=FUSED)
    psi = field(desc, seed=24)
    say("narrow lanes, X=130")
    base = gpu_run(desc, psi, 12, K=1, TY=16)
    for K in (2, 3, 4):
        assert np.array_equal(gpu_run(desc, psi, 12, K=K, TY=32), base), K
    with FDPlan(desc) as p:
        with pytest.raises(abi.SpsError):
            p.tune(0, 0, 0, 4)
    d64 = models.model_b((130, 48), h=1.0, dt=0.01, dtype=F64, path=REF)
    want = oracle_run(orc, d64, psi.astype(np.float32).astype(np.float64), 12)
    assert rel_l2(base.astype(np.float64), want) <= 1e-6


@pytest.mark.parametrize("dtype", [F32, F64])
@pytest.mark.parametrize("model", ["mb", "ac"])
def test_3d_chunking_and_fused_steps_are_bitwise(gpu_lib, dtype, model):
    """Every (steps per launch, rows, planes per chunk) shape gives the bits of
    one step per launch; 11 steps leave a remainder launch for K = 2."""
    st = Stencil3d2h(7, 10, 25)
    if model == "mb":
        desc = models.model_b((70, 33, 21), h=1.0, dt=0.005, stencil=st, dtype=dtype, path=FUSED)
    else:
        desc = models.allen_cahn((70, 33, 21), h=1.0, dt=0.01, stencil=st, dtype=dtype, path=FUSED)
    psi = field(desc, seed=12)
    say("3D z-chunking, rows and fused steps", model, dtype)
    base = gpu_run(desc, psi, 11, K=1, TY=8, TZ=32)
    shapes = [(1, 4), (1, 6), (1, 8), (1, 12), (1, 16), (2, 4)] + ([(2, 6), (2, 8), (2, 12)] if dtype == F32 else [])
    for K, rows in shapes:
        for TZ in (1, 4, 8, 20, 64):
            got = gpu_run(desc, psi, 11, K=K, TY=rows, TZ=TZ)
            assert np.array_equal(got, base), (K, rows, TZ)


def test_3d_fused_steps_match_oracle(gpu_lib, orc):
    """Two steps per launch against the FP64 oracle (FP64 fused <= 1e-12)."""
    desc = models.model_b((40, 36, 20), h=1.0, dt=0.005, stencil=Stencil3d2h(7, 10, 25),
                          dtype=F64, path=FUSED)
    psi = field(desc, seed=23)
    say("3D K=2 vs oracle")
    got = gpu_run(desc, psi, 8, K=2, TY=4, TZ=8)
    want = oracle_run(orc, desc, psi, 8)
    assert rel_l2(got, want) <= 1e-12


def test_graph_replay_matches_eager(gpu_lib):
    desc = models.model_b((256, 256), h=1.0, dt=0.01, dtype=F32, path=FUSED)
    psi = field(desc, seed=13)
    say("graph replay vs eager")
    eager = gpu_run(desc, psi, 200, chunks=[1] * 200)
    graph = gpu_run(desc, psi, 200)
    assert np.array_equal(eager, graph)
    # replays interleaved with other eager work and host synchronisation
    with FDPlan(desc) as p:
        p.upload(psi.astype(np.float32))
        for _ in range(3):
            p.step(64)
            p.reduce()
            p.step(3)
        mixed = p.download()
    assert np.array_equal(mixed, gpu_run(desc, psi, 201, chunks=[1] * 201))


# --------------------------------------------------------------- sizes
@pytest.mark.parametrize("shape", [(3, 3), (5, 7), (7, 3), (9, 17), (66, 6)])
def test_small_and_ragged_2d(gpu_lib, orc, shape):
    desc = models.model_b(shape, h=1.0, dt=0.01, dtype=F64, path=REF)
    psi = field(desc, seed=14)
    say("ragged 2D", shape)
    assert np.array_equal(gpu_run(desc, psi, 10), oracle_run(orc, desc, psi, 10))
    # AUTO fuses when the x extent is a multiple of the narrowest lane width
    # (FP64: one value per lane, so always; FP32: two values per lane)
    desc = models.model_b(shape, h=1.0, dt=0.01, dtype=F64, path=AUTO)
    with FDPlan(desc) as p:
        assert p.query()[0] == FUSED
    got = gpu_run(desc, psi, 10)
    assert rel_l2(got, oracle_run(orc, desc, psi, 10)) <= 1e-12
    d32 = models.model_b(shape, h=1.0, dt=0.01, dtype=F32, path=AUTO)
    with FDPlan(d32) as p:
        assert p.query()[0] == (FUSED if shape[0] % 2 == 0 else REF)


@pytest.mark.parametrize("shape", [(3, 3, 3), (5, 4, 3), (65, 3, 17)])
def test_small_and_ragged_3d(gpu_lib, orc, shape):
    st = Stencil3d2h(7, 10, 25)
    for path in (REF, FUSED):
        desc = models.model_b(shape, h=1.0, dt=0.005, stencil=st, dtype=F64, path=path)
        psi = field(desc, seed=15)
        say("ragged 3D", shape, path)
        got = gpu_run(desc, psi, 6)
        want = oracle_run(orc, desc, psi, 6)
        if path == REF:
            assert np.array_equal(got, want)
        else:
            assert rel_l2(got, want) <= 1e-12


def test_fused_unsupported_width_raises(gpu_lib):
    desc = models.model_b((31, 32), dtype=F32, path=FUSED)  # odd: no lane width fits
    with pytest.raises(abi.SpsError) as ei:
        FDPlan(desc)
    assert ei.value.status == abi.E_UNSUPPORTED


# --------------------------------------------------------------- sums
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("dtype", [F64, F32])
@pytest.mark.parametrize("energy", [None, (0.7, 1.9, 2.5, 0, 0), (0.7, 1.9, 2.5, 1, 0),
                                    (1.3, 0.4, 0.6, 1, 1), (1.0, 1.0, 1.0, 0, 1)])
def test_reductions_match_oracle(gpu_lib, orc, dim, dtype, energy):
    """Mass and the free-energy functional (c1, c2, kappa, gradient, mean)
    against the oracle's sequential FP64 sums (whose functional the CPU tests
    pin by closed forms and summation by parts)."""
    if dim == 2:
        desc = models.model_b((96, 80), h=0.8, dtype=dtype)
    else:
        desc = models.model_b((20, 18, 16), h=0.8, stencil=Stencil3d2h(7, 10, 25), dtype=dtype)
    psi = field(desc, seed=16, lo=-0.7, hi=0.9)
    if dtype == F32:
        psi = psi.astype(np.float32)
    say("reductions", dim, dtype, energy)
    with FDPlan(desc) as p:
        p.upload(psi)
        if energy is not None:
            p.set_energy(abi.Energy(*energy))
        m, e = p.reduce()
    g = orc.refresh(desc, orc.pad(desc, psi.astype(np.float64)))
    args = {} if energy is None else dict(zip(("c1", "c2", "kappa", "gradient", "mean"), energy))
    want = orc.energy(desc, g, **args)
    assert abs(m - orc.mass(desc, g)) <= 1e-12 * np.abs(psi.astype(np.float64)).sum()
    assert abs(e - want) <= 1e-12 * abs(want)


def test_energy_rejects_bad_functional(gpu_lib):
    desc = models.model_b((16, 16), dtype=F64)
    with FDPlan(desc) as p:
        for bad in [(1.0, 1.0, float("nan"), 0, 0), (1.0, 1.0, 1.0, 2, 0), (1.0, 1.0, 1.0, 0, 3)]:
            with pytest.raises(abi.SpsError) as ei:
                p.set_energy(abi.Energy(*bad))
            assert ei.value.status == abi.E_INVALID


def test_fp32_model_b_mass_drift_and_energy(gpu_lib):
    desc = models.model_b((512, 512), h=1.0, dt=0.01, dtype=F32, path=FUSED)
    psi = field(desc, seed=17).astype(np.float32)
    say("FP32 Model B mass drift 512^2 x 2000")
    with FDPlan(desc) as p:
        p.upload(psi)
        m0, e0 = p.reduce()
        energies = [e0]
        for _ in range(4):
            p.step(500)
            energies.append(p.reduce()[1])
        m1, _ = p.reduce()
    drift = abs(m1 - m0) / np.abs(psi.astype(np.float64)).sum()
    print(f"mass drift {drift:.3e}")
    assert drift <= 1e-6
    assert all(b < a for a, b in zip(energies, energies[1:]))


# --------------------------------------------------------------- layout
def test_padded_upload_download(gpu_lib, orc):
    desc = models.model_b((20, 12), dtype=F64, path=REF)
    psi = field(desc, seed=18)
    padded = orc.pad(desc, psi)
    padded[:3, :] = 99.0  # stale ghosts must be ignored
    say("padded layout")
    with FDPlan(desc) as p:
        p.upload(padded, abi.LAYOUT_PADDED)
        out = p.download(abi.LAYOUT_PADDED)
        assert np.array_equal(p.download(), psi)
    assert np.array_equal(out, np.pad(psi, 3, mode="wrap"))


def test_state_and_argument_errors(gpu_lib):
    desc = models.model_b((32, 32), dtype=F32)
    say("error paths")
    with FDPlan(desc) as p:
        with pytest.raises(abi.SpsError) as ei:
            p.step(1)
        assert ei.value.status == abi.E_STATE
        with pytest.raises(abi.SpsError) as ei:
            p.tune(5, 0)
        with pytest.raises(abi.SpsError) as ei:
            p.tune(0, 0, 8)
        assert ei.value.status == abi.E_INVALID
        for vec in (1, 3, 8):
            with pytest.raises(abi.SpsError) as ei:
                p.tune(0, 0, 0, vec)
            assert ei.value.status == abi.E_INVALID, vec
        for rows in (3, 513):
            with pytest.raises(abi.SpsError) as ei:
                p.tune(0, rows)
            assert ei.value.status == abi.E_INVALID, rows
        p.tune(0, 50)
        p.upload(np.zeros((32, 32), np.float32))
        lib = abi.load()
        out = np.empty((32, 32), np.float32)
        import ctypes
        assert lib.sps_fd_download(p._h, out.ctypes.data_as(ctypes.c_void_p), 7) == abi.E_INVALID
        assert lib.sps_fd_apply(p._h, abi.OP_ID, 0, 0, out.ctypes.data_as(ctypes.c_void_p), 0) == abi.E_INVALID
    d3 = models.model_b((16, 16, 16), dtype=F64, stencil=Stencil3d2h(7, 10, 25), path=FUSED)
    with FDPlan(d3) as p:
        for K, rows, planes, vec in ((2, 16, 0, 0), (2, 8, 0, 0), (2, 6, 0, 0), (3, 4, 0, 0),
                                     (1, 10, 0, 0), (1, 8, 2000, 0), (1, 8, 0, 2)):
            with pytest.raises(abi.SpsError) as ei:
                p.tune(K, rows, planes, vec)
            assert ei.value.status == abi.E_INVALID, (K, rows, planes, vec)
        p.tune(2, 4, 8)


# --------------------------------------------------------------- full size
def test_cfg2_full_size_two_steps(gpu_lib, orc):
    """4096^2 Model B, FP32 fused: two steps against the FP64 oracle."""
    desc32 = models.model_b((4096, 4096), h=1.0, dt=0.01, dtype=F32, path=FUSED)
    desc64 = models.model_b((4096, 4096), h=1.0, dt=0.01, dtype=F64, path=REF)
    psi = field(desc64, seed=19).astype(np.float32)
    say("cfg2 4096^2 FP32 two steps")
    got = gpu_run(desc32, psi, 2).astype(np.float64)
    want = oracle_run(orc, desc64, psi.astype(np.float64), 2)
    assert rel_l2(got, want) <= 1e-6


# --------------------------------------------------------------- autotuner
def test_autotune_keeps_field_and_caches(gpu_lib, tmp_path, monkeypatch):
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("SPS_FD_TUNE_CACHE", str(cache))
    desc = models.model_b((512, 384), h=1.0, dt=0.01, dtype=F32, path=FUSED)
    psi = field(desc, seed=20).astype(np.float32)
    say("autotune 2D")
    base = gpu_run(desc, psi, 24, K=1, TY=32)
    with FDPlan(desc) as p:
        p.upload(psi)
        choice = p.autotune(use_cache=False)
        assert np.array_equal(p.download(), psi)
        p.step(24)
        assert np.array_equal(p.download(), base)
    assert choice[0] in (1, 2, 3, 4) and 4 <= choice[1] <= 512
    assert choice[2] == 0 and choice[3] in (4, 2)
    text = cache.read_text()
    assert "dim=2 n=512x384x1 dtype=f32" in text
    with FDPlan(desc) as p:
        p.upload(psi)
        assert p.autotune(use_cache=True) == choice


def test_autotune_3d(gpu_lib):
    desc = models.model_b((96, 64, 40), h=1.0, dt=0.005, stencil=Stencil3d2h(7, 10, 25),
                          dtype=F32, path=FUSED)
    psi = field(desc, seed=21).astype(np.float32)
    say("autotune 3D")
    base = gpu_run(desc, psi, 6, TY=8, TZ=32)
    with FDPlan(desc) as p:
        p.upload(psi)
        k, rows, planes, vec = p.autotune(use_cache=False)
        assert (k, rows) in ((1, 4), (1, 6), (1, 8), (1, 12), (1, 16), (2, 4), (2, 6), (2, 8),
                             (2, 12))
        assert planes in (8, 16, 32, 40) and vec == 1
        assert np.array_equal(p.download(), psi)
        p.step(6)
        assert np.array_equal(p.download(), base)


def test_autotune_reference_path_is_noop(gpu_lib):
    desc = models.model_b((30, 20), dtype=F64, path=REF)
    with FDPlan(desc) as p:
        p.upload(field(desc, seed=22))
        assert p.autotune()[0] == 1
