"""Plot data files (GNU matrix, CSV) and XDR files written and read on the
GPU against the oracle's restatements of the reference's writers
(io/src/writegp.cpp save_grid_plotting, io/src/writecsv.cpp,
io/src/writexdr.cpp save_xdr, io/src/readxdr.cpp), which format with the C
library and encode XDR byte by byte."""
import numpy as np
import pytest

from symphas_amd import abi, inputs, models
from symphas_amd.models import Stencil2d2h, Stencil3d2h
from symphas_amd.solver import FDPlan

pytestmark = pytest.mark.gpu

SPECIAL = [0.0, -0.0, 5e-324, -2.2250738585072014e-308, 1e-300, 1e300, -1.7976931348623157e308,
           float("inf"), -float("inf"), float("nan"), 1.0000005, 9.9999995e-1,
           0.5, 1.25e-5, 123456789012345.0, 2.5e-16, -7.0]


def say(*a):
    print("[case]", *a, flush=True)


def desc_for(n, dtype):
    st = Stencil2d2h(5, 6, 13) if len(n) == 2 else Stencil3d2h(7, 10, 25)
    return models.model_b(n, h=0.7, dt=0.01, stencil=st, dtype=dtype, path=abi.PATH_AUTO)


def field(n, dtype, seed):
    psi = inputs.uniform(tuple(reversed(n)), seed=seed, lo=-2.0, hi=2.0)
    psi.ravel()[:len(SPECIAL)] = SPECIAL
    return psi.astype(np.float32 if dtype == abi.F32 else np.float64)


CASES = [((37, 11), abi.F64), ((64, 24), abi.F32), ((13, 7, 9), abi.F64), ((16, 8, 6), abi.F32)]


@pytest.mark.parametrize("fmt", [abi.PLOT_GNU, abi.PLOT_CSV])
@pytest.mark.parametrize("n,dtype", CASES)
def test_plot_text_is_byte_identical(gpu_lib, orc, n, dtype, fmt):
    say("plot text", n, dtype, fmt)
    psi = field(n, dtype, 53)
    d64 = desc_for(n, abi.F64)
    origin = [0.5, -1.25, 2.0][:len(n)]
    with FDPlan(desc_for(n, dtype)) as p:
        p.upload(psi)
        got = p.plot_text(fmt, origin=origin)
    want = orc.plot_text(d64, orc.pad(d64, psi.astype(np.float64)), fmt, origin=origin)
    assert got == want


def test_plot_text_chunk_boundary_inside_a_slice(gpu_lib, orc):
    """300 x 300 x 200 FP32 (18M values): the second device chunk starts in
    the middle of a z slice, so slice headers and row prefixes must follow
    the global row index."""
    n = (300, 300, 200)
    say("gnu plot text 300x300x200")
    psi = inputs.uniform((200, 300, 300), seed=54, lo=-1.0, hi=1.0).astype(np.float32)
    d64 = desc_for(n, abi.F64)
    with FDPlan(desc_for(n, abi.F32)) as p:
        p.upload(psi)
        got = p.plot_text(abi.PLOT_GNU)
    want = orc.plot_text(d64, orc.pad(d64, psi.astype(np.float64)), abi.PLOT_GNU)
    assert got == want


@pytest.mark.parametrize("n,dtype", CASES)
def test_xdr_is_byte_identical_and_reads_back(gpu_lib, orc, n, dtype):
    say("xdr", n, dtype)
    psi = field(n, dtype, 55)
    d64 = desc_for(n, abi.F64)
    origin = [0.5, -1.25, 2.0][:len(n)]
    want = orc.xdr_write(d64, orc.pad(d64, psi.astype(np.float64)), 21, origin=origin)
    with FDPlan(desc_for(n, dtype)) as p:
        p.upload(psi)
        got = p.checkpoint_xdr(index=21, origin=origin)
        assert got == want
        p.upload(np.zeros_like(psi))
        assert p.read_checkpoint_xdr(got) == 21
        back = p.download()
    assert np.array_equal(back, psi, equal_nan=True)
    assert np.array_equal(np.signbit(back), np.signbit(psi))


def test_xdr_fp64_file_into_fp32_plan(gpu_lib, orc):
    n = (40, 30)
    say("xdr fp64 -> fp32")
    psi = inputs.uniform((30, 40), seed=56, lo=-2.0, hi=2.0)
    d64 = desc_for(n, abi.F64)
    data = orc.xdr_write(d64, orc.pad(d64, psi), 4)
    with FDPlan(desc_for(n, abi.F32)) as p:
        assert p.read_checkpoint_xdr(data) == 4
        assert np.array_equal(p.download(), psi.astype(np.float32))


def test_xdr_two_chunks_full_size(gpu_lib):
    """36M FP32 values: two staged chunks on the device.  The body must be the
    values as big-endian doubles; the header is the XDR encoding of dim,
    counts, bounds and index (h = 0.7)."""
    import struct
    n = (6000, 6000)
    say("xdr 6000x6000")
    psi = inputs.uniform((6000, 6000), seed=57, lo=-1.0, hi=1.0).astype(np.float32)
    with FDPlan(desc_for(n, abi.F32)) as p:
        p.upload(psi)
        got = p.checkpoint_xdr(index=1)
        p.upload(np.zeros_like(psi))
        p.read_checkpoint_xdr(got)
        back = p.download()
    bounds = struct.pack(">dddd", 0.0, 5999 * 0.7, 0.0, 5999 * 0.7)
    header = struct.pack(">iii", 2, 6000, 6000) + bounds + struct.pack(">ii", 1, 2) + bounds
    assert got[:len(header)] == header
    body = np.frombuffer(got[len(header):], dtype=">f8")
    assert np.array_equal(body, psi.ravel().astype(np.float64))
    assert np.array_equal(back, psi)


def test_xdr_read_errors(gpu_lib, orc):
    n = (16, 12)
    psi = inputs.uniform((12, 16), seed=58)
    d64 = desc_for(n, abi.F64)
    data = orc.xdr_write(d64, orc.pad(d64, psi), 2)
    with FDPlan(desc_for(n, abi.F64)) as p:
        with pytest.raises(abi.SpsError):
            p.read_checkpoint_xdr(data[:-8])
        with pytest.raises(abi.SpsError):
            p.read_checkpoint_xdr(data[:10])
    with FDPlan(desc_for((16, 14), abi.F64)) as p:
        with pytest.raises(abi.SpsError):
            p.read_checkpoint_xdr(data)
