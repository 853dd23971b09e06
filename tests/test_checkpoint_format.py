"""The GNU checkpoint format on the CPU: the exact decimal core
(symphas_amd/csrc/decimal.hpp) against the C library, and the oracle's
restatement of the reference's writer and reader against an independent
formatter (Python's correctly rounded '%.10E', David Gay's dtoa)."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from symphas_amd import inputs, models
from symphas_amd.models import Stencil3d2h

ROOT = Path(__file__).resolve().parents[1]


def test_decimal_core_matches_c_library(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    exe = tmp_path / "decimal_check"
    res = subprocess.run([hipcc, "-O2", "-std=c++17", "--offload-arch=gfx950", "-x", "hip",
                          str(ROOT / "tests/host/decimal_check.cpp"), "-o", str(exe)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-3000:]
    res = subprocess.run([str(exe), "400000", "11"], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    words = res.stdout.split()
    checked, formatted = int(words[1]), int(words[3])
    assert checked > 390000 and formatted > 0.6 * checked, res.stdout


def _expected_text(n, h, values, index, origin=(0.0, 0.0, 0.0)):
    dim = len(n)
    head = f"{dim} " + "".join(f"{c} " for c in n)
    head += "".join(f"{origin[a]:f} {origin[a] + (n[a] - 1) * h:f} " for a in range(dim))
    head += f"{index} @\n"
    flat = values.reshape(-1, n[0])
    body = "".join("".join("%.10E " % v for v in row) + "\n" for row in flat)
    return (head + body).encode()


@pytest.mark.parametrize("n", [(7, 5), (4, 3, 6)])
def test_oracle_text_matches_independent_formatter(orc, n):
    st = models.Stencil2d2h() if len(n) == 2 else Stencil3d2h()
    d = models.make_desc(n, st, 0.35, 0.01, [])
    psi = inputs.uniform(tuple(reversed(n)), seed=5, lo=-3.0, hi=3.0)
    psi.ravel()[:6] = [0.0, -0.0, 1e-300, 123456.789, -2.5e-7, 1.00000000005]
    g = orc.pad(d, psi)
    text = orc.checkpoint_text(d, g, index=42)
    assert text == _expected_text(n, 0.35, psi, 42)
    idx, back = orc.checkpoint_read(d, text)
    assert idx == 42
    want = np.array([float("%.10E" % v) for v in psi.ravel()]).reshape(psi.shape)
    assert np.array_equal(orc.unpad(d, back), want)


def test_oracle_header_origin(orc):
    d = models.make_desc((3, 4), models.Stencil2d2h(), 0.5, 0.01, [])
    g = orc.pad(d, np.zeros((4, 3)))
    text = orc.checkpoint_text(d, g, index=0, origin=[-1.0, 2.0])
    assert text.split(b"\n")[0] == b"2 3 4 -1.000000 0.000000 2.000000 3.500000 0 @"


def _expected_plot(n, h, values, fmt, origin=(0.0, 0.0, 0.0)):
    """GNU matrix (fmt 1) or CSV (fmt 2) plot text, formatted by Python's own
    correctly rounded % operator."""
    X, Y = n[0], n[1]
    Z = n[2] if len(n) == 3 else 1
    v = values.reshape(Z, Y, X)
    out = []
    for k in range(Z):
        if fmt == 1:
            out.append("% 13d " % k + "".join("% 13.2f " % (origin[0] + i * h)
                                               for i in range(X)) + "\n")
        for j in range(Y):
            if fmt == 1:
                out.append("% 13.2f " % (origin[1] + j * h)
                           + "".join("% 10E " % x for x in v[k, j]) + "\n")
            else:
                out.append(",".join("% 10E" % x for x in v[k, j]) + "\n")
        if fmt == 1 and k < Z - 1:
            out.append("\n")
    if fmt == 1:
        out.append("\n")
    return "".join(out).encode()


@pytest.mark.parametrize("fmt", [1, 2])
@pytest.mark.parametrize("n", [(7, 5), (4, 3, 6)])
def test_oracle_plot_text_matches_independent_formatter(orc, n, fmt):
    st = models.Stencil2d2h() if len(n) == 2 else Stencil3d2h()
    d = models.make_desc(n, st, 0.35, 0.01, [])
    psi = inputs.uniform(tuple(reversed(n)), seed=6, lo=-3.0, hi=3.0)
    psi.ravel()[:6] = [0.0, -0.0, 1e-300, 123456.789, -2.5e-7, 1.0000005]
    g = orc.pad(d, psi)
    assert orc.plot_text(d, g, fmt) == _expected_plot(n, 0.35, psi, fmt)
    o = (-1.25, 2.005, 0.0)
    assert orc.plot_text(d, g, fmt, origin=o) == _expected_plot(n, 0.35, psi, fmt, o)


@pytest.mark.parametrize("n", [(7, 5), (4, 3, 6)])
def test_oracle_xdr_matches_xdrlib(orc, n):
    """XDR bytes against Python's xdrlib (RFC 4506 encoder), and read back."""
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", DeprecationWarning)
        import xdrlib
    st = models.Stencil2d2h() if len(n) == 2 else Stencil3d2h()
    d = models.make_desc(n, st, 0.5, 0.01, [])
    psi = inputs.uniform(tuple(reversed(n)), seed=8, lo=-3.0, hi=3.0)
    psi.ravel()[:3] = [-0.0, 1e-310, float("inf")]
    g = orc.pad(d, psi)
    origin = (-1.0, 2.0, 0.5)[:len(n)]
    p = xdrlib.Packer()
    p.pack_int(len(n))
    for c in n:
        p.pack_int(c)
    for a in range(len(n)):
        p.pack_double(origin[a])
        p.pack_double(origin[a] + (n[a] - 1) * 0.5)
    p.pack_int(17)
    p.pack_int(len(n))
    for a in range(len(n)):
        p.pack_double(origin[a])
        p.pack_double(origin[a] + (n[a] - 1) * 0.5)
    for x in psi.ravel():
        p.pack_double(float(x))
    data = orc.xdr_write(d, g, index=17, origin=origin)
    assert data == p.get_buffer()
    idx, back = orc.xdr_read(d, data)
    assert idx == 17
    assert np.array_equal(orc.unpad(d, back), psi)
    with pytest.raises(RuntimeError):
        orc.xdr_read(d, data[:-1])
