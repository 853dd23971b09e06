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
