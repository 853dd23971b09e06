"""The C++ drop-in adapter (symphas_amd/host/solverft_mi355x.hpp): its
EVOLUTION lowering produces the same term lists as symphas_amd/models.py
(the reference's MODEL definitions), it rejects what the step cannot run, and
models driven through it on the GPU match the CPU oracle."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from symphas_amd import abi, inputs, models

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="session")
def driver(tmp_path_factory):
    out = tmp_path_factory.mktemp("adapter") / "adapter_driver"
    lib = ROOT / "symphas_amd" / "lib"
    cmd = [shutil.which("g++") or "g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror",
           str(ROOT / "tests" / "host" / "adapter_driver.cpp"), f"-L{lib}", "-lsps_fd",
           f"-Wl,-rpath,{lib}", "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return out


def _parse(text):
    out = {}
    for line in text.strip().splitlines():
        parts = [p.strip() for p in line.split("|")]
        name, rest = parts[0], parts[1:]
        if rest and rest[0].startswith("error"):
            out[name] = rest[0]
            continue
        terms = []
        for t in rest:
            kv = dict(x.split("=", 1) for x in t.split())
            if "pow" in kv:  # a coupled-field term
                terms.append((int(kv["op"]), int(kv["axis"]), float(kv["coeff"]),
                              [int(v) for v in kv["pow"].split(",")]))
                continue
            terms.append((int(kv["op"]), int(kv["axis"]), float(kv["coeff"]),
                          [float(v) for v in kv["poly"].split(",")], int(kv["operand"])))
        out[name] = terms
    return out


def _expected(desc, group=0):
    terms = desc.terms if group == 0 else desc.prov[group - 1]
    n = desc.nterms if group == 0 else desc.prov_nterms[group - 1]
    return [(terms[i].op, terms[i].axis, terms[i].coeff, list(terms[i].poly), terms[i].operand)
            for i in range(n)]


def test_lowering_matches_model_definitions(driver):
    res = subprocess.run([str(driver), "lower"], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    got = _parse(res.stdout)
    assert got["allen_cahn"] == _expected(models.allen_cahn((8, 8)))
    assert got["model_b"] == _expected(models.model_b((8, 8)))
    assert got["model_a"] == _expected(models.make_desc((8, 8), models.Stencil2d2h(), 1, 0.01,
                                                        models.model_a_terms()))
    assert got["model_b_coeffs"] == _expected(models.model_b((8, 8), c1=0.7, c2=1.3))
    assert got["lap_lap"] == [(abi.OP_BILAP, 0, 1.0, [0, 1, 0, 0, 0, 0, 0], 0),
                              (abi.OP_LAP, 0, 0.5, [0, 0, 1, 0, 0, 0, 0], 0)]
    assert got["gradlap_y"] == [(abi.OP_GRADLAP, 1, 2.0, [0, 0, 1, 0, 0, 0, 0], 0)]
    for name in ("derivative_product", "derivative_times_field", "power7", "second_field",
                 "field_times_var", "var3"):
        assert got[name].startswith("error -2"), name
    # provisional variables (PROVISIONAL_DEF) lower to the models' term lists
    mbb = models.model_mbb((8, 8), c1=1.1, c2=0.4)
    bb = models.model_bb((8, 8), a=1.1, b=0.4)
    assert got["mbb_evolution"] == _expected(mbb)
    assert got["mbb_var1"] == _expected(mbb, 1)
    assert got["bb_evolution"] == _expected(bb)
    assert got["bb_var1"] == _expected(bb, 1)
    # 0.5 var(2) + psi - 2 + lap(var(1)^2): one term per operand, psi first
    assert got["mixed_operands"] == [(abi.OP_LAP, 0, 1.0, [0, 0, 1, 0, 0, 0, 0], 1),
                                     (abi.OP_ID, 0, 1.0, [-2, 1, 0, 0, 0, 0, 0], 0),
                                     (abi.OP_ID, 0, 1.0, [0, 0.5, 0, 0, 0, 0, 0], 2)]
    # coupled fields: the right-hand sides of MC, GRAYSCOTT and Turing as
    # written in the reference lower to models.py's term lists
    mc = (1.1, 0.8, 0.6, 1.3, 0.4)
    for name, terms in (("model_c", models.model_c_terms(mc)),
                        ("gray_scott", models.gray_scott_terms(mc[:4])),
                        ("turing", models.turing_terms(mc + (1.0,)))):
        for f, field in enumerate(("psi", "rho")):
            want = [(t.op, t.axis, t.coeff, list(t.pow) + [0] * (abi.MAX_FIELDS - len(t.pow)))
                    for t in terms[f]]
            assert got[f"{name}_{field}"] == want, (name, field)
    for name in ("second_field_single", "field5"):
        assert got[name].startswith("error -2"), name


def _run(driver, case, tmp_path):
    out = tmp_path / f"{case}.bin"
    print(f"[case] adapter {case}", flush=True)
    res = subprocess.run([str(driver), "run", case, str(out)], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    return np.fromfile(out, dtype=np.float64)


@pytest.mark.gpu
def test_adapter_allen_cahn_bitwise(gpu_lib, orc, driver, tmp_path):
    got = _run(driver, "allen_cahn", tmp_path).reshape(64, 64)
    d = models.allen_cahn((64, 64), h=0.5, dt=0.01, dtype=abi.F64, path=abi.PATH_REFERENCE)
    want = orc.unpad(d, orc.run(d, inputs.uniform((64, 64), seed=1), 1000, 8))
    assert np.array_equal(got, want)


@pytest.mark.gpu
def test_adapter_model_b_single_steps_bitwise(gpu_lib, orc, driver, tmp_path):
    got = _run(driver, "model_b_steps", tmp_path).reshape(40, 48)
    d = models.model_b((48, 40), h=1.0, dt=0.01, stencil=models.Stencil2d2h(9, 6, 13),
                       dtype=abi.F64, path=abi.PATH_REFERENCE)
    want = orc.unpad(d, orc.run(d, inputs.uniform((40, 48), seed=2), 50, 8))
    assert np.array_equal(got, want)


@pytest.mark.gpu
def test_adapter_model_b_3d_fp32(gpu_lib, orc, driver, tmp_path):
    got = _run(driver, "model_b_3d_fp32", tmp_path).reshape(16, 24, 32)
    d = models.model_b((32, 24, 16), h=1.0, dt=0.005, stencil=models.Stencil3d2h(7, 10, 25),
                       dtype=abi.F64, path=abi.PATH_REFERENCE)
    psi = inputs.uniform((16, 24, 32), seed=3, lo=-0.05, hi=0.05).astype(np.float32)
    want = orc.unpad(d, orc.run(d, psi.astype(np.float64), 40, 8))
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-5


@pytest.mark.gpu
def test_adapter_mbb_provisional_bitwise(gpu_lib, orc, driver, tmp_path):
    got = _run(driver, "mbb", tmp_path).reshape(28, 40)
    d = models.model_mbb((40, 28), h=0.9, dt=0.002, c1=1.1, c2=0.4)
    want = orc.unpad(d, orc.run(d, inputs.uniform((28, 40), seed=4, lo=-0.5, hi=0.5), 30, 8))
    assert np.array_equal(got, want)


@pytest.mark.gpu
def test_adapter_bb_provisional_fused(gpu_lib, orc, driver, tmp_path):
    got = _run(driver, "bb", tmp_path).reshape(28, 40)
    d = models.model_bb((40, 28), h=0.9, dt=0.002, a=1.1, b=0.4)
    want = orc.unpad(d, orc.run(d, inputs.uniform((28, 40), seed=4, lo=-0.5, hi=0.5), 30, 8))
    assert np.linalg.norm(got - want) / np.linalg.norm(want) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["model_c", "gray_scott", "turing"])
def test_adapter_coupled_bitwise(gpu_lib, orc, driver, tmp_path, case):
    """MultiModel on the reference-order path against the oracle, bit for
    bit, from the right-hand sides as written in the reference."""
    got = _run(driver, case, tmp_path).reshape(2, 28, 40)
    cf, dt = {"model_c": ((1.1, 0.8, 0.6, 1.3, 0.4), 0.002),
              "gray_scott": ((0.16, 0.08, 0.035, 0.065), 0.5),
              "turing": ((1.0, 0.7, 1.2, 0.5, -0.4, 0.9), 0.01)}[case]
    d = getattr(models, case)((40, 28), h=1.0, dt=dt, c=cf, dtype=abi.F64,
                              path=abi.PATH_REFERENCE)
    init = [inputs.uniform((28, 40), seed=5, lo=-0.5, hi=0.5),
            inputs.uniform((28, 40), seed=6, lo=-0.5, hi=0.5)]
    want = orc.mrun(d, init, 30, 8)
    for f in range(2):
        assert np.array_equal(got[f], want[f]), f


@pytest.mark.gpu
def test_adapter_gray_scott_fused(gpu_lib, orc, driver, tmp_path):
    got = _run(driver, "gray_scott_fused", tmp_path).reshape(2, 28, 40)
    d = models.gray_scott((40, 28), h=1.0, dt=0.5, c=(0.16, 0.08, 0.035, 0.065),
                          dtype=abi.F64, path=abi.PATH_FUSED)
    init = [inputs.uniform((28, 40), seed=5, lo=-0.5, hi=0.5),
            inputs.uniform((28, 40), seed=6, lo=-0.5, hi=0.5)]
    want = orc.mrun(d, init, 30, 8)
    for f in range(2):
        assert np.linalg.norm(got[f] - want[f]) / np.linalg.norm(want[f]) <= 1e-12
