"""The SymPhas plugin header (symphas_amd/host/symphas_plugin/solverft_mi355x.h).

The reference's headers need CMake-generated configuration headers, so the
plugin is compiled here against a test double of the SymPhas API it names
(tests/host/symphas_api_double): the expression types with the template
parameters and members the reference declares, the plugin macros and
SolverSystemFD.  tests/host/plugin_driver.cpp writes each model as the
canonical expression type form_expr_one receives.  Checked:
  - form_expr_one lowers Model B (two stencil choices), Allen-Cahn and a
    gradlaplacian-plus-source form to the expected term lists, and Model B's
    equals symphas_amd/models.py's (the reference's MODEL definition);
  - a right-hand side that reads another field is refused at run time, and
    an operator outside the vocabulary (a first derivative) at compile time;
  - on the GPU, Model B and Allen-Cahn stepped through the plugin in
    run_model's order (update, step) equal a plan made from the same
    descriptor bit for bit and the CPU oracle to 1e-12.
"""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import pyoracle as po
from symphas_amd import abi, inputs, models
from symphas_amd.solver import FDPlan

ROOT = Path(__file__).resolve().parent.parent
INC = ["-I", str(ROOT / "tests" / "host" / "symphas_api_double"), "-I", str(ROOT / "symphas_amd" / "host"),
       "-I", str(ROOT / "include")]


@pytest.fixture(scope="session")
def plugin(tmp_path_factory):
    out = tmp_path_factory.mktemp("plugin") / "plugin_driver"
    lib = ROOT / "symphas_amd" / "lib"
    cmd = [shutil.which("g++") or "g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", *INC,
           str(ROOT / "tests" / "host" / "plugin_driver.cpp"), f"-L{lib}", "-lsps_fd",
           f"-Wl,-rpath,{lib}", "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return out


def _lowered(plugin, model):
    res = subprocess.run([str(plugin), "lower", model], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    lines = res.stdout.strip().splitlines()
    head = lines[0].split()
    info = {"dim": int(head[1]), "n": (int(head[3]), int(head[4])), "L": int(head[6]),
            "G": int(head[8]), "B": int(head[10]), "h": float(head[12]), "dt": float(head[14])}
    terms = []
    for ln in lines[1:]:
        p = ln.split()
        terms.append((int(p[1]), int(p[2]), float(p[3]), [float(v) for v in p[4:]]))
    return info, terms


def _terms(desc):
    return [(desc.terms[i].op, desc.terms[i].axis, desc.terms[i].coeff, list(desc.terms[i].poly))
            for i in range(desc.nterms)]


def test_model_b_lowers_to_the_model_definition(plugin):
    info, terms = _lowered(plugin, "mb")
    want = models.model_b((16, 12), h=1.0, dt=0.01, c1=0.7, c2=1.3)
    assert info == {"dim": 2, "n": (16, 12), "L": 5, "G": 6, "B": 13, "h": 1.0, "dt": 0.01}
    assert terms == _terms(want)
    info9, terms9 = _lowered(plugin, "mb9")
    want9 = models.model_b((16, 12), h=1.0, dt=0.005, stencil=models.Stencil2d2h(9, 6, 17))
    assert (info9["L"], info9["G"], info9["B"], info9["dt"]) == (9, 6, 17, 0.005)
    assert terms9 == _terms(want9)


def test_allen_cahn_and_gradlap_forms_lower_term_by_term(plugin):
    info, terms = _lowered(plugin, "ac")
    assert info["h"] == 0.5
    z = [0.0] * 7
    lin = list(z); lin[1] = 1.0
    cub = list(z); cub[3] = -1.0
    assert terms == [(abi.OP_LAP, 0, 1.0, lin), (abi.OP_ID, 0, 1.0, lin), (abi.OP_ID, 0, 1.0, cub)]
    _, gl = _lowered(plugin, "gl")
    sq = list(z); sq[2] = 1.0
    const = list(z); const[0] = 0.25
    assert gl == [(abi.OP_GRADLAP, abi.AXIS_Y, 0.5, sq), (abi.OP_ID, 0, 1.0, const)]


def test_coupled_right_hand_side_is_refused(plugin):
    res = subprocess.run([str(plugin), "coupled"], capture_output=True, text=True)
    assert res.returncode == 5
    assert "reads field 1" in res.stdout and "sps_fdm" in res.stdout


def test_operator_outside_the_vocabulary_fails_to_compile(tmp_path):
    cmd = [shutil.which("g++") or "g++", "-std=c++17", "-fsyntax-only", *INC,
           str(ROOT / "tests" / "host" / "plugin_unsupported.cpp")]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode != 0
    assert "SolverFT_MI355X lowers laplacians (order 2)" in res.stderr


def _run(plugin, tmp_path, model, n0, n1, steps, psi):
    fin, fout = tmp_path / f"{model}_in.bin", tmp_path / f"{model}_out.bin"
    np.ascontiguousarray(psi, dtype=np.float64).tofile(fin)
    print(f"plugin run {model} {n0}x{n1} {steps} steps", flush=True)
    res = subprocess.run([str(plugin), "run", model, str(n0), str(n1), str(steps), str(fin),
                          str(fout)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    return np.fromfile(fout, dtype=np.float64).reshape(n1, n0)


@pytest.mark.gpu
def test_model_b_through_the_plugin(gpu_lib, plugin, tmp_path):
    n0, n1, steps = 64, 48, 300
    d = models.model_b((n0, n1), h=1.0, dt=0.01, dtype=abi.F64)
    psi = inputs.uniform((n1, n0), seed=41)
    got = _run(plugin, tmp_path, "mb", n0, n1, steps, psi)
    with FDPlan(d) as p:
        p.upload(psi)
        p.step(steps)
        plan = p.download()
    assert np.array_equal(got, plan)
    want = po.unpad(d, po.run(d, psi, steps))
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-12


@pytest.mark.gpu
def test_allen_cahn_through_the_plugin(gpu_lib, plugin, tmp_path):
    n0, n1, steps = 64, 64, 500
    # the plugin's lowering: lap(psi) + psi - psi^3 as three terms
    d = models.make_desc((n0, n1), models.Stencil2d2h(5, 6, 13), 0.5, 0.01,
                         [models.TermSpec(abi.OP_LAP, 1.0, models.PSI),
                          models.TermSpec(abi.OP_ID, 1.0, models.PSI),
                          models.TermSpec(abi.OP_ID, 1.0, (0.0, 0.0, 0.0, -1.0))], abi.F64)
    psi = inputs.uniform((n1, n0), seed=42)
    got = _run(plugin, tmp_path, "ac", n0, n1, steps, psi)
    with FDPlan(d) as p:
        p.upload(psi)
        p.step(steps)
        plan = p.download()
    assert np.array_equal(got, plan)
    want = po.unpad(d, po.run(d, psi, steps))
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-12
