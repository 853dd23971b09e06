"""The C ABI library loads and exports every symbol include/sps_fd.h declares,
its structures match the header, and descriptor validation fails cleanly.
No compute call is made: these run on machines without a GPU."""
import ctypes
import re
from pathlib import Path

import pytest

from symphas_amd import abi, models

ROOT = Path(__file__).resolve().parent.parent


def _declared_functions():
    text = (ROOT / "include" / "sps_fd.h").read_text()
    return sorted(set(re.findall(r"\b(sps_fd_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    names = _declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_descriptor_layout_matches_library_and_oracle(orc):
    lib = abi.load()
    assert lib.sps_fd_desc_size() == ctypes.sizeof(abi.Desc)
    assert orc.load().orc_desc_size() == ctypes.sizeof(abi.Desc)
    assert lib.sps_fd_abi_version() == 3


def _create(desc):
    lib = abi.load()
    h = ctypes.c_void_p()
    rc = lib.sps_fd_create(ctypes.byref(desc), ctypes.byref(h))
    return rc, lib.sps_fd_last_error().decode()


@pytest.mark.parametrize("mutate,needle", [
    (lambda d: setattr(d, "dim", 4), "dim"),
    (lambda d: setattr(d, "dtype", 7), "dtype"),
    (lambda d: d.n.__setitem__(0, 2), "n[0]"),
    (lambda d: setattr(d, "h", 0.0), "h must be"),
    (lambda d: setattr(d, "lap_points", 7), "Stencil2d2h"),
    (lambda d: setattr(d, "nterms", 9), "nterms"),
    (lambda d: setattr(d.terms[0], "op", 9), "unknown op"),
    (lambda d: setattr(d, "path", 5), "path"),
])
def test_invalid_descriptors_rejected_before_device(mutate, needle):
    d = models.model_b((64, 64))
    mutate(d)
    rc, msg = _create(d)
    assert rc == abi.E_INVALID
    assert needle in msg


def test_gradlap_component_validation():
    d = models.make_desc((16, 16), models.Stencil2d2h(5, 8, 13), 1.0, 0.01,
                         [models.TermSpec(abi.OP_GRADLAP, 1.0, models.PSI, axis=0, component=1)])
    rc, msg = _create(d)
    assert rc == abi.E_INVALID and "component" in msg


def test_no_device_reports_cleanly():
    lib = abi.load()
    if lib.sps_fd_device_count() > 0:
        pytest.skip("a HIP device is present")
    rc, msg = _create(models.model_b((64, 64)))
    assert rc == abi.E_HIP and "no HIP device" in msg


def test_null_arguments():
    lib = abi.load()
    assert lib.sps_fd_create(None, None) == abi.E_INVALID
    assert lib.sps_fd_step(None, 1) == abi.E_INVALID
    assert lib.sps_fd_destroy(None) == abi.OK


def test_models_build_reference_term_lists():
    d = models.model_b((8, 8))
    ops = [d.terms[i].op for i in range(d.nterms)]
    assert ops == [abi.OP_BILAP, abi.OP_LAP, abi.OP_LAP]
    assert [d.terms[i].coeff for i in range(3)] == [-1.0, -1.0, 1.0]
    assert list(d.terms[2].poly)[:4] == [0, 0, 0, 1]
    d = models.allen_cahn((8, 8))
    assert [d.terms[i].op for i in range(d.nterms)] == [abi.OP_LAP, abi.OP_ID]
    assert list(d.terms[1].poly)[:4] == [0, 1, 0, -1]
