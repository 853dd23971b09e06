"""ctypes view of include/sps_fd.h.

The structures mirror the C declarations field by field; `Desc` is also the
layout of the CPU oracle's problem descriptor, so one instance can describe a
problem to both.
"""
from __future__ import annotations

import ctypes
from ctypes import c_double, c_int, c_int32, c_size_t, c_void_p, POINTER
from pathlib import Path

MAX_DEGREE = 6
MAX_TERMS = 8
MAX_PROV = 2
MAX_FIELDS = 4
MAX_MTERMS = 12
BOUNDARY_DEPTH = 3

OK, E_INVALID, E_UNSUPPORTED, E_HIP, E_NOMEM, E_STATE = 0, -1, -2, -3, -4, -5
F32, F64 = 0, 1
OP_ID, OP_LAP, OP_BILAP, OP_GRADLAP = 0, 1, 2, 3
AXIS_X, AXIS_Y, AXIS_Z = 0, 1, 2
LAYOUT_INTERIOR, LAYOUT_PADDED = 0, 1
PATH_AUTO, PATH_REFERENCE, PATH_FUSED = 0, 1, 2
PLOT_GNU, PLOT_CSV = 1, 2
GRAD_FORWARD, GRAD_CENTRAL = 0, 1

STATUS_NAMES = {OK: "SPS_OK", E_INVALID: "SPS_E_INVALID", E_UNSUPPORTED: "SPS_E_UNSUPPORTED",
                E_HIP: "SPS_E_HIP", E_NOMEM: "SPS_E_NOMEM", E_STATE: "SPS_E_STATE"}


class Term(ctypes.Structure):
    # operand 0 is the field psi, k >= 1 the provisional variable var(k)
    _fields_ = [("op", c_int32), ("axis", c_int32), ("component", c_int32),
                ("operand", c_int32), ("coeff", c_double),
                ("poly", c_double * (MAX_DEGREE + 1))]


class Desc(ctypes.Structure):
    _fields_ = [("dim", c_int32), ("dtype", c_int32), ("n", c_int32 * 3),
                ("lap_points", c_int32), ("gradlap_points", c_int32),
                ("bilap_points", c_int32), ("nterms", c_int32), ("path", c_int32),
                ("h", c_double), ("dt", c_double), ("terms", Term * MAX_TERMS),
                ("nprov", c_int32), ("prov_nterms", c_int32 * MAX_PROV),
                ("prov_reserved", c_int32), ("prov", (Term * MAX_TERMS) * MAX_PROV)]


class MTerm(ctypes.Structure):
    # coeff * D[prod_f psi_f^pow[f]]
    _fields_ = [("op", c_int32), ("axis", c_int32), ("component", c_int32),
                ("reserved", c_int32), ("coeff", c_double), ("pow", c_int32 * MAX_FIELDS)]


class MDesc(ctypes.Structure):
    _fields_ = [("dim", c_int32), ("dtype", c_int32), ("n", c_int32 * 3),
                ("lap_points", c_int32), ("gradlap_points", c_int32),
                ("bilap_points", c_int32), ("nfields", c_int32), ("path", c_int32),
                ("h", c_double), ("dt", c_double), ("nterms", c_int32 * MAX_FIELDS),
                ("terms", (MTerm * MAX_MTERMS) * MAX_FIELDS)]


class Regional(ctypes.Structure):
    # sps_fd_regional; defaults are the reference's parameters
    _fields_ = [("cutoff_eps", c_double), ("resize_factor", c_double),
                ("resize_time", c_double), ("fixed_resize", c_int32),
                ("steps_per_launch", c_int32), ("dims_relative_eps", c_double)]


def regional_defaults() -> "Regional":
    return Regional(1e-6, 2.0, 3.0, 1, 0, 0.05)


class RegionInfo(ctypes.Structure):
    _fields_ = [("origin", c_int32 * 3), ("dims", c_int32 * 3), ("empty", c_double),
                ("cutoff", c_double), ("time", c_double), ("index", c_int32),
                ("reserved", c_int32)]


class Energy(ctypes.Structure):
    # -c1/2 psi^2 + c2/4 psi^4 + kappa/2 |grad psi|^2 (sps_fd_energy)
    _fields_ = [("c1", c_double), ("c2", c_double), ("kappa", c_double),
                ("gradient", c_int32), ("mean", c_int32)]


LIB_PATH = Path(__file__).resolve().parent / "lib" / "libsps_fd.so"

_lib = None


class SpsError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def _one_hip_runtime():
    """torch's ROCm wheel ships its own libamdhip64 (and HSA runtime) under the
    same SONAME as /opt/rocm's.  A library loaded first binds that SONAME for
    every later library that names it, but torch's libraries name their copy
    by file, so a torch imported after this library loads a second HIP
    runtime, and that one finds no device.  Importing torch first, when it is
    installed, leaves one HIP runtime in the process (torch's), which this
    library then shares; without torch it uses /opt/rocm's."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load() -> ctypes.CDLL:
    """Load the in-tree HIP library; fails loudly when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    _one_hip_runtime()
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " (there is no CPU fallback for the solver step)")
    lib = ctypes.CDLL(str(LIB_PATH))
    P = c_void_p
    sig = {
        "sps_fd_last_error": (ctypes.c_char_p, []),
        "sps_fd_abi_version": (c_int, []),
        "sps_fd_desc_size": (c_size_t, []),
        "sps_fd_device_count": (c_int, []),
        "sps_fd_create": (c_int, [POINTER(Desc), POINTER(P)]),
        "sps_fd_destroy": (c_int, [P]),
        "sps_fd_upload": (c_int, [P, P, c_int]),
        "sps_fd_download": (c_int, [P, P, c_int]),
        "sps_fd_step": (c_int, [P, c_int]),
        "sps_fd_equation": (c_int, [P, P, c_int]),
        "sps_fd_apply": (c_int, [P, c_int, c_int, c_int, P, c_int]),
        "sps_fd_reduce": (c_int, [P, POINTER(c_double), POINTER(c_double)]),
        "sps_fd_set_energy": (c_int, [P, P]),
        "sps_fd_reduce_device": (c_int, [c_int, c_int, POINTER(c_int32), c_double, P, P, P, P]),
        "sps_fd_batch_set_energy": (c_int, [P, P]),
        "sps_fd_sync": (c_int, [P]),
        "sps_fd_device_field": (c_int, [P, POINTER(P)]),
        "sps_fd_set_stream": (c_int, [P, P]),
        "sps_fd_copy_device": (c_int, [P, P, c_int]),
        "sps_fd_batch_copy_device": (c_int, [P, c_int, P, c_int]),
        "sps_fd_query": (c_int, [P, POINTER(c_int), POINTER(c_int)]),
        "sps_fd_checkpoint_text": (c_int, [P, c_int, P, P, c_size_t, POINTER(c_size_t)]),
        "sps_fd_checkpoint_read": (c_int, [P, ctypes.c_char_p, c_size_t, POINTER(c_int)]),
        "sps_fd_plot_text": (c_int, [P, c_int, P, c_void_p, c_size_t, POINTER(c_size_t)]),
        "sps_fd_checkpoint_xdr": (c_int, [P, c_int, P, c_void_p, c_size_t, POINTER(c_size_t)]),
        "sps_fd_checkpoint_read_xdr": (c_int, [P, c_void_p, c_size_t, POINTER(c_int)]),
        "sps_fdm_create": (c_int, [P, POINTER(c_void_p)]),
        "sps_fdm_destroy": (c_int, [P]),
        "sps_fdm_upload": (c_int, [P, c_int, P, c_int]),
        "sps_fdm_download": (c_int, [P, c_int, P, c_int]),
        "sps_fdm_step": (c_int, [P, c_int]),
        "sps_fdm_query": (c_int, [P, POINTER(c_int), POINTER(c_int)]),
        "sps_fdm_sync": (c_int, [P]),
        "sps_fdm_tune": (c_int, [P, c_int, c_int, c_int]),
        "sps_fdm_autotune": (c_int, [P, POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
        "sps_fdm_checkpoint_text": (c_int, [P, c_int, c_int, P, c_void_p, c_size_t, POINTER(c_size_t)]),
        "sps_fdm_checkpoint_read": (c_int, [P, c_int, ctypes.c_char_p, c_size_t, POINTER(c_int)]),
        "sps_fdm_plot_text": (c_int, [P, c_int, c_int, P, c_void_p, c_size_t, POINTER(c_size_t)]),
        "sps_fdm_checkpoint_xdr": (c_int, [P, c_int, c_int, P, c_void_p, c_size_t, POINTER(c_size_t)]),
        "sps_fdm_checkpoint_read_xdr": (c_int, [P, c_int, c_void_p, c_size_t, POINTER(c_int)]),
        "sps_fdm_set_stream": (c_int, [P, P]),
        "sps_fd_checkpoint_save": (c_int, [P, ctypes.c_char_p, c_int, c_int, c_int, c_int, P]),
        "sps_fd_checkpoint_load": (c_int, [P, ctypes.c_char_p, c_int, c_int, c_int,
                                           POINTER(c_int)]),
        "sps_fd_checkpoint_progress": (c_int, [ctypes.c_char_p, POINTER(c_int), POINTER(c_int),
                                               POINTER(c_int)]),
        "sps_fdm_checkpoint_save": (c_int, [P, ctypes.c_char_p, c_int, c_int, c_int, P]),
        "sps_fdm_checkpoint_load": (c_int, [P, ctypes.c_char_p, c_int, c_int, POINTER(c_int)]),
        "sps_fd_tune": (c_int, [P, c_int, c_int, c_int, c_int]),
        "sps_fd_autotune": (c_int, [P, c_int, POINTER(c_int), POINTER(c_int), POINTER(c_int),
                                    POINTER(c_int)]),
        "sps_fd_batch_autotune": (c_int, [P, POINTER(c_int), POINTER(c_int)]),
        "sps_fd_batch_create": (c_int, [POINTER(Desc), c_int, POINTER(P)]),
        "sps_fd_batch_destroy": (c_int, [P]),
        "sps_fd_batch_upload": (c_int, [P, c_int, P, c_int]),
        "sps_fd_batch_download": (c_int, [P, c_int, P, c_int]),
        "sps_fd_batch_step": (c_int, [P, c_int]),
        "sps_fd_batch_reduce": (c_int, [P, POINTER(c_double), POINTER(c_double)]),
        "sps_fd_batch_sync": (c_int, [P]),
        "sps_fd_batch_tune": (c_int, [P, c_int, c_int]),
        "sps_fd_batch_set_stream": (c_int, [P, P]),
        "sps_fdr_create": (c_int, [POINTER(Desc), P, POINTER(P)]),
        "sps_fdr_destroy": (c_int, [P]),
        "sps_fdr_upload": (c_int, [P, P, c_int]),
        "sps_fdr_step": (c_int, [P, c_int]),
        "sps_fdr_download": (c_int, [P, P, c_int]),
        "sps_fdr_region": (c_int, [P, POINTER(RegionInfo)]),
        "sps_fdr_sync": (c_int, [P]),
    }
    I, F = c_int, ctypes.c_float
    sig.update({
        "sps_nn_conv3x3": (c_int, [P, I, I, I, I, P, P, I, I, I, P, P, P, P]),
        "sps_nn_conv3x3_wgrad": (c_int, [P, P, I, I, I, I, I, P, P, c_size_t, P]),
        "sps_nn_wgrad_workspace": (c_size_t, [I, I, I, I, I, I]),
        "sps_nn_conv3x3_in1": (c_int, [P, I, I, I, P, P, I, I, P, P, P]),
        "sps_nn_conv3x3_in1_wgrad": (c_int, [P, P, I, I, I, I, P, P, c_size_t, P]),
        "sps_nn_conv3x3_out1": (c_int, [P, I, I, I, I, P, P, P, P]),
        "sps_nn_conv3x3_out1_bwd": (c_int, [P, P, I, I, I, I, P, P, P, P, c_size_t, P]),
        "sps_nn_silu_bwd": (c_int, [P, P, P, c_size_t, P]),
        "sps_nn_pool2": (c_int, [P, I, I, I, I, I, P, P]),
        "sps_nn_unpool2": (c_int, [P, I, I, I, I, F, P, P]),
        "sps_nn_center": (c_int, [P, P, I, ctypes.c_int64, P]),
    })
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != OK:
        msg = load().sps_fd_last_error()
        raise SpsError(status, msg.decode() if msg else "")
