"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, always as the checker.  It takes the same descriptor structure as the
product's C ABI (the two layouts are identical).
"""
from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, c_double, c_int, c_int64, c_size_t, c_uint64, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libsps_oracle.so"
DEPTH = 3

_lib = None


def build() -> Path:
    res = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stderr}")
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = ctypes.CDLL(str(LIB_PATH))
    P = c_void_p
    for name, res, args in [
        ("orc_desc_size", c_size_t, []),
        ("orc_check", c_int, [P]),
        ("orc_padded_len", c_size_t, [P]),
        ("orc_pad", None, [P, P, P]),
        ("orc_unpad", None, [P, P, P]),
        ("orc_refresh", None, [P, P]),
        ("orc_refresh_modular", None, [P, P]),
        ("orc_stencil", c_double, [P, c_int, c_int, c_int, P, c_int64]),
        ("orc_apply", None, [P, c_int, c_int, c_int, P, P, c_int]),
        ("orc_equation", c_int, [P, P, P, c_int]),
        ("orc_provisional", c_int, [P, P, c_int, P, c_int]),
        ("orc_euler", None, [P, P, P, c_int]),
        ("orc_run", c_int, [P, P, c_int, c_int]),
        ("orc_mass", c_double, [P, P]),
        ("orc_energy", c_double, [P, P]),
        ("orc_energy_ex", c_double, [P, P, c_double, c_double, c_double, c_int, c_int]),
        ("orc_fill_uniform", None, [P, c_size_t, c_uint64, c_double, c_double]),
        ("orc_checkpoint_text", c_int64, [P, P, c_int, P, P, c_size_t]),
        ("orcm_desc_size", c_size_t, []),
        ("orcm_check", c_int, [P]),
        ("orcm_equation", c_int, [P, P, P, c_int]),
        ("orcm_run", c_int, [P, P, c_int, c_int]),
        ("orc_checkpoint_read", c_int, [P, ctypes.c_char_p, P, POINTER(c_int)]),
        ("orc_plot_text", c_int64, [P, P, c_int, P, P, c_size_t]),
        ("orc_xdr_write", c_int64, [P, P, c_int, P, P, c_size_t]),
        ("orc_xdr_read", c_int, [P, ctypes.c_char_p, c_size_t, P, POINTER(c_int)]),
        ("orc_cover_interval", c_int, [P, c_int, POINTER(c_int), POINTER(c_int)]),
        ("orc_regional_run", c_int, [P, P, P, c_int, P, c_int]),
    ]:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(c_void_p)


def _d(desc):
    return ctypes.cast(ctypes.byref(desc), c_void_p)


def padded_shape(desc):
    return tuple(int(desc.n[a]) + 2 * DEPTH for a in reversed(range(desc.dim)))


def interior_shape(desc):
    return tuple(int(desc.n[a]) for a in reversed(range(desc.dim)))


def pad(desc, interior):
    out = np.zeros(padded_shape(desc))
    load().orc_pad(_d(desc), _p(np.ascontiguousarray(interior, dtype=np.float64)), _p(out))
    return out


def unpad(desc, padded):
    out = np.empty(interior_shape(desc))
    load().orc_unpad(_d(desc), _p(np.ascontiguousarray(padded, dtype=np.float64)), _p(out))
    return out


def refresh(desc, padded, modular=False):
    fn = load().orc_refresh_modular if modular else load().orc_refresh
    fn(_d(desc), _p(padded))
    return padded


def apply(desc, op, axis, component, padded, nthreads=1):
    out = np.zeros_like(padded)
    load().orc_apply(_d(desc), op, axis, component, _p(padded), _p(out), nthreads)
    return out


def stencil_at(desc, op, axis, component, padded, idx):
    return load().orc_stencil(_d(desc), op, axis, component, _p(padded), int(idx))


def equation(desc, padded, nthreads=1):
    out = np.zeros_like(padded)
    rc = load().orc_equation(_d(desc), _p(padded), _p(out), nthreads)
    if rc:
        raise RuntimeError(f"orc_equation failed ({rc})")
    return out


def provisional(desc, padded, k, nthreads=1):
    """var(k) over the padded layout, ghosts refreshed (padded psi must have
    fresh ghosts)."""
    out = np.zeros_like(padded)
    rc = load().orc_provisional(_d(desc), _p(padded), int(k), _p(out), nthreads)
    if rc:
        raise RuntimeError(f"orc_provisional failed ({rc})")
    return out


def run(desc, interior, nsteps, nthreads=1):
    """Return the padded field after nsteps (ghosts refreshed)."""
    g = pad(desc, interior)
    rc = load().orc_run(_d(desc), _p(g), int(nsteps), nthreads)
    if rc:
        raise RuntimeError(f"orc_run failed ({rc})")
    return g


def mass(desc, padded):
    return load().orc_mass(_d(desc), _p(padded))


def energy(desc, padded, c1=1.0, c2=1.0, kappa=1.0, gradient=0, mean=False):
    """Landau free energy of a padded field with fresh ghosts (gradient 0:
    forward differences, 1: central; mean: divide by the site count instead
    of multiplying by h^D)."""
    return load().orc_energy_ex(_d(desc), _p(padded), float(c1), float(c2),
                                float(kappa), int(gradient), int(bool(mean)))


def fill_uniform(count, seed, lo, hi):
    out = np.empty(count)
    load().orc_fill_uniform(_p(out), count, seed, lo, hi)
    return out


def checkpoint_text(desc, padded, index=0, origin=None):
    """The reference's GNU checkpoint text of the interior (bytes)."""
    o = None if origin is None else _p(np.asarray(origin, dtype=np.float64))
    n = load().orc_checkpoint_text(_d(desc), _p(padded), int(index), o, None, 0)
    if n < 0:
        raise RuntimeError("orc_checkpoint_text failed")
    buf = ctypes.create_string_buffer(int(n) + 1)
    load().orc_checkpoint_text(_d(desc), _p(padded), int(index), o,
                               ctypes.cast(buf, c_void_p), int(n) + 1)
    return buf.raw[:int(n)]


def checkpoint_read(desc, text):
    """(index, padded field) read the way the reference reads a checkpoint."""
    g = np.zeros(padded_shape(desc))
    idx = c_int(0)
    rc = load().orc_checkpoint_read(_d(desc), text, _p(g), ctypes.byref(idx))
    if rc:
        raise RuntimeError(f"orc_checkpoint_read failed ({rc})")
    return idx.value, g


def plot_text(desc, padded, fmt, origin=None):
    """The reference's plot data text (fmt 1: GNU matrix, 2: CSV) of the
    interior (bytes)."""
    o = None if origin is None else _p(np.asarray(origin, dtype=np.float64))
    n = load().orc_plot_text(_d(desc), _p(padded), int(fmt), o, None, 0)
    if n < 0:
        raise RuntimeError("orc_plot_text failed")
    buf = ctypes.create_string_buffer(int(n) + 1)
    load().orc_plot_text(_d(desc), _p(padded), int(fmt), o, ctypes.cast(buf, c_void_p),
                         int(n) + 1)
    return buf.raw[:int(n)]


def xdr_write(desc, padded, index=0, origin=None):
    """The reference's XDR file of the interior (bytes)."""
    o = None if origin is None else _p(np.asarray(origin, dtype=np.float64))
    n = load().orc_xdr_write(_d(desc), _p(padded), int(index), o, None, 0)
    if n < 0:
        raise RuntimeError("orc_xdr_write failed")
    buf = ctypes.create_string_buffer(int(n))
    load().orc_xdr_write(_d(desc), _p(padded), int(index), o, ctypes.cast(buf, c_void_p),
                         int(n))
    return buf.raw[:int(n)]


def xdr_read(desc, data):
    """(index, padded field) read the way the reference reads an XDR file."""
    g = np.zeros(padded_shape(desc))
    idx = c_int(0)
    rc = load().orc_xdr_read(_d(desc), data, len(data), _p(g), ctypes.byref(idx))
    if rc:
        raise RuntimeError(f"orc_xdr_read failed ({rc})")
    return idx.value, g


def mrun(mdesc, fields, nsteps, nthreads=1):
    """Coupled fields: `fields` is a sequence of interior arrays; returns the
    list of interior arrays after nsteps."""
    g = [pad(mdesc, f) for f in fields]
    buf = np.ascontiguousarray(np.stack(g))
    rc = load().orcm_run(_d(mdesc), _p(buf), int(nsteps), nthreads)
    if rc:
        raise RuntimeError(f"orcm_run failed ({rc})")
    return [unpad(mdesc, buf[f]) for f in range(len(fields))]


def mequation(mdesc, fields, nthreads=1):
    """Right-hand sides of all fields (interior arrays), ghosts refreshed first."""
    g = np.ascontiguousarray(np.stack([refresh(mdesc, pad(mdesc, f)) for f in fields]))
    out = np.zeros_like(g)
    rc = load().orcm_equation(_d(mdesc), _p(g), _p(out), nthreads)
    if rc:
        raise RuntimeError(f"orcm_equation failed ({rc})")
    return [unpad(mdesc, out[f]) for f in range(len(fields))]


def mcheck(mdesc):
    return load().orcm_check(_d(mdesc))


def check(desc):
    return load().orc_check(_d(desc))


class Regional(ctypes.Structure):
    """orc_regional (same layout as the product's sps_fd_regional)."""
    _fields_ = [("cutoff_eps", c_double), ("resize_factor", c_double),
                ("resize_time", c_double), ("fixed_resize", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("dims_relative_eps", c_double)]


class Region(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_int32 * 3), ("dims", ctypes.c_int32 * 3),
                ("empty", c_double), ("cutoff", c_double), ("time", c_double)]


def regional_defaults():
    """The reference's defaults (sol/src/boundarysystem.cpp:24-27,
    sol/inc/boundarysystem.h:137)."""
    return Regional(1e-6, 2.0, 3.0, 1, 0, 0.05)


def cover_interval(occ):
    """(start, length) of the shortest periodic interval covering the set
    coordinates of occ."""
    a = np.ascontiguousarray(np.asarray(occ, dtype=np.uint8))
    s, n = c_int(0), c_int(0)
    load().orc_cover_interval(_p(a), int(a.size), ctypes.byref(s), ctypes.byref(n))
    return s.value, n.value


def regional_run(desc, params, interior, nsteps, nthreads=1):
    """(whole-domain interior field, Region) after nsteps on a regional grid."""
    f = np.ascontiguousarray(interior, dtype=np.float64).copy()
    st = Region()
    rc = load().orc_regional_run(_d(desc), ctypes.cast(ctypes.byref(params), c_void_p), _p(f),
                                 int(nsteps), ctypes.cast(ctypes.byref(st), c_void_p), nthreads)
    if rc:
        raise RuntimeError(f"orc_regional_run failed ({rc})")
    return f, st
