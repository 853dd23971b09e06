"""Host-side handle of the HIP finite-difference step (SolverFT semantics).

`FDPlan` owns one periodic field on the device and advances it with the
hand-written gfx950 kernels through the C ABI in include/sps_fd.h.  Arrays
cross the boundary as numpy arrays of the plan's dtype, shaped (ny, nx) or
(nz, ny, nx) for the interior layout and with 2*3 more cells per axis for the
SymPhas padded layout (datatypes/inc/gridinfo.h:46).
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_double, c_int, c_size_t, c_void_p

import numpy as np

from . import abi


def _np_dtype(d: abi.Desc):
    return np.float32 if d.dtype == abi.F32 else np.float64


def interior_shape(d: abi.Desc):
    return tuple(int(d.n[a]) for a in reversed(range(d.dim)))


def padded_shape(d: abi.Desc):
    return tuple(int(d.n[a]) + 2 * abi.BOUNDARY_DEPTH for a in reversed(range(d.dim)))


class FDPlan:
    def __init__(self, desc: abi.Desc):
        self.lib = abi.load()
        self.desc = desc
        self.dtype = _np_dtype(desc)
        h = c_void_p()
        abi.check(self.lib.sps_fd_create(byref(desc), byref(h)))
        self._h = h

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self.lib.sps_fd_destroy(self._h)
            self._h = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- data ---------------------------------------------------------------
    def _shape(self, layout):
        return interior_shape(self.desc) if layout == abi.LAYOUT_INTERIOR else padded_shape(self.desc)

    def upload(self, field: np.ndarray, layout: int = abi.LAYOUT_INTERIOR):
        arr = np.ascontiguousarray(field, dtype=self.dtype)
        if arr.shape != self._shape(layout):
            raise ValueError(f"field shape {arr.shape} != {self._shape(layout)}")
        abi.check(self.lib.sps_fd_upload(self._h, arr.ctypes.data_as(c_void_p), layout))

    def download(self, layout: int = abi.LAYOUT_INTERIOR) -> np.ndarray:
        out = np.empty(self._shape(layout), dtype=self.dtype)
        abi.check(self.lib.sps_fd_download(self._h, out.ctypes.data_as(c_void_p), layout))
        return out

    def step(self, nsteps: int = 1):
        abi.check(self.lib.sps_fd_step(self._h, int(nsteps)))

    def equation(self, layout: int = abi.LAYOUT_INTERIOR) -> np.ndarray:
        out = np.empty(self._shape(layout), dtype=self.dtype)
        abi.check(self.lib.sps_fd_equation(self._h, out.ctypes.data_as(c_void_p), layout))
        return out

    def apply(self, op: int, axis: int = 0, component: int = 0,
              layout: int = abi.LAYOUT_INTERIOR) -> np.ndarray:
        out = np.empty(self._shape(layout), dtype=self.dtype)
        abi.check(self.lib.sps_fd_apply(self._h, op, axis, component,
                                        out.ctypes.data_as(c_void_p), layout))
        return out

    def reduce(self):
        m, e = c_double(), c_double()
        abi.check(self.lib.sps_fd_reduce(self._h, byref(m), byref(e)))
        return m.value, e.value

    def set_energy(self, energy: "abi.Energy | None" = None):
        """Functional of reduce() (None: c1 = c2 = kappa = 1, forward
        differences, integral)."""
        arg = None if energy is None else ctypes.cast(byref(energy), c_void_p)
        abi.check(self.lib.sps_fd_set_energy(self._h, arg))

    def sync(self):
        abi.check(self.lib.sps_fd_sync(self._h))

    def set_stream(self, stream_handle: int | None):
        """Enqueue later work on a caller-owned hipStream_t (0/None: the null
        stream, e.g. torch's default stream)."""
        abi.check(self.lib.sps_fd_set_stream(self._h, c_void_p(stream_handle or 0)))

    def copy_to_device(self, dev_ptr: int):
        abi.check(self.lib.sps_fd_copy_device(self._h, c_void_p(dev_ptr), 0))

    def copy_from_device(self, dev_ptr: int):
        abi.check(self.lib.sps_fd_copy_device(self._h, c_void_p(dev_ptr), 1))

    def device_ptr(self) -> int:
        p = c_void_p()
        abi.check(self.lib.sps_fd_device_field(self._h, byref(p)))
        return p.value

    def query(self):
        path, spl = c_int(), c_int()
        abi.check(self.lib.sps_fd_query(self._h, byref(path), byref(spl)))
        return path.value, spl.value

    def checkpoint_text(self, index: int = 0, origin=None) -> bytes:
        """The field as the reference's GNU checkpoint file (formatted on the
        device; byte-identical to fprintf("%.10E "))."""
        o = None
        if origin is not None:
            o_arr = np.ascontiguousarray(origin, dtype=np.float64)
            o = o_arr.ctypes.data_as(c_void_p)
        # upper bound: 19 bytes per value ("-d.ddddddddddE+ddd "), one newline
        # per row, the header line
        sites = int(np.prod(interior_shape(self.desc)))
        cap = 19 * sites + sites // int(self.desc.n[0]) + 4096
        buf = np.empty(cap, dtype=np.uint8)
        n = c_size_t(0)
        abi.check(self.lib.sps_fd_checkpoint_text(self._h, int(index), o,
                                                  buf.ctypes.data_as(c_void_p), cap, byref(n)))
        return buf[:n.value].tobytes()

    def read_checkpoint(self, text: bytes) -> int:
        """Load a GNU checkpoint file into the field; returns its index."""
        idx = c_int(0)
        abi.check(self.lib.sps_fd_checkpoint_read(self._h, text, len(text), byref(idx)))
        return idx.value

    def plot_text(self, fmt: int = abi.PLOT_GNU, origin=None) -> bytes:
        """The field as the reference's plot data file: abi.PLOT_GNU (GNU
        matrix with positions) or abi.PLOT_CSV, formatted on the device."""
        o = None
        if origin is not None:
            o_arr = np.ascontiguousarray(origin, dtype=np.float64)
            o = o_arr.ctypes.data_as(c_void_p)
        # upper bound: 15 bytes per value ("-d.ddddddE+ddd "), 15 per row
        # (position column and newline), a header line of 14 bytes per x
        # position per z slice, blank lines
        n_ = [int(self.desc.n[a]) for a in range(self.desc.dim)]
        sites = int(np.prod(n_))
        rows = sites // n_[0]
        slices = n_[2] if self.desc.dim == 3 else 1
        cap = 15 * sites + 15 * rows + slices * (14 * (n_[0] + 1) + 2) + 16
        buf = np.empty(cap, dtype=np.uint8)
        n = c_size_t(0)
        abi.check(self.lib.sps_fd_plot_text(self._h, int(fmt), o, buf.ctypes.data_as(c_void_p),
                                            cap, byref(n)))
        return buf[:n.value].tobytes()

    def checkpoint_xdr(self, index: int = 0, origin=None) -> bytes:
        """The field as the reference's XDR file (big-endian doubles)."""
        o = None
        if origin is not None:
            o_arr = np.ascontiguousarray(origin, dtype=np.float64)
            o = o_arr.ctypes.data_as(c_void_p)
        n = c_size_t(0)
        abi.check(self.lib.sps_fd_checkpoint_xdr(self._h, int(index), o, None, 0, byref(n)))
        buf = np.empty(n.value, dtype=np.uint8)
        abi.check(self.lib.sps_fd_checkpoint_xdr(self._h, int(index), o,
                                                 buf.ctypes.data_as(c_void_p), buf.size,
                                                 byref(n)))
        return buf.tobytes()

    def read_checkpoint_xdr(self, data: bytes) -> int:
        """Load an XDR file into the field; returns its index."""
        idx = c_int(0)
        arr = np.frombuffer(data, dtype=np.uint8)
        abi.check(self.lib.sps_fd_checkpoint_read_xdr(self._h, arr.ctypes.data_as(c_void_p),
                                                      arr.size, byref(idx)))
        return idx.value

    def checkpoint_save(self, directory: str, index: int, field_id: int = 0,
                        single_file: bool = True, done: bool = False, origin=None):
        """Checkpoint the field as field `field_id` of the run in `directory`
        (<directory>/checkpoint/data<id>[_<index>] plus the progress record)."""
        keep, optr = _origin_arg(origin)
        abi.check(self.lib.sps_fd_checkpoint_save(self._h, str(directory).encode(), int(field_id),
                                                  int(index), int(bool(single_file)),
                                                  int(bool(done)), optr))

    def checkpoint_load(self, directory: str, index: int, field_id: int = 0,
                        single_file: bool = True) -> int:
        """Resume from the checkpoint of `index` (single file: the last block
        whose index does not pass it); returns the index loaded."""
        got = c_int(-1)
        abi.check(self.lib.sps_fd_checkpoint_load(self._h, str(directory).encode(), int(field_id),
                                                  int(index), int(bool(single_file)), byref(got)))
        return got.value

    def tune(self, steps_per_launch: int = 0, rows_per_tile: int = 0, planes_per_chunk: int = 0,
             values_per_lane: int = 0):
        abi.check(self.lib.sps_fd_tune(self._h, steps_per_launch, rows_per_tile, planes_per_chunk,
                                       values_per_lane))

    def autotune(self, use_cache: bool = True):
        """Time candidate launch shapes on the device and keep the fastest;
        returns (steps_per_launch, rows_per_tile, planes_per_chunk,
        values_per_lane)."""
        k, r, z, v = c_int(), c_int(), c_int(), c_int()
        abi.check(self.lib.sps_fd_autotune(self._h, int(use_cache), byref(k), byref(r), byref(z),
                                           byref(v)))
        return k.value, r.value, z.value, v.value


def _origin_arg(origin):
    if origin is None:
        return None, None
    arr = np.ascontiguousarray(origin, dtype=np.float64)
    return arr, arr.ctypes.data_as(c_void_p)


def _text_cap(desc, per_value: int, per_row: int, per_slice_header: int = 0) -> int:
    n = [int(desc.n[a]) for a in range(desc.dim)]
    sites = int(np.prod(n))
    rows = sites // n[0]
    slices = n[2] if desc.dim == 3 else 1
    return per_value * sites + per_row * rows + slices * per_slice_header + 4096


def checkpoint_progress(directory: str):
    """(last index or -1, number of indices recorded, done) of the run's
    progress record <directory>/checkpoint/progress.txt."""
    lib = abi.load()
    last, count, done = c_int(), c_int(), c_int()
    abi.check(lib.sps_fd_checkpoint_progress(str(directory).encode(), byref(last), byref(count),
                                             byref(done)))
    return last.value, count.value, bool(done.value)


def device_count() -> int:
    return int(abi.load().sps_fd_device_count())


class FDBatch:
    """Many independent 2D FP32 fields advanced together by one persistent
    launch per step group (sps_fd_batch_* in include/sps_fd.h)."""

    def __init__(self, descs):
        self.lib = abi.load()
        self.descs = list(descs)
        arr = (abi.Desc * len(self.descs))(*self.descs)
        h = c_void_p()
        abi.check(self.lib.sps_fd_batch_create(arr, len(self.descs), byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self.lib.sps_fd_batch_destroy(self._h)
            self._h = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, index: int, field: np.ndarray, layout: int = abi.LAYOUT_INTERIOR):
        d = self.descs[index]
        shape = interior_shape(d) if layout == abi.LAYOUT_INTERIOR else padded_shape(d)
        arr = np.ascontiguousarray(field, dtype=np.float32)
        if arr.shape != shape:
            raise ValueError(f"field shape {arr.shape} != {shape}")
        abi.check(self.lib.sps_fd_batch_upload(self._h, index, arr.ctypes.data_as(c_void_p), layout))

    def download(self, index: int, layout: int = abi.LAYOUT_INTERIOR) -> np.ndarray:
        d = self.descs[index]
        shape = interior_shape(d) if layout == abi.LAYOUT_INTERIOR else padded_shape(d)
        out = np.empty(shape, dtype=np.float32)
        abi.check(self.lib.sps_fd_batch_download(self._h, index, out.ctypes.data_as(c_void_p), layout))
        return out

    def step(self, nsteps: int = 1):
        abi.check(self.lib.sps_fd_batch_step(self._h, int(nsteps)))

    def reduce(self):
        n = len(self.descs)
        m = (c_double * n)()
        e = (c_double * n)()
        abi.check(self.lib.sps_fd_batch_reduce(self._h, m, e))
        return np.array(m[:]), np.array(e[:])

    def set_energy(self, energies=None):
        """Per-instance functionals of reduce(): a sequence of abi.Energy, one
        per instance (None: the default for every instance)."""
        if energies is None:
            abi.check(self.lib.sps_fd_batch_set_energy(self._h, None))
            return
        arr = (abi.Energy * len(self.descs))(*energies)
        abi.check(self.lib.sps_fd_batch_set_energy(self._h, ctypes.cast(arr, c_void_p)))

    def sync(self):
        abi.check(self.lib.sps_fd_batch_sync(self._h))

    def tune(self, steps_per_launch: int = 0, rows_per_item: int = 0):
        abi.check(self.lib.sps_fd_batch_tune(self._h, steps_per_launch, rows_per_item))

    def autotune(self):
        k, r = c_int(), c_int()
        abi.check(self.lib.sps_fd_batch_autotune(self._h, byref(k), byref(r)))
        return k.value, r.value

    def set_stream(self, stream_handle):
        abi.check(self.lib.sps_fd_batch_set_stream(self._h, c_void_p(stream_handle or 0)))

    def copy_to_device(self, index: int, dev_ptr: int):
        abi.check(self.lib.sps_fd_batch_copy_device(self._h, index, c_void_p(dev_ptr), 0))

    def copy_from_device(self, index: int, dev_ptr: int):
        abi.check(self.lib.sps_fd_batch_copy_device(self._h, index, c_void_p(dev_ptr), 1))


class MultiPlan:
    """A plan of coupled fields (sps_fdm_*): field f is uploaded and
    downloaded by index; step() advances all of them."""

    def __init__(self, desc: abi.MDesc):
        self.lib = abi.load()
        self.desc = desc
        h = c_void_p()
        abi.check(self.lib.sps_fdm_create(byref(desc), byref(h)))
        self._h = h
        self.dtype = np.float32 if desc.dtype == abi.F32 else np.float64

    def close(self):
        if getattr(self, "_h", None):
            self.lib.sps_fdm_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shape(self, layout):
        return interior_shape(self.desc) if layout == abi.LAYOUT_INTERIOR else padded_shape(self.desc)

    def upload(self, field: int, values: np.ndarray, layout: int = abi.LAYOUT_INTERIOR):
        a = np.ascontiguousarray(values, dtype=self.dtype)
        if a.shape != self._shape(layout):
            raise ValueError(f"field shape {a.shape}, expected {self._shape(layout)}")
        abi.check(self.lib.sps_fdm_upload(self._h, int(field), a.ctypes.data_as(c_void_p), layout))

    def download(self, field: int, layout: int = abi.LAYOUT_INTERIOR) -> np.ndarray:
        out = np.empty(self._shape(layout), dtype=self.dtype)
        abi.check(self.lib.sps_fdm_download(self._h, int(field), out.ctypes.data_as(c_void_p), layout))
        return out


    def checkpoint_text(self, field: int, index: int = 0, origin=None) -> bytes:
        """Field `field` as the reference's GNU checkpoint file (device-formatted)."""
        keep, o = _origin_arg(origin)
        cap = _text_cap(self.desc, 19, 1)
        buf = np.empty(cap, dtype=np.uint8)
        n = c_size_t(0)
        abi.check(self.lib.sps_fdm_checkpoint_text(self._h, int(field), int(index), o,
                                                   buf.ctypes.data_as(c_void_p), cap, byref(n)))
        return buf[:n.value].tobytes()

    def read_checkpoint(self, field: int, text: bytes) -> int:
        idx = c_int(0)
        abi.check(self.lib.sps_fdm_checkpoint_read(self._h, int(field), text, len(text),
                                                   byref(idx)))
        return idx.value

    def plot_text(self, field: int, fmt: int = abi.PLOT_GNU, origin=None) -> bytes:
        keep, o = _origin_arg(origin)
        cap = _text_cap(self.desc, 15, 15, 14 * (int(self.desc.n[0]) + 1) + 2)
        buf = np.empty(cap, dtype=np.uint8)
        n = c_size_t(0)
        abi.check(self.lib.sps_fdm_plot_text(self._h, int(field), int(fmt), o,
                                             buf.ctypes.data_as(c_void_p), cap, byref(n)))
        return buf[:n.value].tobytes()

    def checkpoint_xdr(self, field: int, index: int = 0, origin=None) -> bytes:
        keep, o = _origin_arg(origin)
        n = c_size_t(0)
        abi.check(self.lib.sps_fdm_checkpoint_xdr(self._h, int(field), int(index), o, None, 0,
                                                  byref(n)))
        buf = np.empty(n.value, dtype=np.uint8)
        abi.check(self.lib.sps_fdm_checkpoint_xdr(self._h, int(field), int(index), o,
                                                  buf.ctypes.data_as(c_void_p), buf.size,
                                                  byref(n)))
        return buf.tobytes()

    def read_checkpoint_xdr(self, field: int, data: bytes) -> int:
        idx = c_int(0)
        arr = np.frombuffer(data, dtype=np.uint8)
        abi.check(self.lib.sps_fdm_checkpoint_read_xdr(self._h, int(field),
                                                       arr.ctypes.data_as(c_void_p), arr.size,
                                                       byref(idx)))
        return idx.value

    def checkpoint_save(self, directory: str, index: int, single_file: bool = True,
                        done: bool = False, origin=None):
        """Every field's checkpoint (field f is id f) and one progress line."""
        keep, optr = _origin_arg(origin)
        abi.check(self.lib.sps_fdm_checkpoint_save(self._h, str(directory).encode(), int(index),
                                                   int(bool(single_file)), int(bool(done)), optr))

    def checkpoint_load(self, directory: str, index: int, single_file: bool = True) -> int:
        got = c_int(-1)
        abi.check(self.lib.sps_fdm_checkpoint_load(self._h, str(directory).encode(), int(index),
                                                   int(bool(single_file)), byref(got)))
        return got.value

    def step(self, nsteps: int = 1):
        abi.check(self.lib.sps_fdm_step(self._h, int(nsteps)))

    def sync(self):
        abi.check(self.lib.sps_fdm_sync(self._h))

    def query(self):
        path, spl = c_int(), c_int()
        abi.check(self.lib.sps_fdm_query(self._h, byref(path), byref(spl)))
        return path.value, spl.value

    def tune(self, steps_per_launch: int = 0, rows_per_tile: int = 0,
             values_per_lane: int = 0):
        abi.check(self.lib.sps_fdm_tune(self._h, steps_per_launch, rows_per_tile,
                                        values_per_lane))

    def autotune(self):
        """Time launch shapes on the device (fields kept); returns
        (steps_per_launch, rows_per_tile, values_per_lane)."""
        k, r, v = c_int(), c_int(), c_int()
        abi.check(self.lib.sps_fdm_autotune(self._h, byref(k), byref(r), byref(v)))
        return k.value, r.value, v.value

    def set_stream(self, stream_handle: int):
        abi.check(self.lib.sps_fdm_set_stream(self._h, c_void_p(stream_handle)))


class RegionalPlan:
    """A field on a regional grid (SolverSystemFDwSD): stored and stepped only
    on a box that follows the sites above a cutoff (sps_fdr_* in
    include/sps_fd.h).  Upload and download take the whole domain."""

    def __init__(self, desc: abi.Desc, params: "abi.Regional | None" = None):
        self.lib = abi.load()
        self.desc = desc
        self.dtype = _np_dtype(desc)
        h = c_void_p()
        pp = byref(params) if params is not None else None
        abi.check(self.lib.sps_fdr_create(byref(desc), pp, byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self.lib.sps_fdr_destroy(self._h)
            self._h = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shape(self, layout):
        return interior_shape(self.desc) if layout == abi.LAYOUT_INTERIOR else padded_shape(self.desc)

    def upload(self, field: np.ndarray, layout: int = abi.LAYOUT_INTERIOR):
        arr = np.ascontiguousarray(field, dtype=self.dtype)
        if arr.shape != self._shape(layout):
            raise ValueError(f"field shape {arr.shape} != {self._shape(layout)}")
        abi.check(self.lib.sps_fdr_upload(self._h, arr.ctypes.data_as(c_void_p), layout))

    def download(self, layout: int = abi.LAYOUT_INTERIOR) -> np.ndarray:
        out = np.empty(self._shape(layout), dtype=self.dtype)
        abi.check(self.lib.sps_fdr_download(self._h, out.ctypes.data_as(c_void_p), layout))
        return out

    def step(self, nsteps: int = 1):
        abi.check(self.lib.sps_fdr_step(self._h, int(nsteps)))

    def region(self) -> "abi.RegionInfo":
        info = abi.RegionInfo()
        abi.check(self.lib.sps_fdr_region(self._h, byref(info)))
        return info

    def sync(self):
        abi.check(self.lib.sps_fdr_sync(self._h))
