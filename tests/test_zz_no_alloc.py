"""The allocation-free reduction entry point (sps_fd_reduce_device) against
sps_fd_reduce on the same fields.

Both run the same two kernels in the same summation order, so the sums must be
bit-identical.  Two calls in a row share the GPU's scratch array: the second,
on a smaller field, must not see anything the first left there.  Argument
checks run before any HIP call and are tested without a GPU.
"""
import ctypes

import numpy as np
import pytest

from symphas_amd import abi, inputs, models
from symphas_amd.solver import FDPlan, interior_shape


def _call(lib, dim, dtype, n, h, energy, field_ptr, sums_ptr, stream=None):
    nn = (ctypes.c_int32 * 3)(*(list(n) + [1] * (3 - len(n))))
    eptr = ctypes.byref(energy) if energy is not None else None
    return lib.sps_fd_reduce_device(dim, dtype, nn, h, eptr, field_ptr, sums_ptr, stream)


def test_rejects_bad_arguments_before_any_hip_call():
    lib = abi.load()
    buf = ctypes.c_double(0.0)
    p = ctypes.cast(ctypes.byref(buf), ctypes.c_void_p)
    cases = [
        (4, abi.F32, (8, 8), 1.0, "dim must be 2 or 3"),
        (2, 7, (8, 8), 1.0, "dtype"),
        (2, abi.F32, (0, 8), 1.0, "out of range"),
        (2, abi.F32, (8, (1 << 20) + 1), 1.0, "out of range"),
        (3, abi.F64, (1 << 11, 1 << 11, 1 << 10), 1.0, "exceeds the maximum"),
        (2, abi.F64, (8, 8), 0.0, "h must be positive"),
    ]
    for dim, dtype, n, h, msg in cases:
        rc = _call(lib, dim, dtype, n, h, None, p, p)
        assert rc == abi.E_INVALID, (dim, dtype, n, h)
        assert msg in lib.sps_fd_last_error().decode()
    rc = _call(lib, 2, abi.F32, (8, 8), 1.0, None, None, p)
    assert rc == abi.E_INVALID
    bad = abi.Energy(1.0, 1.0, 1.0, 5, 0)
    rc = _call(lib, 2, abi.F32, (8, 8), 1.0, bad, p, p)
    assert rc == abi.E_INVALID
    assert "gradient" in lib.sps_fd_last_error().decode()


def _plan_sums(desc, psi, energy):
    with FDPlan(desc) as p:
        p.upload(psi)
        p.set_energy(energy)
        return p.reduce()


@pytest.mark.gpu
def test_no_alloc_reduce_matches_plan_reduce(gpu_lib):
    import torch

    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    # first call: Model B 2D 256 x 192 FP32 with a non-default functional;
    # second call: a smaller 3D FP64 field with the default functional
    d1 = models.model_b((256, 192), dtype=abi.F32)
    d2 = models.model_b((12, 10, 9), dtype=abi.F64,
                        stencil=models.Stencil3d2h(7, 10, 25), dt=0.005)
    cases = [
        (d1, inputs.uniform(interior_shape(d1), seed=31).astype(np.float32),
         abi.Energy(0.7, 1.3, 1.9, abi.GRAD_CENTRAL, 1), torch.float32),
        (d2, inputs.uniform(interior_shape(d2), seed=32), None, torch.float64),
    ]
    sums = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
    for desc, psi, energy, tdtype in cases:
        want_m, want_e = _plan_sums(desc, psi, energy)
        field = torch.from_numpy(np.ascontiguousarray(psi)).to("cuda", tdtype)
        n = [desc.n[a] for a in range(desc.dim)]
        print(f"no-alloc reduce: dim {desc.dim} n {n} dtype {desc.dtype}", flush=True)
        rc = _call(gpu_lib, desc.dim, desc.dtype, n, desc.h, energy,
                   ctypes.c_void_p(field.data_ptr()), ctypes.c_void_p(sums.data_ptr()),
                   ctypes.c_void_p(stream.cuda_stream))
        abi.check(rc)
        torch.cuda.synchronize()
        got_m, got_e = sums.cpu().tolist()
        assert got_m == want_m, (got_m, want_m)
        assert got_e == want_e, (got_e, want_e)
