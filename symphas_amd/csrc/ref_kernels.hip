// ref_kernels.hip - reference-order finite-difference kernels.
//
// Compiled with -ffp-contract=off: every product and sum rounds as written,
// so the FP64 instantiation reproduces the CPU SolverFT arithmetic bit for bit
// (same stencil groupings, same term order, same Euler form).
//
// One thread per interior site of the unpadded field; periodic neighbours are
// found by wrapping coordinates, which replaces the reference's ghost refresh
// (datatypes/inc/boundaryupdate.h:154-356,650-708) with modular indexing - the
// two are the same map for every ghost a stencil of radius <= 3 can touch.
#include "common.hpp"
#include "ref_kernels.hpp"
#include "stencil_ref.hpp"

namespace sps {
namespace {

template <typename T, int DIM>
struct Window {
  const T* f;
  int64_t xo[5], yo[5], zo[5];  // element offsets of coordinate c-2 .. c+2
};

// Operand = the field itself.
template <typename T>
struct IdentityOp {
  __device__ __forceinline__ T operator()(T x) const { return x; }
};

// Operand = sum_k c[k] x^k in increasing degree, x^k as a product of k copies
// (lib/inc/spslibmath.h:255-269); zero coefficients are skipped and the first
// non-zero one starts the sum.  This is the value the reference stores in the
// temporary grid of OpDerivative (sym/inc/expressionderivatives.h:1583-1588).
template <typename T>
__device__ __forceinline__ T poly_eval(const T* c, T x) {
  T acc = T(0);
  bool started = false;
  for (int k = 0; k <= kMaxDeg; ++k) {
    if (c[k] == T(0)) continue;
    T p = T(1);
    if (k >= 1) {
      p = x;
      for (int j = 1; j < k; ++j) p = p * x;
    }
    T t = c[k] * p;
    acc = started ? acc + t : t;
    started = true;
  }
  return acc;
}

template <typename T>
struct PolyOp {
  const T* c;
  __device__ __forceinline__ T operator()(T x) const { return poly_eval(c, x); }
};

// Maps a stencil offset (i, j, k) along get_stride<AX>'s direction vectors to
// grid offsets (dx, dy, dz) (datatypes/inc/grid.h:537-578):
//   X: s = (+x, +y, +z);  Y: s = (+y, -x, +z);  Z: s = (-z, +y, -x).
template <typename T, int DIM, int AX, class Op>
struct Acc {
  const Window<T, DIM>* w;
  Op op;
  __device__ __forceinline__ T operator()(int i, int j, int k) const {
    int dx, dy, dz;
    if (AX == 0) {
      dx = i; dy = j; dz = k;
    } else if (AX == 1) {
      dx = -j; dy = i; dz = k;
    } else {
      dx = -k; dy = j; dz = -i;
    }
    int64_t off = w->xo[dx + 2] + w->yo[dy + 2];
    if (DIM == 3) off += w->zo[dz + 2];
    return op(w->f[off]);
  }
};

template <typename T, int DIM, int AX, class Op>
__device__ __forceinline__ T gradlap_axis(const RefParams<T>& p,
                                          const Window<T, DIM>& w, Op op,
                                          int component) {
  Acc<T, DIM, AX, Op> a{&w, op};
  if (DIM == 2) {
    switch (p.G) {
      case 6: return ref::gradlap2_6<T>(a, p.divh3, component);
      case 8: return ref::gradlap2_8<T>(a, p.divh3);
      case 12: return ref::gradlap2_12<T>(a, p.divh3);
      default: return ref::gradlap2_16<T>(a, p.divh3);
    }
  }
  switch (p.G) {
    case 10: return ref::gradlap3_10<T>(a, p.divh3);
    case 12: return ref::gradlap3_12<T>(a, p.divh3);
    case 28: return ref::gradlap3_28<T>(a, p.divh3);
    case 36: return ref::gradlap3_36<T>(a, p.divh3);
    default: return ref::gradlap3_40<T>(a, p.divh3);
  }
}

template <typename T, int DIM, class Op>
__device__ __forceinline__ T derivative(const RefParams<T>& p,
                                        const Window<T, DIM>& w, Op op,
                                        int dop, int axis, int component) {
  if (dop == SPS_OP_LAP) {
    Acc<T, DIM, 0, Op> a{&w, op};
    if (DIM == 2) {
      return p.L == 5 ? ref::lap2_5<T>(a, p.divh2) : ref::lap2_9<T>(a, p.divh2);
    }
    switch (p.L) {
      case 7: return ref::lap3_7<T>(a, p.divh2);
      case 15: return ref::lap3_15<T>(a, p.divh2);
      case 19: return ref::lap3_19<T>(a, p.divh2);
      case 21: return ref::lap3_21<T>(a, p.divh2);
      default: return ref::lap3_27<T>(a, p.divh2);
    }
  }
  if (dop == SPS_OP_BILAP) {
    Acc<T, DIM, 0, Op> a{&w, op};
    if (DIM == 2) {
      switch (p.B) {
        case 13: return ref::bilap2_13<T>(a, p.divh4);
        case 17: return ref::bilap2_17<T>(a, p.divh4);
        default: return ref::bilap2_21<T>(a, p.divh4);
      }
    }
    switch (p.B) {
      case 21: return ref::bilap3_21<T>(a, p.divh4);
      case 25: return ref::bilap3_25<T>(a, p.divh4);
      case 41: return ref::bilap3_41<T>(a, p.divh4);
      case 52: return ref::bilap3_52<T>(a, p.divh4);
      default: return ref::bilap3_57<T>(a, p.divh4);
    }
  }
  if (axis == 0) return gradlap_axis<T, DIM, 0>(p, w, op, component);
  if (axis == 1 || DIM == 2) return gradlap_axis<T, DIM, 1>(p, w, op, component);
  return gradlap_axis<T, DIM, 2>(p, w, op, component);
}

__device__ __forceinline__ int wrap(int c, int n) {
  c %= n;
  return c < 0 ? c + n : c;
}

// c in [-2, n + 2): a neighbour at most two sites away.  For n >= 3 one
// conditional add or subtract gives the same index as wrap without an
// integer division.
__device__ __forceinline__ int wrap_near(int c, int n) {
  if (n < 3) return wrap(c, n);
  return c < 0 ? c + n : (c >= n ? c - n : c);
}

// The value one site takes: mode 0 psi + dt*rhs, 1 rhs, 2 one derivative of
// psi.  `in` may point to global memory or to LDS (flat addressing); the
// arithmetic is the same either way.
// BOX: p.n are the dims of a box whose boundary layer (BOUNDARY_DEPTH cells
// per side) holds the neighbours of its interior sites, so offsets are taken
// as they are; otherwise the array is the unpadded periodic field and
// coordinates wrap.
template <typename T, int DIM, bool BOX = false>
__device__ __forceinline__ T site_value(const RefParams<T>& p, const T* in,
                                        int x, int y, int z) {
  const int64_t nx = p.n[0], nxy = (int64_t)p.n[0] * p.n[1];
  Window<T, DIM> w;
  w.f = in;
#pragma unroll
  for (int d = -2; d <= 2; ++d) {
    if (BOX) {
      w.xo[d + 2] = x + d;
      w.yo[d + 2] = (int64_t)(y + d) * nx;
      w.zo[d + 2] = DIM == 3 ? (int64_t)(z + d) * nxy : 0;
    } else {
      w.xo[d + 2] = wrap_near(x + d, p.n[0]);
      w.yo[d + 2] = (int64_t)wrap_near(y + d, p.n[1]) * nx;
      w.zo[d + 2] = DIM == 3 ? (int64_t)wrap_near(z + d, p.n[2]) * nxy : 0;
    }
  }
  const int64_t idx = x + nx * y + nxy * z;
  if (p.mode == 2)
    return derivative<T, DIM>(p, w, IdentityOp<T>{}, p.apply_op, p.apply_axis,
                              p.apply_component);
  T acc = T(0);
  for (int t = 0; t < p.nterms; ++t) {
    const DevTerm<T>& tm = p.terms[t];
    // the operand's grid: the field, or a provisional variable's grid,
    // which the reference refreshes periodically like the field
    // (sol/inc/modelspecialized.h:430-434); same offsets either way
    Window<T, DIM> wo = w;
    if (tm.operand > 0) wo.f = p.vars[tm.operand];
    T val;
    if (tm.op == SPS_OP_ID) {
      val = tm.coeff * poly_eval(tm.poly, wo.f[idx]);
    } else if (tm.identity) {
      val = tm.coeff * derivative<T, DIM>(p, wo, IdentityOp<T>{}, tm.op,
                                          tm.axis, tm.component);
    } else {
      val = tm.coeff * derivative<T, DIM>(p, wo, PolyOp<T>{tm.poly}, tm.op,
                                          tm.axis, tm.component);
    }
    acc = t == 0 ? val : acc + val;
  }
  // forward Euler psi + dt*dframe (examples/solvers/solverft.h:46-50)
  return p.mode == 0 ? in[idx] + p.dt * acc : acc;
}

template <typename T, int DIM>
__global__ __launch_bounds__(256) void ref_kernel(const T* __restrict__ in,
                                                  T* __restrict__ out,
                                                  RefParams<T> p) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= p.n[0] || y >= p.n[1]) return;
  const int64_t idx = x + (int64_t)p.n[0] * (y + (int64_t)p.n[1] * z);
  out[idx] = site_value<T, DIM>(p, in, x, y, z);
}

// One thread per site of a box (a regional grid's stored region): interior
// sites take psi + dt*rhs, boundary-layer sites keep their value.
template <typename T, int DIM>
__global__ __launch_bounds__(256) void ref_box_kernel(const T* __restrict__ in,
                                                      T* __restrict__ out,
                                                      RefParams<T> p) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= p.n[0] || y >= p.n[1]) return;
  const int64_t idx = x + (int64_t)p.n[0] * (y + (int64_t)p.n[1] * z);
  constexpr int b = SPS_BOUNDARY_DEPTH;
  const bool interior = x >= b && x < p.n[0] - b && y >= b && y < p.n[1] - b &&
                        (DIM == 2 || (z >= b && z < p.n[2] - b));
  out[idx] = interior ? site_value<T, DIM, true>(p, in, x, y, z) : in[idx];
}

}  // namespace

template <typename T>
int launch_ref_box(const T* in, T* out, const RefParams<T>& p, int dim,
                   hipStream_t stream) {
  dim3 block(64, 4, 1);
  dim3 grid((p.n[0] + 63) / 64, (p.n[1] + 3) / 4, dim == 3 ? p.n[2] : 1);
  if (dim == 2)
    ref_box_kernel<T, 2><<<grid, block, 0, stream>>>(in, out, p);
  else
    ref_box_kernel<T, 3><<<grid, block, 0, stream>>>(in, out, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("regional step launch failed: %s", hipGetErrorString(e));
    return SPS_E_HIP;
  }
  return SPS_OK;
}

template int launch_ref_box<float>(const float*, float*,
                                   const RefParams<float>&, int, hipStream_t);
template int launch_ref_box<double>(const double*, double*,
                                    const RefParams<double>&, int,
                                    hipStream_t);

template <typename T>
int launch_ref(const T* in, T* out, const RefParams<T>& p, int dim,
               hipStream_t stream) {
  dim3 block(64, 4, 1);
  dim3 grid((p.n[0] + 63) / 64, (p.n[1] + 3) / 4, dim == 3 ? p.n[2] : 1);
  if (dim == 2)
    ref_kernel<T, 2><<<grid, block, 0, stream>>>(in, out, p);
  else
    ref_kernel<T, 3><<<grid, block, 0, stream>>>(in, out, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("reference kernel launch failed: %s", hipGetErrorString(e));
    return SPS_E_HIP;
  }
  return SPS_OK;
}

template int launch_ref<float>(const float*, float*, const RefParams<float>&,
                               int, hipStream_t);
template int launch_ref<double>(const double*, double*,
                                const RefParams<double>&, int, hipStream_t);

}  // namespace sps
