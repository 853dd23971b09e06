// multi_ref.hip - reference-order kernel of the coupled-field step.
//
// Compiled with -ffp-contract=off like the single-field reference kernels:
// every product and sum rounds as written, so the FP64 instantiation gives
// the oracle's bits.  One thread per interior site computes every field's
// right-hand side from the current fields (terms in descriptor order, the
// reference's stencil groupings, a monomial operand evaluated at each stencil
// point as its temporary grid would hold it) and writes psi_f + dt * rhs_f.
// Periodic neighbours are found by wrapping coordinates.
#include "multi.hpp"
#include "stencil_ref.hpp"

namespace sps {
namespace {

struct Offsets {
  int64_t xo[5], yo[5], zo[5];  // element offsets of coordinate c-2 .. c+2
};

// The monomial of the term at element offset off: the left-fold product of
// its factors in field order (psi*rho*rho = (psi*rho)*rho), 1 if empty.
template <typename T>
struct MonoSrc {
  const MRefParams<T>* p;
  const MDevTerm<T>* t;
  __device__ __forceinline__ T operator()(int64_t off) const {
    if (t->single >= 0) return p->in[t->single][off];
    T acc = T(1);
    bool started = false;
    for (int f = 0; f < p->nfields; ++f)
      for (int k = 0; k < t->pow[f]; ++k) {
        const T x = p->in[f][off];
        acc = started ? acc * x : x;
        started = true;
      }
    return acc;
  }
};

// (i, j, k) along get_stride<AX>'s direction vectors to grid offsets
// (datatypes/inc/grid.h:537-578): X: (+x, +y, +z); Y: (+y, -x, +z);
// Z: (-z, +y, -x).
template <typename T, int DIM, int AX, class Src>
struct MAcc {
  const Offsets* o;
  Src src;
  __device__ __forceinline__ T operator()(int i, int j, int k) const {
    int dx, dy, dz;
    if (AX == 0) {
      dx = i; dy = j; dz = k;
    } else if (AX == 1) {
      dx = -j; dy = i; dz = k;
    } else {
      dx = -k; dy = j; dz = -i;
    }
    int64_t off = o->xo[dx + 2] + o->yo[dy + 2];
    if (DIM == 3) off += o->zo[dz + 2];
    return src(off);
  }
};

template <typename T, int DIM, int AX, class Src>
__device__ __forceinline__ T gradlap_m(const MRefParams<T>& p, const Offsets& o,
                                       Src s, int component) {
  MAcc<T, DIM, AX, Src> a{&o, s};
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

template <typename T, int DIM, class Src>
__device__ __forceinline__ T derivative_m(const MRefParams<T>& p,
                                          const Offsets& o, Src s, int dop,
                                          int axis, int component) {
  if (dop == SPS_OP_LAP) {
    MAcc<T, DIM, 0, Src> a{&o, s};
    if (DIM == 2)
      return p.L == 5 ? ref::lap2_5<T>(a, p.divh2) : ref::lap2_9<T>(a, p.divh2);
    switch (p.L) {
      case 7: return ref::lap3_7<T>(a, p.divh2);
      case 15: return ref::lap3_15<T>(a, p.divh2);
      case 19: return ref::lap3_19<T>(a, p.divh2);
      case 21: return ref::lap3_21<T>(a, p.divh2);
      default: return ref::lap3_27<T>(a, p.divh2);
    }
  }
  if (dop == SPS_OP_BILAP) {
    MAcc<T, DIM, 0, Src> a{&o, s};
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
  if (axis == 0) return gradlap_m<T, DIM, 0>(p, o, s, component);
  if (axis == 1 || DIM == 2) return gradlap_m<T, DIM, 1>(p, o, s, component);
  return gradlap_m<T, DIM, 2>(p, o, s, component);
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

template <typename T, int DIM>
__global__ __launch_bounds__(256) void ref_multi_kernel(MRefParams<T> p) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= p.n[0] || y >= p.n[1]) return;
  const int64_t nx = p.n[0], nxy = (int64_t)p.n[0] * p.n[1];
  Offsets o;
#pragma unroll
  for (int d = -2; d <= 2; ++d) {
    o.xo[d + 2] = wrap_near(x + d, p.n[0]);
    o.yo[d + 2] = (int64_t)wrap_near(y + d, p.n[1]) * nx;
    o.zo[d + 2] = DIM == 3 ? (int64_t)wrap_near(z + d, p.n[2]) * nxy : 0;
  }
  const int64_t idx = x + nx * y + nxy * z;
  for (int f = 0; f < p.nfields; ++f) {
    T acc = T(0);
    for (int t = 0; t < p.nterms[f]; ++t) {
      const MDevTerm<T>& tm = p.terms[f][t];
      MonoSrc<T> src{&p, &tm};
      T val;
      if (tm.op == SPS_OP_ID) {
        val = tm.coeff * src(idx);
      } else {
        val = tm.coeff * derivative_m<T, DIM>(p, o, src, tm.op, tm.axis,
                                               tm.component);
      }
      acc = t == 0 ? val : acc + val;
    }
    // forward Euler per field (examples/solvers/solverft.h:46-50)
    p.out[f][idx] = p.in[f][idx] + p.dt * acc;
  }
}

}  // namespace

template <typename T>
int launch_ref_multi(const MRefParams<T>& p, int dim, hipStream_t stream) {
  dim3 block(64, 4, 1);
  dim3 grid((p.n[0] + 63) / 64, (p.n[1] + 3) / 4, dim == 3 ? p.n[2] : 1);
  if (dim == 2)
    ref_multi_kernel<T, 2><<<grid, block, 0, stream>>>(p);
  else
    ref_multi_kernel<T, 3><<<grid, block, 0, stream>>>(p);
  return SPS_OK;
}

template int launch_ref_multi<float>(const MRefParams<float>&, int, hipStream_t);
template int launch_ref_multi<double>(const MRefParams<double>&, int,
                                      hipStream_t);

}  // namespace sps
