// multi.hip - coupled-field plans (sps_fdm_*): two-field fused streaming
// kernel, plan state and the C entry points.
//
// The fused kernel streams rows like the single-field one (fused_tile.hpp): a
// wave owns a strip of 64 lanes x VEC columns, halo lanes feed the
// neighbours, and K time steps are pipelined per launch; here every stage
// carries both fields, because each field's right-hand side reads both.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

#include "ckpt.hpp"
#include "fused_tile.hpp"
#include "lanes.hpp"
#include "multi.hpp"
#include "plan.hpp"
#include "textio.hpp"

namespace sps {
namespace {

// sum_i x^i * sum_j c[i][j] y^j, two-level Horner
template <typename T>
__device__ __forceinline__ T bipoly(const T (&c)[kMDeg + 1][kMDeg + 1], T x,
                                    T y) {
  T r = T(0);
#pragma unroll
  for (int i = kMDeg; i >= 0; --i) {
    T ci = c[i][kMDeg];
#pragma unroll
    for (int j = kMDeg - 1; j >= 0; --j) ci = fmad(ci, y, c[i][j]);
    r = i == kMDeg ? ci : fmad(r, x, ci);
  }
  return r;
}

// Total degree <= 3: sum_i x^i * sum_{j <= 3-i} c[i][j] y^j.
template <typename T>
__device__ __forceinline__ T tripoly(const T (&c)[kMDeg + 1][kMDeg + 1], T x,
                                     T y) {
  const T r2 = fmad(c[2][1], y, c[2][0]);
  const T r1 = fmad(fmad(c[1][2], y, c[1][1]), y, c[1][0]);
  const T r0 = fmad(fmad(fmad(c[0][3], y, c[0][2]), y, c[0][1]), y, c[0][0]);
  return fmad(fmad(fmad(c[3][0], x, r2), x, r1), x, r0);
}

// The polynomial of class cls at every (x[j], y[j]), into p[j].
template <typename T, int VEC>
__device__ __forceinline__ void eval_poly(int cls,
                                          const T (&c)[kMDeg + 1][kMDeg + 1],
                                          const T* x, const T* y, T* p) {
  if (cls == kPolyZero) {
#pragma unroll
    for (int j = 0; j < VEC; ++j) p[j] = T(0);
  } else if (cls == kPolyLin) {
#pragma unroll
    for (int j = 0; j < VEC; ++j)
      p[j] = fmad(c[1][0], x[j], fmad(c[0][1], y[j], c[0][0]));
  } else if (cls == kPolyXLin) {
    // (c00 + c01 y + c02 y^2 + c03 y^3) + x (c10 + c11 y + c12 y^2)
#pragma unroll
    for (int j = 0; j < VEC; ++j) {
      const T r1 = fmad(fmad(c[1][2], y[j], c[1][1]), y[j], c[1][0]);
      const T r0 =
          fmad(fmad(fmad(c[0][3], y[j], c[0][2]), y[j], c[0][1]), y[j], c[0][0]);
      p[j] = fmad(r1, x[j], r0);
    }
  } else if (cls == kPolyYLin) {
    // (c00 + c10 x + c20 x^2 + c30 x^3) + y (c01 + c11 x + c21 x^2)
#pragma unroll
    for (int j = 0; j < VEC; ++j) {
      const T r1 = fmad(fmad(c[2][1], x[j], c[1][1]), x[j], c[0][1]);
      const T r0 =
          fmad(fmad(fmad(c[3][0], x[j], c[2][0]), x[j], c[1][0]), x[j], c[0][0]);
      p[j] = fmad(r1, y[j], r0);
    }
  } else if (cls == kPolySep) {
    // (c10 x + c20 x^2 + c30 x^3) + (c00 + c01 y + c02 y^2 + c03 y^3)
#pragma unroll
    for (int j = 0; j < VEC; ++j) {
      const T g =
          fmad(fmad(fmad(c[0][3], y[j], c[0][2]), y[j], c[0][1]), y[j], c[0][0]);
      p[j] = fmad(fmad(fmad(c[3][0], x[j], c[2][0]), x[j], c[1][0]), x[j], g);
    }
  } else if (cls == kPolyTri) {
#pragma unroll
    for (int j = 0; j < VEC; ++j) p[j] = tripoly(c, x[j], y[j]);
  } else {
#pragma unroll
    for (int j = 0; j < VEC; ++j) p[j] = bipoly(c, x[j], y[j]);
  }
}

// LIN (radius 2 only): 1 when field 1, 2 when field 0 has no bilaplacian and
// B_f = b_f * psi_f (+ a constant).  That field's mu_f is b_f * psi_f, so its
// first stage keeps the unscaled D(psi_f) and its update is
// psi_f + Adt_f + k1_f * D(psi_f), k1_f = kd * b_f: the radius-one form, run
// in step with the other field's two stages.
template <typename T, int VEC, int R, int LIN>
struct Stage2 {
  T P[2][3][VEC];  // input rows of both fields (slots rotate with the phase)
  T M[2][3][VEC];  // mu rows (radius 2)

  __device__ __forceinline__ void clear() {
#pragma unroll
    for (int f = 0; f < 2; ++f)
#pragma unroll
      for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int j = 0; j < VEC; ++j) P[f][s][j] = M[f][s][j] = T(0);
  }

  template <int C, int U, int D>
  __device__ __forceinline__ void diff5(const T (&A)[3][VEC],
                                        T (&d)[VEC]) const {
    const T left0 = from_prev_lane(A[C][VEC - 1]);
    const T rightN = from_next_lane(A[C][0]);
#pragma unroll
    for (int j = 0; j < VEC; ++j) {
      const T l = j == 0 ? left0 : A[C][j - 1];
      const T r = j == VEC - 1 ? rightN : A[C][j + 1];
      d[j] = fmad(T(-4), A[C][j], (l + r) + (A[U][j] + A[D][j]));
    }
  }

  template <int PH>
  __device__ __forceinline__ void push(const MFusedCoef<T>& c,
                                       T (&v)[2][VEC]) {
    constexpr int S0 = PH % 3, S1 = (PH + 1) % 3, S2 = (PH + 2) % 3;
#pragma unroll
    for (int f = 0; f < 2; ++f)
#pragma unroll
      for (int j = 0; j < VEC; ++j) P[f][S2][j] = v[f][j];
    T d[2][VEC], a[2][VEC];
    diff5<S1, S2, S0>(P[0], d[0]);
    diff5<S1, S2, S0>(P[1], d[1]);
    if constexpr (R == 1) {
      eval_poly<T, VEC>(c.adt_class[0], c.adt[0], P[0][S1], P[1][S1], a[0]);
      eval_poly<T, VEC>(c.adt_class[1], c.adt[1], P[0][S1], P[1][S1], a[1]);
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        v[0][j] = fmad(c.k1[0], d[0][j], P[0][S1][j] + a[0][j]);
        v[1][j] = fmad(c.k1[1], d[1][j], P[1][S1][j] + a[1][j]);
      }
    } else {
      constexpr bool lin0 = LIN == 2, lin1 = LIN == 1;
      if constexpr (!lin0)
        eval_poly<T, VEC>(c.b_class[0], c.b[0], P[0][S1], P[1][S1], a[0]);
      if constexpr (!lin1)
        eval_poly<T, VEC>(c.b_class[1], c.b[1], P[0][S1], P[1][S1], a[1]);
#pragma unroll
      for (int f = 0; f < 2; ++f) {
        const bool lin = f == 0 ? lin0 : lin1;
#pragma unroll
        for (int j = 0; j < VEC; ++j)
          M[f][S2][j] = lin ? d[f][j] : fmad(c.kmu[f], d[f][j], a[f][j]);
      }
      if constexpr (!lin0) diff5<S1, S2, S0>(M[0], d[0]);
      if constexpr (!lin1) diff5<S1, S2, S0>(M[1], d[1]);
      eval_poly<T, VEC>(c.adt_class[0], c.adt[0], P[0][S0], P[1][S0], a[0]);
      eval_poly<T, VEC>(c.adt_class[1], c.adt[1], P[0][S0], P[1][S0], a[1]);
#pragma unroll
      for (int j = 0; j < VEC; ++j) {
        v[0][j] = lin0 ? fmad(c.k1[0], M[0][S1][j], P[0][S0][j] + a[0][j])
                       : fmad(c.kd, d[0][j], P[0][S0][j] + a[0][j]);
        v[1][j] = lin1 ? fmad(c.k1[1], M[1][S1][j], P[1][S0][j] + a[1][j])
                       : fmad(c.kd, d[1][j], P[1][S0][j] + a[1][j]);
      }
    }
  }
};

template <typename T, int VEC, int K, int R, int LIN>
__global__ __launch_bounds__(256) void fused_multi(
    const T* __restrict__ in0, const T* __restrict__ in1, T* __restrict__ out0,
    T* __restrict__ out1, int X, int Y, int TY, int nstrips, int nitems,
    MFusedCoef<T> c) {
  constexpr int HL = (K * R + VEC - 1) / VEC;
  constexpr int W = (64 - 2 * HL) * VEC;
  constexpr int LAG = K * R;
  const int lane = threadIdx.x & 63;
  const int item = __builtin_amdgcn_readfirstlane(
      xcd_remap(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6));
  if (item >= nitems) return;
  const int strip = item % nstrips;
  const int chunk = item / nstrips;
  const int cb = strip * W + (lane - HL) * VEC;
  int col = cb % X;
  if (col < 0) col += X;
  const bool store_lane = lane >= HL && lane < 64 - HL && cb < X;
  const int y0 = chunk * TY;
  const int ty = min(TY, Y - y0);
  const int NI = ty + 2 * LAG;
  int rload = (y0 - LAG) % Y;
  if (rload < 0) rload += Y;

  Stage2<T, VEC, R, LIN> st[K];
#pragma unroll
  for (int s = 0; s < K; ++s) st[s].clear();

  // The schedule of the single-field stream (stream_tile): the fill rows
  // (0 .. 2*K*R-1, stage s from row 2*R*s on) unrolled, then groups of D
  // rows with no control flow, then the last group with a check per row.
  // Rows are prefetched D ahead into a register ring (both fields), and
  // loaded and stored through a wave-uniform buffer descriptor per field
  // and row; lanes that do not store get an offset past the row, whose
  // stores the hardware drops.
  constexpr int FILL = 2 * LAG;
  constexpr int D = 3;  // a multiple of 3 (phase = row mod 3)
  const int row_bytes = X * (int)sizeof(T);
  const unsigned ld_off = (unsigned)col * (unsigned)sizeof(T);
  const unsigned st_off = store_lane ? ld_off : 0x80000000u;
  T ring[D][2][VEC];
  auto load_row = [&](T (&b)[2][VEC]) {
    const int64_t o = (int64_t)rload * X;
    buf_load<T, VEC>(buffer_rsrc(in0 + o, row_bytes), ld_off, 0, b[0]);
    buf_load<T, VEC>(buffer_rsrc(in1 + o, row_bytes), ld_off, 0, b[1]);
    rload = rload + 1 == Y ? 0 : rload + 1;
  };
  static_for<0, D>([&](auto u) { load_row(ring[decltype(u)::value]); });

  auto run_row = [&](auto ic, auto nsc, T (&v)[2][VEC]) {
    constexpr int I = decltype(ic)::value, NS = decltype(nsc)::value;
#pragma unroll
    for (int f = 0; f < 2; ++f)
#pragma unroll
      for (int j = 0; j < VEC; ++j) v[f][j] = ring[I % D][f][j];
    static_for<0, NS>([&](auto sc) {
      st[decltype(sc)::value].template push<I % 3>(c, v);
    });
  };
  auto store_row = [&](int orow, const T (&v)[2][VEC]) {
    const int64_t o = (int64_t)(y0 + orow) * X;
    buf_store<T, VEC>(buffer_rsrc(out0 + o, row_bytes), st_off, 0, v[0]);
    buf_store<T, VEC>(buffer_rsrc(out1 + o, row_bytes), st_off, 0, v[1]);
  };

  static_for<0, FILL>([&](auto ic) {
    constexpr int I = decltype(ic)::value;
    constexpr int NS = I / (2 * R) + 1 < K ? I / (2 * R) + 1 : K;
    T v[2][VEC];
    run_row(ic, std::integral_constant<int, NS>{}, v);
    load_row(ring[I % D]);
  });
  int i = FILL;
  for (; i + 2 * D <= NI; i += D) {
    static_for<0, D>([&](auto uc) {
      constexpr int U = decltype(uc)::value, I = FILL + U;
      T v[2][VEC];
      run_row(std::integral_constant<int, I>{}, std::integral_constant<int, K>{}, v);
      load_row(ring[I % D]);
      store_row(i + U - FILL, v);
    });
  }
  for (; i < NI; i += D) {
    static_for<0, D>([&](auto uc) {
      constexpr int U = decltype(uc)::value, I = FILL + U;
      if (i + U < NI) {
        T v[2][VEC];
        run_row(std::integral_constant<int, I>{}, std::integral_constant<int, K>{}, v);
        if (i + U + D < NI) load_row(ring[I % D]);
        store_row(i + U - FILL, v);
      }
    });
  }
}

template <typename T, int VEC, int K, int R, int LIN>
int launch_multi_v(const T* const* in, T* const* out, int X, int Y, int TY,
                   const MFusedCoef<T>& c, hipStream_t s) {
  constexpr int HL = (K * R + VEC - 1) / VEC;
  constexpr int W = (64 - 2 * HL) * VEC;
  const int nstrips = (X + W - 1) / W;
  const int nchunks = (Y + TY - 1) / TY;
  const int nitems = nstrips * nchunks;
  fused_multi<T, VEC, K, R, LIN><<<(nitems + 3) / 4, 256, 0, s>>>(
      in[0], in[1], out[0], out[1], X, Y, TY, nstrips, nitems, c);
  return SPS_OK;
}

// 16-byte lanes (VEC = 16 / sizeof(T)) or 8-byte lanes (VEC = 8 / sizeof(T))
template <typename T, int K, int R, int LIN>
int launch_multi_k(const T* const* in, T* const* out, int X, int Y, int TY,
                   int vec, const MFusedCoef<T>& c, hipStream_t s) {
  if (vec == 16 / (int)sizeof(T))
    return launch_multi_v<T, 16 / sizeof(T), K, R, LIN>(in, out, X, Y, TY, c, s);
  return launch_multi_v<T, 8 / sizeof(T), K, R, LIN>(in, out, X, Y, TY, c, s);
}

template <typename T, int K>
int launch_multi_r(const T* const* in, T* const* out, int X, int Y, int TY,
                   int vec, const MFusedCoef<T>& c, hipStream_t s) {
  if (c.radius == 1) return launch_multi_k<T, K, 1, 0>(in, out, X, Y, TY, vec, c, s);
  if (c.lin == 1) return launch_multi_k<T, K, 2, 1>(in, out, X, Y, TY, vec, c, s);
  if (c.lin == 2) return launch_multi_k<T, K, 2, 2>(in, out, X, Y, TY, vec, c, s);
  return launch_multi_k<T, K, 2, 0>(in, out, X, Y, TY, vec, c, s);
}

// Private-segment (scratch) bytes per lane of one kernel instance, as the
// compiler's resource report records it: nonzero when the instance spills
// its registers to scratch memory.
template <typename T, int VEC, int K, int R, int LIN>
size_t multi_scratch_v() {
  hipFuncAttributes a{};
  if (hipFuncGetAttributes(&a, (const void*)fused_multi<T, VEC, K, R, LIN>) !=
      hipSuccess)
    return 0;
  return a.localSizeBytes;
}

template <typename T, int K, int R, int LIN>
size_t multi_scratch_k(int vec) {
  if (vec == 16 / (int)sizeof(T)) return multi_scratch_v<T, 16 / sizeof(T), K, R, LIN>();
  return multi_scratch_v<T, 8 / sizeof(T), K, R, LIN>();
}

template <typename T, int K>
size_t multi_scratch_r(int vec, const MFusedCoef<T>& c) {
  if (c.radius == 1) return multi_scratch_k<T, K, 1, 0>(vec);
  if (c.lin == 1) return multi_scratch_k<T, K, 2, 1>(vec);
  if (c.lin == 2) return multi_scratch_k<T, K, 2, 2>(vec);
  return multi_scratch_k<T, K, 2, 0>(vec);
}

}  // namespace

// Scratch bytes per lane of the 2D coupled kernel a launch of this shape
// would run (0 for a shape it does not build).
template <typename T>
size_t fused_multi_scratch(int K, int vec, const MFusedCoef<T>& c) {
  switch (K) {
    case 1: return multi_scratch_r<T, 1>(vec, c);
    case 2: return multi_scratch_r<T, 2>(vec, c);
    case 3: return multi_scratch_r<T, 3>(vec, c);
    case 4: return multi_scratch_r<T, 4>(vec, c);
    default: return 0;
  }
}
template size_t fused_multi_scratch<float>(int, int, const MFusedCoef<float>&);
template size_t fused_multi_scratch<double>(int, int, const MFusedCoef<double>&);

template <typename T>
int launch_fused_multi(const T* const* in, T* const* out, int X, int Y, int K,
                       int TY, int vec, const MFusedCoef<T>& c, hipStream_t s) {
  if ((vec != 8 / (int)sizeof(T) && vec != 16 / (int)sizeof(T)) ||
      X % vec != 0 || TY < 1) {
    set_error("fused coupled step: X=%d must be a multiple of %d values per "
              "lane (8 or 16 bytes)", X, vec);
    return SPS_E_UNSUPPORTED;
  }
  int rc;
  switch (K) {
    case 1: rc = launch_multi_r<T, 1>(in, out, X, Y, TY, vec, c, s); break;
    case 2: rc = launch_multi_r<T, 2>(in, out, X, Y, TY, vec, c, s); break;
    case 3: rc = launch_multi_r<T, 3>(in, out, X, Y, TY, vec, c, s); break;
    case 4: rc = launch_multi_r<T, 4>(in, out, X, Y, TY, vec, c, s); break;
    default:
      set_error("fused coupled step: steps per launch %d not in 1..4", K);
      return SPS_E_INVALID;
  }
  if (rc != SPS_OK) return rc;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("fused coupled launch failed: %s", hipGetErrorString(e));
    return SPS_E_HIP;
  }
  return SPS_OK;
}

template int launch_fused_multi<float>(const float* const*, float* const*, int,
                                       int, int, int, int, const MFusedCoef<float>&,
                                       hipStream_t);
template int launch_fused_multi<double>(const double* const*, double* const*,
                                        int, int, int, int, int,
                                        const MFusedCoef<double>&, hipStream_t);

namespace {

// ---------------------------------------------------------------- 3D

// radius-one 7-point difference (unscaled) at output row r of a ring whose
// plane S1 is centred on row r + 1; each lane move pairs with a register
// operand so both fold into v_add_f32_dpp
template <int S0, int S1, int S2, typename T, int NI>
__device__ __forceinline__ T d7m(const T (&a)[3][NI], int r) {
  const T ctr = a[S1][r + 1];
  const T xl = from_prev_lane(ctr), xr = from_next_lane(ctr);
  return fmad(T(-6), ctr,
              ((xl + a[S1][r]) + (xr + a[S1][r + 2])) +
                  (a[S0][r + 1] + a[S2][r + 1]));
}

template <int N>
struct MRowsOf {
  static constexpr int value = N > 0 ? N : 1;
};

// Two coupled fields in 3D, streamed along z like fused3d_stream: one wave
// owns 64 columns (one per lane, R*K halo lanes per side) x TYW rows of a
// z-chunk; K steps are a chain of radius-one stages (radius two: mu, then
// update), every stage carrying both fields in rings of three planes.
template <typename T, int TYW, int R, int K, int LIN>
__global__ __launch_bounds__(256) void fused3d_multi(
    const T* __restrict__ in0, const T* __restrict__ in1, T* __restrict__ out0,
    T* __restrict__ out1, int X, int Y, int Z, int TZ, int nstrips, int nrowblk,
    int nitems, MFusedCoef<T> c) {
  constexpr int RT = R * K;
  constexpr int NS = (R == 2 ? 2 : 1) * K;
  static_assert(NS >= 1 && NS <= 2, "at most two stages");
  constexpr int HL = RT;
  constexpr int W = 64 - 2 * HL;
  constexpr int NP = TYW + 2 * RT;
  const int lane = threadIdx.x & 63;
  const int item = __builtin_amdgcn_readfirstlane(
      xcd_remap(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6));
  if (item >= nitems) return;
  const int strip = item % nstrips;
  const int rb = (item / nstrips) % nrowblk;
  const int zc = item / (nstrips * nrowblk);
  const int x = strip * W + lane - HL;
  int col = x % X;
  if (col < 0) col += X;
  const bool store_lane = lane >= HL && lane < 64 - HL && x < X;
  const int y0 = rb * TYW;
  const int z0 = zc * TZ;
  const int tz = min(TZ, Z - z0);
  const int64_t XY = (int64_t)X * Y;
  int rowoff[NP];
#pragma unroll
  for (int r = 0; r < NP; ++r) {
    int y = (y0 - RT + r) % Y;
    if (y < 0) y += Y;
    rowoff[r] = __builtin_amdgcn_readfirstlane(y * X);
  }
  const int NI = tz + 2 * RT;
  int zl = (z0 - RT) % Z;
  if (zl < 0) zl += Z;

  T A0[2][3][NP];
  T A1[2][3][MRowsOf<NS >= 2 ? NP - 2 : 0>::value];
#pragma unroll
  for (int f = 0; f < 2; ++f)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
#pragma unroll
      for (int r = 0; r < NP; ++r) A0[f][q][r] = T(0);
#pragma unroll
      for (int r = 0; r < MRowsOf<NS >= 2 ? NP - 2 : 0>::value; ++r)
        A1[f][q][r] = T(0);
    }

  T NX[2][NP];
  auto fetch = [&]() {
    const int64_t po = (int64_t)zl * XY;
#pragma unroll
    for (int r = 0; r < NP; ++r) {
      NX[0][r] = (in0 + po + rowoff[r])[col];
      NX[1][r] = (in1 + po + rowoff[r])[col];
    }
    zl = zl + 1 == Z ? 0 : zl + 1;
  };
  fetch();

  auto stage = [&](auto sc, auto PHc, auto& a, auto& b, auto& o, int i) {
    constexpr int S = decltype(sc)::value;
    constexpr int PH = decltype(PHc)::value;
    constexpr int S0 = PH % 3, S1 = (PH + 1) % 3, S2 = (PH + 2) % 3;
    constexpr int NO = NP - 2 * S;
    constexpr bool MU = R == 2 && (S % 2 == 1);
    constexpr bool LAST = S == NS;
    if constexpr (LAST) {
      if (i < 2 * NS) return;
    } else {
      if (i < 2 * (S - 1)) return;
    }
    T res[2][NO];
#pragma unroll
    for (int r = 0; r < NO; ++r) {
      // a radius-one field (LIN) keeps its unscaled D from the first stage
      // and its update reads that value instead of a second difference
      constexpr bool lin0 = R == 2 && LIN == 2, lin1 = R == 2 && LIN == 1;
      T d0 = T(0), d1 = T(0);
      if constexpr (MU || !lin0) d0 = d7m<S0, S1, S2>(a[0], r);
      if constexpr (MU || !lin1) d1 = d7m<S0, S1, S2>(a[1], r);
      if constexpr (MU) {
        const T xv = a[0][S1][r + 1], yv = a[1][S1][r + 1];
        T p0 = T(0), p1 = T(0);
        if constexpr (!lin0) eval_poly<T, 1>(c.b_class[0], c.b[0], &xv, &yv, &p0);
        if constexpr (!lin1) eval_poly<T, 1>(c.b_class[1], c.b[1], &xv, &yv, &p1);
        res[0][r] = lin0 ? d0 : fmad(c.kmu[0], d0, p0);
        res[1][r] = lin1 ? d1 : fmad(c.kmu[1], d1, p1);
      } else {
        T xv, yv;
        if constexpr (R == 2) {
          xv = b[0][S0][r + 2];
          yv = b[1][S0][r + 2];
        } else {
          xv = a[0][S1][r + 1];
          yv = a[1][S1][r + 1];
        }
        T p0, p1;
        eval_poly<T, 1>(c.adt_class[0], c.adt[0], &xv, &yv, &p0);
        eval_poly<T, 1>(c.adt_class[1], c.adt[1], &xv, &yv, &p1);
        res[0][r] = lin0 ? fmad(c.k1[0], a[0][S1][r + 1], xv + p0)
                         : fmad(R == 2 ? c.kd : c.k1[0], d0, xv + p0);
        res[1][r] = lin1 ? fmad(c.k1[1], a[1][S1][r + 1], yv + p1)
                         : fmad(R == 2 ? c.kd : c.k1[1], d1, yv + p1);
      }
    }
    if constexpr (LAST) {
      if (store_lane) {
        const int64_t po = (int64_t)(z0 + i - 2 * NS) * XY + col;
        if (y0 + NO <= Y) {
#pragma unroll
          for (int r = 0; r < NO; ++r) {
            out0[po + rowoff[r + RT]] = res[0][r];
            out1[po + rowoff[r + RT]] = res[1][r];
          }
        } else {
#pragma unroll
          for (int r = 0; r < NO; ++r)
            if (y0 + r < Y) {
              out0[po + rowoff[r + RT]] = res[0][r];
              out1[po + rowoff[r + RT]] = res[1][r];
            }
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < NO; ++r) {
        o[0][S2][r] = res[0][r];
        o[1][S2][r] = res[1][r];
      }
    }
  };

  auto phase = [&](auto PHc, int i) {
    constexpr int PH = decltype(PHc)::value;
    constexpr int S2 = (PH + 2) % 3;
#pragma unroll
    for (int f = 0; f < 2; ++f)
#pragma unroll
      for (int r = 0; r < NP; ++r) A0[f][S2][r] = NX[f][r];
    if (i + 1 < NI) fetch();
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    stage(I1{}, PHc, A0, A0, A1, i);
    if constexpr (NS >= 2) stage(I2{}, PHc, A1, A0, A0, i);
  };
  for (int i = 0; i < NI; i += 3) {
    phase(std::integral_constant<int, 0>{}, i);
    if (i + 1 < NI) phase(std::integral_constant<int, 1>{}, i + 1);
    if (i + 2 < NI) phase(std::integral_constant<int, 2>{}, i + 2);
  }
}

template <typename T, int TYW, int R, int K, int LIN>
int launch3m(const T* const* in, T* const* out, int X, int Y, int Z, int TZ,
             const MFusedCoef<T>& c, hipStream_t s) {
  constexpr int W = 64 - 2 * R * K;
  const int nstrips = (X + W - 1) / W;
  const int nrowblk = (Y + TYW - 1) / TYW;
  const int nzc = (Z + TZ - 1) / TZ;
  const int64_t nitems = (int64_t)nstrips * nrowblk * nzc;
  if (nitems > ((int64_t)1 << 31) - 4) {
    set_error("fused coupled 3D step: too many work items");
    return SPS_E_INVALID;
  }
  fused3d_multi<T, TYW, R, K, LIN><<<(unsigned)((nitems + 3) / 4), 256, 0, s>>>(
      in[0], in[1], out[0], out[1], X, Y, Z, TZ, nstrips, nrowblk, (int)nitems,
      c);
  return SPS_OK;
}

}  // namespace

// 3D shapes: radius one K in {1, 2}, radius two K = 1; 4 or 8 rows per wave
bool fused3d_multi_shape_ok(int radius, int K, int rows) {
  return (rows == 4 || rows == 8) && K >= 1 && K <= (radius == 1 ? 2 : 1);
}

template <typename T>
int launch_fused3d_multi(const T* const* in, T* const* out, int X, int Y,
                         int Z, int K, int TY, int TZ, const MFusedCoef<T>& c,
                         hipStream_t s) {
  if (!fused3d_multi_shape_ok(c.radius, K, TY) || TZ < 1) {
    set_error("fused coupled 3D step: unsupported shape (%d steps, %d rows)", K,
              TY);
    return SPS_E_INVALID;
  }
  int rc;
  if (c.radius == 1) {
    if (K == 1)
      rc = TY == 4 ? launch3m<T, 4, 1, 1, 0>(in, out, X, Y, Z, TZ, c, s)
                   : launch3m<T, 8, 1, 1, 0>(in, out, X, Y, Z, TZ, c, s);
    else
      rc = TY == 4 ? launch3m<T, 4, 1, 2, 0>(in, out, X, Y, Z, TZ, c, s)
                   : launch3m<T, 8, 1, 2, 0>(in, out, X, Y, Z, TZ, c, s);
  } else if (c.lin == 1) {
    rc = TY == 4 ? launch3m<T, 4, 2, 1, 1>(in, out, X, Y, Z, TZ, c, s)
                 : launch3m<T, 8, 2, 1, 1>(in, out, X, Y, Z, TZ, c, s);
  } else if (c.lin == 2) {
    rc = TY == 4 ? launch3m<T, 4, 2, 1, 2>(in, out, X, Y, Z, TZ, c, s)
                 : launch3m<T, 8, 2, 1, 2>(in, out, X, Y, Z, TZ, c, s);
  } else {
    rc = TY == 4 ? launch3m<T, 4, 2, 1, 0>(in, out, X, Y, Z, TZ, c, s)
                 : launch3m<T, 8, 2, 1, 0>(in, out, X, Y, Z, TZ, c, s);
  }
  if (rc != SPS_OK) return rc;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("fused coupled 3D launch failed: %s", hipGetErrorString(e));
    return SPS_E_HIP;
  }
  return SPS_OK;
}

template int launch_fused3d_multi<float>(const float* const*, float* const*,
                                         int, int, int, int, int, int,
                                         const MFusedCoef<float>&, hipStream_t);
template int launch_fused3d_multi<double>(const double* const*, double* const*,
                                          int, int, int, int, int, int,
                                          const MFusedCoef<double>&,
                                          hipStream_t);

}  // namespace sps

// ------------------------------------------------------------------ plans

struct sps_fdm_plan {
  sps_fdm_desc d;
  sps::MLowered low;
  int path = SPS_PATH_REFERENCE;
  hipStream_t stream = nullptr, own_stream = nullptr;
  size_t nsites = 0, esize = 0;
  void* buf[SPS_MAX_FIELDS][2] = {};
  int cur = 0;
  bool has_data = false;
  int K = 1, TY = 32, VEC = 1;
  int TZ = 64;  // 3D: z planes per streamed chunk
  sps::MRefParams<float> rp32;
  sps::MRefParams<double> rp64;
  sps::MFusedCoef<float> fc32;
  sps::MFusedCoef<double> fc64;

  int launch_one(int k);
  int steps_per_launch() const { return path == SPS_PATH_FUSED ? K : 1; }
};

namespace sps {
namespace {

int validate_mdesc(const sps_fdm_desc* d) {
  if (!d) {
    set_error("descriptor is NULL");
    return SPS_E_INVALID;
  }
  // geometry and stencil checks of the single-field descriptor
  sps_fd_desc g;
  std::memset(&g, 0, sizeof(g));
  g.dim = d->dim;
  g.dtype = d->dtype;
  for (int a = 0; a < 3; ++a) g.n[a] = d->n[a];
  g.lap_points = d->lap_points;
  g.gradlap_points = d->gradlap_points;
  g.bilap_points = d->bilap_points;
  g.path = d->path;
  g.h = d->h;
  g.dt = d->dt;
  SPS_TRY(validate_desc(&g));
  if (d->nfields < 1 || d->nfields > SPS_MAX_FIELDS) {
    set_error("nfields %d not in [1, %d]", d->nfields, SPS_MAX_FIELDS);
    return SPS_E_INVALID;
  }
  for (int f = 0; f < d->nfields; ++f) {
    if (d->nterms[f] < 0 || d->nterms[f] > SPS_MAX_MTERMS) {
      set_error("field %d: nterms %d not in [0, %d]", f, d->nterms[f],
                SPS_MAX_MTERMS);
      return SPS_E_INVALID;
    }
    for (int t = 0; t < d->nterms[f]; ++t) {
      const sps_fdm_term& tm = d->terms[f][t];
      if (tm.op < SPS_OP_ID || tm.op > SPS_OP_GRADLAP) {
        set_error("field %d term %d: unknown op %d", f, t, tm.op);
        return SPS_E_INVALID;
      }
      int deg = 0;
      for (int g2 = 0; g2 < SPS_MAX_FIELDS; ++g2) {
        if (tm.pow[g2] < 0 || (g2 >= d->nfields && tm.pow[g2] != 0)) {
          set_error("field %d term %d: exponent of field %d invalid", f, t, g2);
          return SPS_E_INVALID;
        }
        deg += tm.pow[g2];
      }
      if (deg > SPS_MAX_DEGREE) {
        set_error("field %d term %d: degree %d above %d", f, t, deg,
                  SPS_MAX_DEGREE);
        return SPS_E_INVALID;
      }
      if (tm.op == SPS_OP_GRADLAP) {
        const int ncomp = (d->dim == 2 && d->gradlap_points == 6) ? 2 : 1;
        if (tm.axis < 0 || tm.axis >= d->dim || tm.component < 0 ||
            tm.component >= ncomp) {
          set_error("field %d term %d: gradlaplacian axis/component", f, t);
          return SPS_E_INVALID;
        }
      }
      if (!std::isfinite(tm.coeff)) {
        set_error("field %d term %d: coefficient not finite", f, t);
        return SPS_E_INVALID;
      }
    }
  }
  return SPS_OK;
}

MLowered lower_multi(const sps_fdm_desc& d) {
  MLowered lw;
  if (d.nfields != 2) {
    lw.why_not = "the fused coupled kernels take two fields";
    return lw;
  }
  // lap5 o lap5 = bilap13 in 2D, lap7 o lap7 = bilap25 in 3D
  const int lap_pts = d.dim == 2 ? 5 : 7, bilap_pts = d.dim == 2 ? 13 : 25;
  bool any_lap = false, any_bilap = false;
  for (int f = 0; f < 2; ++f)
    for (int t = 0; t < d.nterms[f]; ++t) {
      const sps_fdm_term& tm = d.terms[f][t];
      const int p0 = tm.pow[0], p1 = tm.pow[1];
      if (p0 > kMDeg || p1 > kMDeg) {
        lw.why_not = "a field power above 3";
        return lw;
      }
      switch (tm.op) {
        case SPS_OP_ID:
          lw.a[f][p0][p1] += tm.coeff;
          break;
        case SPS_OP_LAP:
          any_lap = true;
          if (p0 + p1 > 0) lw.b[f][p0][p1] += tm.coeff;
          break;
        case SPS_OP_BILAP:
          if (tm.pow[f] != 1 || p0 + p1 != 1) {
            lw.why_not = "bilaplacian of other than the field itself";
            return lw;
          }
          any_bilap = true;
          lw.kc[f] += tm.coeff;
          break;
        default:
          lw.why_not = "gradlaplacian terms";
          return lw;
      }
    }
  if (any_lap && d.lap_points != lap_pts) {
    lw.why_not = "laplacian stencil is not the 5-point (2D) / 7-point (3D) one";
    return lw;
  }
  if (any_bilap && d.bilap_points != bilap_pts) {
    lw.why_not = "bilaplacian stencil is not the 13-point (2D) / 25-point (3D) one";
    return lw;
  }
  // radius one when every B_f is b_f * psi_f and there is no bilaplacian
  bool r1 = lw.kc[0] == 0.0 && lw.kc[1] == 0.0;
  for (int f = 0; f < 2 && r1; ++f)
    for (int i = 0; i <= kMDeg; ++i)
      for (int j = 0; j <= kMDeg; ++j) {
        const bool own = (f == 0 && i == 1 && j == 0) || (f == 1 && i == 0 && j == 1);
        if (!own && lw.b[f][i][j] != 0.0) r1 = false;
      }
  lw.radius = r1 ? 1 : 2;
  lw.fusable = true;
  return lw;
}

// The smallest class whose evaluation covers every nonzero coefficient.
int poly_class(const double (&c)[kMDeg + 1][kMDeg + 1]) {
  int maxdeg = -1;
  for (int i = 0; i <= kMDeg; ++i)
    for (int j = 0; j <= kMDeg; ++j)
      if (c[i][j] != 0.0) maxdeg = std::max(maxdeg, i + j);
  if (maxdeg < 0) return kPolyZero;
  if (maxdeg <= 1) return kPolyLin;
  if (maxdeg > 3) return kPolyDense;
  bool xlin = true, ylin = true, sep = true;
  for (int i = 0; i <= kMDeg; ++i)
    for (int j = 0; j <= kMDeg; ++j)
      if (c[i][j] != 0.0) {
        if (i > 1) xlin = false;
        if (j > 1) ylin = false;
        if (i > 0 && j > 0) sep = false;
      }
  if (xlin) return kPolyXLin;
  if (ylin) return kPolyYLin;
  if (sep) return kPolySep;
  return kPolyTri;
}

template <typename T>
void make_mfused_coef(const sps_fdm_desc& d, const MLowered& lw,
                      MFusedCoef<T>& c) {
  std::memset(&c, 0, sizeof(c));
  const double divh = 1.0 / d.h, divh2 = divh * divh;
  for (int f = 0; f < 2; ++f) {
    for (int i = 0; i <= kMDeg; ++i)
      for (int j = 0; j <= kMDeg; ++j) {
        c.adt[f][i][j] = T(d.dt * lw.a[f][i][j]);
        c.b[f][i][j] = T(lw.b[f][i][j]);
      }
    c.kmu[f] = T(lw.kc[f] * divh2);
    const double bown = f == 0 ? lw.b[0][1][0] : lw.b[1][0][1];
    c.k1[f] = T(d.dt * divh2 * bown);
  }
  c.kd = T(d.dt * divh2);
  c.radius = lw.radius;
  // a radius-two plan's field without a bilaplacian whose B is b*psi_f (plus
  // a constant, whose laplacian is zero) takes the radius-one form
  c.lin = 0;
  if (lw.radius == 2)
    for (int f = 0; f < 2; ++f) {
      bool own = lw.kc[f] == 0.0 && (f == 0 ? lw.b[0][1][0] : lw.b[1][0][1]) != 0.0;
      for (int i = 0; i <= kMDeg && own; ++i)
        for (int j = 0; j <= kMDeg; ++j)
          if (i + j >= 1 && !(f == 0 ? (i == 1 && j == 0) : (i == 0 && j == 1)) &&
              lw.b[f][i][j] != 0.0)
            own = false;
      if (own) c.lin = f == 1 ? 1 : 2;
    }
  for (int f = 0; f < 2; ++f) {
    c.adt_class[f] = poly_class(lw.a[f]);
    c.b_class[f] = poly_class(lw.b[f]);
  }
}

template <typename T>
void make_mref_params(const sps_fdm_desc& d, MRefParams<T>& p) {
  std::memset(&p, 0, sizeof(p));
  for (int a = 0; a < 3; ++a) p.n[a] = a < d.dim ? d.n[a] : 1;
  p.L = d.lap_points;
  p.G = d.gradlap_points;
  p.B = d.bilap_points;
  p.nfields = d.nfields;
  const double divh = 1.0 / d.h, divh2 = divh * divh;
  p.divh2 = T(divh2);
  p.divh3 = T(divh * divh2);
  p.divh4 = T(divh2 * divh2);
  p.dt = T(d.dt);
  for (int f = 0; f < d.nfields; ++f) {
    p.nterms[f] = d.nterms[f];
    for (int t = 0; t < d.nterms[f]; ++t) {
      const sps_fdm_term& tm = d.terms[f][t];
      MDevTerm<T>& dt = p.terms[f][t];
      dt.op = tm.op;
      dt.axis = tm.axis;
      dt.component = tm.component;
      int deg = 0, which = -1;
      for (int g2 = 0; g2 < SPS_MAX_FIELDS; ++g2) {
        dt.pow[g2] = tm.pow[g2];
        deg += tm.pow[g2];
        if (tm.pow[g2] == 1) which = g2;
      }
      dt.single = deg == 1 ? which : -1;
      dt.coeff = T(tm.coeff);
    }
  }
}

}  // namespace
}  // namespace sps

int sps_fdm_plan::launch_one(int k) {
  using namespace sps;
  int rc;
  if (path == SPS_PATH_REFERENCE) {
    if (d.dtype == SPS_F32) {
      MRefParams<float> p = rp32;
      for (int f = 0; f < d.nfields; ++f) {
        p.in[f] = (const float*)buf[f][cur];
        p.out[f] = (float*)buf[f][1 - cur];
      }
      rc = launch_ref_multi<float>(p, d.dim, stream);
    } else {
      MRefParams<double> p = rp64;
      for (int f = 0; f < d.nfields; ++f) {
        p.in[f] = (const double*)buf[f][cur];
        p.out[f] = (double*)buf[f][1 - cur];
      }
      rc = launch_ref_multi<double>(p, d.dim, stream);
    }
    if (rc == SPS_OK) {
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) {
        set_error("coupled reference launch failed: %s", hipGetErrorString(e));
        rc = SPS_E_HIP;
      }
    }
  } else if (d.dim == 3) {
    if (d.dtype == SPS_F32) {
      const float* in[2] = {(const float*)buf[0][cur], (const float*)buf[1][cur]};
      float* out[2] = {(float*)buf[0][1 - cur], (float*)buf[1][1 - cur]};
      rc = launch_fused3d_multi<float>(in, out, d.n[0], d.n[1], d.n[2], k, TY,
                                       TZ, fc32, stream);
    } else {
      const double* in[2] = {(const double*)buf[0][cur],
                             (const double*)buf[1][cur]};
      double* out[2] = {(double*)buf[0][1 - cur], (double*)buf[1][1 - cur]};
      rc = launch_fused3d_multi<double>(in, out, d.n[0], d.n[1], d.n[2], k, TY,
                                        TZ, fc64, stream);
    }
  } else if (d.dtype == SPS_F32) {
    const float* in[2] = {(const float*)buf[0][cur], (const float*)buf[1][cur]};
    float* out[2] = {(float*)buf[0][1 - cur], (float*)buf[1][1 - cur]};
    rc = launch_fused_multi<float>(in, out, d.n[0], d.n[1], k, TY, VEC, fc32,
                                   stream);
  } else {
    const double* in[2] = {(const double*)buf[0][cur],
                           (const double*)buf[1][cur]};
    double* out[2] = {(double*)buf[0][1 - cur], (double*)buf[1][1 - cur]};
    rc = launch_fused_multi<double>(in, out, d.n[0], d.n[1], k, TY, VEC, fc64,
                                    stream);
  }
  if (rc == SPS_OK) cur = 1 - cur;
  return rc;
}

using namespace sps;

extern "C" {

int sps_fdm_create(const sps_fdm_desc* desc, sps_fdm_plan** out) {
  if (!out) {
    set_error("output pointer is NULL");
    return SPS_E_INVALID;
  }
  *out = nullptr;
  SPS_TRY(validate_mdesc(desc));
  if (sps_fd_device_count() < 1) {
    set_error("no HIP device: this library runs the step on an MI355X only");
    return SPS_E_HIP;
  }
  sps_fdm_plan* p = new (std::nothrow) sps_fdm_plan();
  if (!p) {
    set_error("host allocation failed");
    return SPS_E_NOMEM;
  }
  p->d = *desc;
  p->esize = desc->dtype == SPS_F32 ? 4 : 8;
  p->nsites = (size_t)desc->n[0] * desc->n[1] * (desc->dim == 3 ? desc->n[2] : 1);
  p->low = lower_multi(*desc);
  const int vec = desc->dim == 3 ? 1 : 8 / (int)p->esize;
  const bool fusable = p->low.fusable && desc->n[0] % vec == 0;
  if (desc->path == SPS_PATH_FUSED && !fusable) {
    set_error("fused coupled path unavailable: %s",
              p->low.fusable ? "x extent not a multiple of the lane width"
                             : p->low.why_not);
    delete p;
    return SPS_E_UNSUPPORTED;
  }
  p->path = (desc->path == SPS_PATH_REFERENCE || !fusable) ? SPS_PATH_REFERENCE
                                                           : SPS_PATH_FUSED;
  make_mref_params(*desc, p->rp32);
  make_mref_params(*desc, p->rp64);
  if (p->path == SPS_PATH_FUSED) {
    make_mfused_coef(*desc, p->low, p->fc32);
    make_mfused_coef(*desc, p->low, p->fc64);
    if (desc->dim == 3) {
      // two fields per stage: radius one fuses 2 steps, radius two 1; the
      // z-chunk splits the grid into about one round of resident waves
      p->K = p->low.radius == 1 ? 2 : 1;
      p->TY = p->esize == 8 ? 4 : 8;  // FP64 8-row rings exceed 256 VGPRs
      p->VEC = 1;
      const int halo = p->low.radius * p->K;
      const int64_t per_chunk =
          (int64_t)((desc->n[0] + 63 - 2 * halo) / (64 - 2 * halo)) *
          ((desc->n[1] + p->TY - 1) / p->TY);
      const int64_t m = std::max<int64_t>(1, 2048 / per_chunk);
      p->TZ = (int)std::min<int64_t>(
          1024, std::max<int64_t>(16, (desc->n[2] + m - 1) / m));
    } else {
      p->K = p->low.radius == 1 ? 4 : 2;
      p->TY = 32;
      // 8-byte lanes: the 16-byte variant holds twice the rows in registers
      // and measured slower at every launch shape on 4096^2 FP32
      p->VEC = 8 / (int)p->esize;
    }
  }
  hipError_t e = hipStreamCreateWithFlags(&p->own_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    set_error("stream creation failed: %s", hipGetErrorString(e));
    delete p;
    return SPS_E_HIP;
  }
  p->stream = p->own_stream;
  for (int f = 0; f < desc->nfields; ++f)
    for (int i = 0; i < 2; ++i) {
      e = hipMalloc(&p->buf[f][i], p->nsites * p->esize);
      if (e != hipSuccess) {
        p->buf[f][i] = nullptr;
        set_error("hipMalloc of %zu bytes failed: %s", p->nsites * p->esize,
                  hipGetErrorString(e));
        sps_fdm_destroy(p);
        return SPS_E_NOMEM;
      }
    }
  *out = p;
  return SPS_OK;
}

int sps_fdm_destroy(sps_fdm_plan* p) {
  if (!p) return SPS_OK;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  if (p->own_stream && p->own_stream != p->stream)
    (void)hipStreamSynchronize(p->own_stream);
  for (auto& fb : p->buf)
    for (void* b : fb)
      if (b) (void)hipFree(b);
  if (p->own_stream) (void)hipStreamDestroy(p->own_stream);
  delete p;
  return SPS_OK;
}

static int check_field(const sps_fdm_plan* p, int field, const void* host,
                       int layout) {
  if (!p || !host) {
    set_error("plan or host pointer is NULL");
    return SPS_E_INVALID;
  }
  if (field < 0 || field >= p->d.nfields) {
    set_error("field %d not in [0, %d)", field, p->d.nfields);
    return SPS_E_INVALID;
  }
  if (layout != SPS_LAYOUT_INTERIOR && layout != SPS_LAYOUT_PADDED) {
    set_error("unknown layout %d", layout);
    return SPS_E_INVALID;
  }
  return SPS_OK;
}

// interior <-> padded copies on the host (ghosts written periodic)
static void padded_to_interior_m(const sps_fdm_desc& d, const char* in,
                                 char* out, size_t es) {
  const int D = SPS_BOUNDARY_DEPTH;
  const int64_t X = d.n[0] + 2 * D, Y = d.n[1] + 2 * D;
  const int nz = d.dim == 3 ? d.n[2] : 1;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < d.n[1]; ++y) {
      const int64_t src = ((int64_t)(d.dim == 3 ? z + D : 0) * Y + (y + D)) * X + D;
      const int64_t dst = ((int64_t)z * d.n[1] + y) * d.n[0];
      std::memcpy(out + dst * es, in + src * es, (size_t)d.n[0] * es);
    }
}

static void interior_to_padded_m(const sps_fdm_desc& d, const char* in,
                                 char* out, size_t es) {
  const int D = SPS_BOUNDARY_DEPTH;
  const int64_t X = d.n[0] + 2 * D, Y = d.n[1] + 2 * D;
  const int64_t Z = d.dim == 3 ? d.n[2] + 2 * D : 1;
  const int64_t nx = d.n[0], ny = d.n[1], nz = d.dim == 3 ? d.n[2] : 1;
  for (int64_t z = 0; z < Z; ++z)
    for (int64_t y = 0; y < Y; ++y)
      for (int64_t x = 0; x < X; ++x) {
        const int64_t ix = ((x - D) % nx + nx) % nx;
        const int64_t iy = ((y - D) % ny + ny) % ny;
        const int64_t iz = d.dim == 3 ? ((z - D) % nz + nz) % nz : 0;
        std::memcpy(out + ((z * Y + y) * X + x) * es,
                    in + ((iz * ny + iy) * nx + ix) * es, es);
      }
}

int sps_fdm_upload(sps_fdm_plan* p, int field, const void* host, int layout) {
  SPS_TRY(check_field(p, field, host, layout));
  const size_t bytes = p->nsites * p->esize;
  const void* src = host;
  std::vector<char> tmp;
  if (layout == SPS_LAYOUT_PADDED) {
    tmp.resize(bytes);
    padded_to_interior_m(p->d, (const char*)host, tmp.data(), p->esize);
    src = tmp.data();
  }
  SPS_HIP_TRY(hipMemcpyAsync(p->buf[field][p->cur], src, bytes,
                             hipMemcpyHostToDevice, p->stream));
  SPS_HIP_TRY(hipStreamSynchronize(p->stream));
  p->has_data = true;
  return SPS_OK;
}

int sps_fdm_download(sps_fdm_plan* p, int field, void* host, int layout) {
  SPS_TRY(check_field(p, field, host, layout));
  if (!p->has_data) {
    set_error("no field uploaded");
    return SPS_E_STATE;
  }
  const size_t bytes = p->nsites * p->esize;
  std::vector<char> tmp;
  void* dst = host;
  if (layout == SPS_LAYOUT_PADDED) {
    tmp.resize(bytes);
    dst = tmp.data();
  }
  SPS_HIP_TRY(hipMemcpyAsync(dst, p->buf[field][p->cur], bytes,
                             hipMemcpyDeviceToHost, p->stream));
  SPS_HIP_TRY(hipStreamSynchronize(p->stream));
  if (layout == SPS_LAYOUT_PADDED)
    interior_to_padded_m(p->d, tmp.data(), (char*)host, p->esize);
  return SPS_OK;
}

int sps_fdm_step(sps_fdm_plan* p, int nsteps) {
  if (!p || nsteps < 0) {
    set_error("plan is NULL or nsteps < 0");
    return SPS_E_INVALID;
  }
  if (!p->has_data) {
    set_error("no field uploaded");
    return SPS_E_STATE;
  }
  const int spl = p->steps_per_launch();
  for (int done = 0; done < nsteps;) {
    const int k = std::min(spl, nsteps - done);
    SPS_TRY(p->launch_one(k));
    done += k;
  }
  return SPS_OK;
}

int sps_fdm_query(const sps_fdm_plan* p, int* path, int* spl) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  if (path) *path = p->path;
  if (spl) *spl = p->steps_per_launch();
  return SPS_OK;
}

int sps_fdm_sync(sps_fdm_plan* p) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  SPS_HIP_TRY(hipStreamSynchronize(p->stream));
  return SPS_OK;
}

int sps_fdm_set_stream(sps_fdm_plan* p, void* hip_stream) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  SPS_HIP_TRY(hipStreamSynchronize(p->stream));
  p->stream = (hipStream_t)hip_stream;
  return SPS_OK;
}

int sps_fdm_tune(sps_fdm_plan* p, int steps_per_launch, int rows_per_tile,
                 int values_per_lane) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  if (steps_per_launch < 0 || steps_per_launch > 4 || rows_per_tile < 0 ||
      rows_per_tile > 512 || (rows_per_tile != 0 && rows_per_tile < 4)) {
    set_error("tune (coupled): steps_per_launch 0..4, rows_per_tile 0 or "
              "4..512");
    return SPS_E_INVALID;
  }
  if (p->d.dim == 3) {
    const int k = steps_per_launch ? steps_per_launch : p->K;
    const int rows = rows_per_tile ? rows_per_tile : p->TY;
    if ((values_per_lane != 0 && values_per_lane != 1) ||
        (p->path == SPS_PATH_FUSED &&
         !fused3d_multi_shape_ok(p->low.radius, k, rows))) {
      set_error("tune (coupled 3D): steps_per_launch 1..2 (radius one) or 1 "
                "(radius two), rows_per_tile 4 or 8, values_per_lane 0 or 1");
      return SPS_E_INVALID;
    }
    p->K = k;
    p->TY = rows;
    return SPS_OK;
  }
  if (values_per_lane != 0) {
    const int narrow = 8 / (int)p->esize, wide = 16 / (int)p->esize;
    if ((values_per_lane != narrow && values_per_lane != wide) ||
        p->d.n[0] % values_per_lane != 0) {
      set_error("tune (coupled): values_per_lane must be %d or %d and divide "
                "the x extent %d", narrow, wide, p->d.n[0]);
      return SPS_E_INVALID;
    }
  }
  if (steps_per_launch) p->K = steps_per_launch;
  if (rows_per_tile) p->TY = rows_per_tile;
  if (values_per_lane) p->VEC = values_per_lane;
  return SPS_OK;
}

}  // extern "C"

// field f of a coupled plan as a file-writer view (dim, n, h, dtype)
static int field_view(sps_fdm_plan* p, int field, bool need_data,
                      FieldView* v) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  if (field < 0 || field >= p->d.nfields) {
    set_error("field %d not in [0, %d)", field, p->d.nfields);
    return SPS_E_INVALID;
  }
  if (need_data && !p->has_data) {
    set_error("no field uploaded");
    return SPS_E_STATE;
  }
  std::memset(&v->d, 0, sizeof(v->d));
  v->d.dim = p->d.dim;
  v->d.dtype = p->d.dtype;
  for (int a = 0; a < 3; ++a) v->d.n[a] = p->d.n[a];
  v->d.h = p->d.h;
  v->data = p->buf[field][p->cur];
  v->nsites = p->nsites;
  v->stream = p->stream;
  return SPS_OK;
}

extern "C" {

int sps_fdm_checkpoint_text(sps_fdm_plan* p, int field, int index,
                            const double* origin, char* text, size_t capacity,
                            size_t* length) {
  FieldView v;
  SPS_TRY(field_view(p, field, true, &v));
  return file_checkpoint_text(v, index, origin, text, capacity, length);
}

int sps_fdm_checkpoint_read(sps_fdm_plan* p, int field, const char* text,
                            size_t length, int* index) {
  FieldView v;
  SPS_TRY(field_view(p, field, false, &v));
  SPS_TRY(file_checkpoint_read(v, text, length, index));
  p->has_data = true;
  return SPS_OK;
}

int sps_fdm_plot_text(sps_fdm_plan* p, int field, int format,
                      const double* origin, char* text, size_t capacity,
                      size_t* length) {
  FieldView v;
  SPS_TRY(field_view(p, field, true, &v));
  return file_plot_text(v, format, origin, text, capacity, length);
}

int sps_fdm_checkpoint_xdr(sps_fdm_plan* p, int field, int index,
                           const double* origin, void* data, size_t capacity,
                           size_t* length) {
  FieldView v;
  SPS_TRY(field_view(p, field, true, &v));
  return file_xdr_write(v, index, origin, data, capacity, length);
}

int sps_fdm_checkpoint_read_xdr(sps_fdm_plan* p, int field, const void* data,
                                size_t length, int* index) {
  FieldView v;
  SPS_TRY(field_view(p, field, false, &v));
  SPS_TRY(file_xdr_read(v, data, length, index));
  p->has_data = true;
  return SPS_OK;
}

int sps_fdm_checkpoint_save(sps_fdm_plan* p, const char* dir, int index,
                            int single_file, int done, const double* origin) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  // one file per field, the field's index being its id (io/inc/io.h:185-188)
  for (int f = 0; f < p->d.nfields; ++f) {
    SPS_TRY(ckpt_write_field(
        dir, f, index, single_file, [&](int idx, std::string& text) {
          size_t len = 0;
          SPS_TRY(sps_fdm_checkpoint_text(p, f, idx, origin, nullptr, 0, &len));
          text.assign(len, '\0');
          return sps_fdm_checkpoint_text(p, f, idx, origin, &text[0], len, &len);
        }));
  }
  return ckpt_record(dir, index, done);
}

int sps_fdm_checkpoint_load(sps_fdm_plan* p, const char* dir, int index,
                            int single_file, int* index_read) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  int first = -1;
  for (int f = 0; f < p->d.nfields; ++f) {
    int got = -1;
    SPS_TRY(ckpt_read_field(dir, f, index, single_file, &got,
                            [&](const char* text, size_t len, int* idx) {
                              return sps_fdm_checkpoint_read(p, f, text, len, idx);
                            }));
    if (f == 0) {
      first = got;
    } else if (got != first) {
      set_error("checkpoint of field %d is at index %d, field 0 at %d", f, got,
                first);
      return SPS_E_INVALID;
    }
  }
  if (index_read) *index_read = first;
  return SPS_OK;
}

int sps_fdm_autotune(sps_fdm_plan* p, int* out_k, int* out_rows,
                     int* out_vec) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  auto report = [&]() {
    if (out_k) *out_k = p->steps_per_launch();
    if (out_rows) *out_rows = p->TY;
    if (out_vec) *out_vec = p->VEC;
    return SPS_OK;
  };
  if (p->path != SPS_PATH_FUSED) return report();
  if (!p->has_data) {
    set_error("autotune needs uploaded fields");
    return SPS_E_STATE;
  }
  const size_t bytes = p->nsites * p->esize;
  const int nf = p->d.nfields;
  std::vector<void*> saved(nf, nullptr);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = SPS_OK;
  const int oldK = p->K, oldTY = p->TY, oldVEC = p->VEC, cur0 = p->cur;
  int bestK = oldK, bestTY = oldTY, bestVEC = oldVEC;
  std::vector<int> vecs;
  if (p->d.dim == 3)
    vecs.push_back(1);
  else
    for (int v : {16 / (int)p->esize, 8 / (int)p->esize})
      if (p->d.n[0] % v == 0) vecs.push_back(v);
  double best = 1e300;
  do {
    bool ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
    for (int f = 0; f < nf && ok; ++f)
      ok = hipMalloc(&saved[f], bytes) == hipSuccess &&
           hipMemcpyAsync(saved[f], p->buf[f][p->cur], bytes,
                          hipMemcpyDeviceToDevice, p->stream) == hipSuccess;
    if (!ok) {
      set_error("coupled autotune set-up failed");
      rc = SPS_E_HIP;
      break;
    }
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                              dev) != hipSuccess) {
      set_error("device query failed");
      rc = SPS_E_HIP;
      break;
    }
    for (int vec : vecs)
    for (int k = 1; k <= 4 && rc == SPS_OK; ++k) {
      std::vector<int> cand_rows;
      if (p->d.dim == 3) {
        for (int rows : {4, 8})
          if (fused3d_multi_shape_ok(p->low.radius, k, rows)) cand_rows.push_back(rows);
      } else {
        cand_rows = {16, 32, 64};
        // chunk heights that give every CU the same number m of 4-wave
        // workgroups (a CU with one more sets the launch's length)
        const int hl = (k * p->low.radius + vec - 1) / vec;
        const int w = (64 - 2 * hl) * vec;
        const int nstrips = (p->d.n[0] + w - 1) / w;
        for (int m = 1; m <= 6; ++m) {
          const long nch = (long)m * cus * 4 / nstrips;
          if (nch < 1) continue;
          const long ty = (p->d.n[1] + nch - 1) / nch;
          if (ty >= 8 && ty <= 512 &&
              std::find(cand_rows.begin(), cand_rows.end(), (int)ty) == cand_rows.end())
            cand_rows.push_back((int)ty);
        }
      }
      // an instance the compiler had to spill to scratch memory is left
      // out instead of timed (it runs several times slower than its
      // neighbours)
      if (p->d.dim == 2 &&
          (p->d.dtype == SPS_F32 ? fused_multi_scratch<float>(k, vec, p->fc32)
                                 : fused_multi_scratch<double>(k, vec, p->fc64)) > 0)
        continue;
      for (int rows : cand_rows) {
        p->K = k;
        p->TY = rows;
        p->VEC = vec;
        double us = 1e300;
        for (int rep = 0; rep < 3 && rc == SPS_OK; ++rep) {
          if ((rc = p->launch_one(k)) != SPS_OK) break;  // warm
          if (hipEventRecord(e0, p->stream) != hipSuccess) rc = SPS_E_HIP;
          for (int i = 0; i < 12 / k && rc == SPS_OK; ++i) rc = p->launch_one(k);
          if (rc != SPS_OK || hipEventRecord(e1, p->stream) != hipSuccess ||
              hipEventSynchronize(e1) != hipSuccess) {
            if (rc == SPS_OK) rc = SPS_E_HIP;
            break;
          }
          float ms = 0;
          if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = SPS_E_HIP;
            break;
          }
          if (rep > 0) us = std::min(us, 1e3 * ms / ((12 / k) * k));
        }
        if (rc != SPS_OK) break;
        if (us < best) {
          best = us;
          bestK = k;
          bestTY = rows;
          bestVEC = vec;
        }
      }
    }
    // restore the caller's fields
    p->cur = cur0;
    for (int f = 0; f < nf; ++f)
      if (saved[f] && hipMemcpyAsync(p->buf[f][p->cur], saved[f], bytes,
                                     hipMemcpyDeviceToDevice, p->stream) != hipSuccess &&
          rc == SPS_OK)
        rc = SPS_E_HIP;
    if (hipStreamSynchronize(p->stream) != hipSuccess && rc == SPS_OK) rc = SPS_E_HIP;
  } while (false);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (void* v : saved)
    if (v) (void)hipFree(v);
  if (rc != SPS_OK) {
    p->K = oldK;
    p->TY = oldTY;
    p->VEC = oldVEC;
    return rc;
  }
  p->K = bestK;
  p->TY = bestTY;
  p->VEC = bestVEC;
  return report();
}

}  // extern "C"
