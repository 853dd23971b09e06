// fused3d.hip - fused 3D steps: register streaming along z.
//
// One wave owns a column block of 64 lanes (one x column per lane) by TYW
// rows and streams the planes of one z-chunk.  K steps are fused per launch
// as a chain of radius-one stages, each holding a ring of three planes in
// registers: x neighbours come from adjacent lanes (DPP), y and z neighbours
// are registers, and there is no LDS and no barrier.
//
// With the 7-point laplacian, lap7∘lap7 is the 25-point bilaplacian of
// Stencil3d2h<7,*,25> exactly, so a radius-two step is the canonical form
//   mu = B(psi) + kmu*D7(psi),  psi' = psi + Adt(psi) + kd*D7(mu)
// (two stages: mu, then update), which is the reference's
// -bilap psi - lap psi + lap psi^3 up to rounding; a radius-one step is one
// stage psi' = psi + Adt(psi) + k1*D7(psi).  The stages use the folded
// coefficients of FusedCoef (S6 = sum of the 6 neighbours):
//   mu~ = S6 + psi*p(psi),  psi' = psi + q(psi) + k2*(S6[mu~] - 6*mu~),
//   radius one: psi' = psi + q(psi) + k1*(S6 - 6*psi).
// Every stage is evaluated with the same expression whatever K, so K fused
// steps are bitwise K launches.
//
// Plane pipeline: while input plane i is loaded, stage s produces plane i - s
// of its output from planes i-s-1 .. i-s+1 of its input.  Input plane 0 lies
// R*K planes below the chunk and the NS stages total radius R*K, so the last
// stage stores chunk plane i - 2*NS.  Rows shrink by one per stage on each
// side (TYW + 2*R*K rows are loaded) and so do the valid lanes (R*K halo
// lanes per side are loaded but not stored).
#include "fused.hpp"
#include "lanes.hpp"

#include <type_traits>

namespace sps {
namespace {

// sum of the 6 neighbours at output row r of a ring whose plane S1 is
// centred on row r + 1
template <int S0, int S1, int S2, typename T, int NI>
__device__ __forceinline__ T n6(const T (&a)[3][NI], int r) {
  const T ctr = a[S1][r + 1];
  const T xl = from_prev_lane(ctr), xr = from_next_lane(ctr);
  // each lane move pairs with a register operand, so both fold into
  // v_add_f32_dpp
  return ((xl + a[S1][r]) + (xr + a[S1][r + 2])) +
         (a[S0][r + 1] + a[S2][r + 1]);
}

template <int N>
struct RowsOf {
  static constexpr int value = N > 0 ? N : 1;
};

template <typename T, int TYW, int R, int K, bool HASA, int POLY>
__global__ __launch_bounds__(256) void fused3d_stream(
    const T* __restrict__ in, T* __restrict__ out, int X, int Y, int Z, int TZ,
    int nstrips, int nrowblk, int nitems, FusedCoef<T> c) {
  constexpr int RT = R * K;                 // radius of K steps
  constexpr int NS = (R == 2 ? 2 : 1) * K;  // radius-one stages
  static_assert(NS >= 1 && NS <= 4, "at most four stages");
  constexpr int HL = RT;  // halo lanes per side
  constexpr int W = 64 - 2 * HL;
  constexpr int NP = TYW + 2 * RT;  // rows of the loaded planes
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
  // wave-uniform row offsets (scalar registers); the lane's column is added
  // in the address only
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

  // rings: A0 = psi planes, A(s) = output planes of stage s < NS
  T A0[3][NP];
  T A1[3][RowsOf<NS >= 2 ? NP - 2 : 0>::value];
  T A2[3][RowsOf<NS >= 3 ? NP - 4 : 0>::value];
  T A3[3][RowsOf<NS >= 4 ? NP - 6 : 0>::value];
  auto zero = [](auto& a) {
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int r = 0; r < (int)(sizeof(a[0]) / sizeof(T)); ++r) a[q][r] = T(0);
  };
  zero(A0);
  zero(A1);
  zero(A2);
  zero(A3);

  // the next plane is loaded one phase ahead into NX, so each wave's loads
  // overlap the arithmetic of the plane before
  T NX[NP];
  auto fetch = [&]() {
    const T* src = in + (int64_t)zl * XY;
#pragma unroll
    for (int r = 0; r < NP; ++r) NX[r] = (src + rowoff[r])[col];
    zl = zl + 1 == Z ? 0 : zl + 1;
  };
  fetch();

  // stage s (1-based) reads ring `a` (its input) and, for an update stage of
  // a radius-two step, ring `b` (psi of the step, two stages back)
  auto stage = [&](auto sc, auto PHc, auto& a, auto& b, auto& o, int i) {
    constexpr int S = decltype(sc)::value;
    constexpr int PH = decltype(PHc)::value;
    constexpr int S0 = PH % 3, S1 = (PH + 1) % 3, S2 = (PH + 2) % 3;
    constexpr int NO = NP - 2 * S;  // rows produced
    constexpr bool MU = R == 2 && (S % 2 == 1);
    constexpr bool LAST = S == NS;
    if constexpr (LAST) {
      if (i < 2 * NS) return;
    } else {
      // input planes are valid from i = 2*(S-1) on; earlier ones would only
      // fill the ring with garbage that no valid output reads
      if (i < 2 * (S - 1)) return;
    }
    T res[NO];
#pragma unroll
    for (int r = 0; r < NO; ++r) {
      const T sn = n6<S0, S1, S2>(a, r);
      const T x = a[S1][r + 1];
      T v;
      if constexpr (MU) {
        const T t = horner<POLY - 1>(c.p, x);
        v = fmad(t, x, POLY == 3 ? sn : c.ks * sn);
      } else if constexpr (R == 2) {
        T base = b[S0][r + 2];
        if constexpr (HASA) base = base + horner<POLY>(c.q, base);
        v = fmad(c.k2, fmad(T(-6), x, sn), base);
      } else {
        T base = x;
        if constexpr (HASA) base = base + horner<POLY>(c.q, x);
        v = fmad(c.k1, fmad(T(-6), x, sn), base);
      }
      if constexpr (LAST)
        res[r] = v;
      else
        o[S2][r] = v;
    }
    if constexpr (LAST) {
      // every row computed by all lanes first, then one masked region of
      // stores; the row checks are wave-uniform
      if (store_lane) {
        T* dst = out + (int64_t)(z0 + i - 2 * NS) * XY + col;
        if (y0 + NO <= Y) {
#pragma unroll
          for (int r = 0; r < NO; ++r) dst[rowoff[r + RT]] = res[r];
        } else {
#pragma unroll
          for (int r = 0; r < NO; ++r)
            if (y0 + r < Y) dst[rowoff[r + RT]] = res[r];
        }
      }
    }
  };

  auto phase = [&](auto PHc, int i) {
    constexpr int PH = decltype(PHc)::value;
    constexpr int S2 = (PH + 2) % 3;
#pragma unroll
    for (int r = 0; r < NP; ++r) A0[S2][r] = NX[r];
    if (i + 1 < NI) fetch();
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using I4 = std::integral_constant<int, 4>;
    // ring arguments that a stage does not read are passed as A0
    stage(I1{}, PHc, A0, A0, A1, i);
    if constexpr (NS >= 2) stage(I2{}, PHc, A1, A0, A2, i);
    if constexpr (NS >= 3) stage(I3{}, PHc, A2, A1, A3, i);
    if constexpr (NS >= 4) stage(I4{}, PHc, A3, A2, A0, i);
  };
  for (int i = 0; i < NI; i += 3) {
    phase(std::integral_constant<int, 0>{}, i);
    if (i + 1 < NI) phase(std::integral_constant<int, 1>{}, i + 1);
    if (i + 2 < NI) phase(std::integral_constant<int, 2>{}, i + 2);
  }
}

template <typename T, int TYW, int R, int K, bool HASA, int POLY>
int launch3s(const T* in, T* out, int X, int Y, int Z, int TZ,
             const FusedCoef<T>& c, hipStream_t s) {
  constexpr int W = 64 - 2 * R * K;
  const int nstrips = (X + W - 1) / W;
  const int nrowblk = (Y + TYW - 1) / TYW;
  const int nzc = (Z + TZ - 1) / TZ;
  const int64_t nitems = (int64_t)nstrips * nrowblk * nzc;
  if (nitems > ((int64_t)1 << 31) - 4) {
    set_error("fused 3D step: too many work items; raise planes per chunk");
    return SPS_E_INVALID;
  }
  fused3d_stream<T, TYW, R, K, HASA, POLY><<<(unsigned)((nitems + 3) / 4), 256, 0, s>>>(
      in, out, X, Y, Z, TZ, nstrips, nrowblk, (int)nitems, c);
  return SPS_OK;
}

template <typename T, int R, bool HASA, int POLY>
int launch_variant(int K, int rows, const T* in, T* out, int X, int Y, int Z,
                   int TZ, const FusedCoef<T>& c, hipStream_t s) {
  if (K == 1) {
    switch (rows) {
      case 4: return launch3s<T, 4, R, 1, HASA, POLY>(in, out, X, Y, Z, TZ, c, s);
      // 6 rows also serve the one-step remainder launch of a (2, 6) shape
      case 6: return launch3s<T, 6, R, 1, HASA, POLY>(in, out, X, Y, Z, TZ, c, s);
      case 8: return launch3s<T, 8, R, 1, HASA, POLY>(in, out, X, Y, Z, TZ, c, s);
      case 12: return launch3s<T, 12, R, 1, HASA, POLY>(in, out, X, Y, Z, TZ, c, s);
      case 16: return launch3s<T, 16, R, 1, HASA, POLY>(in, out, X, Y, Z, TZ, c, s);
      default: break;
    }
  } else if (K == 2) {
    if (rows == 4) return launch3s<T, 4, R, 2, HASA, POLY>(in, out, X, Y, Z, TZ, c, s);
    if constexpr (sizeof(T) == 4) {
      // 6 rows: the rings of four radius-one stages fit three waves per SIMD
      if (rows == 6) return launch3s<T, 6, R, 2, HASA, POLY>(in, out, X, Y, Z, TZ, c, s);
      if (rows == 8) return launch3s<T, 8, R, 2, HASA, POLY>(in, out, X, Y, Z, TZ, c, s);
      if (rows == 12) return launch3s<T, 12, R, 2, HASA, POLY>(in, out, X, Y, Z, TZ, c, s);
    }
  }
  set_error("fused 3D step: unsupported launch shape (%d steps, %d rows)", K, rows);
  return SPS_E_INVALID;
}

template <typename T, int R>
int dispatch3(int K, int rows, const T* in, T* out, int X, int Y, int Z, int TZ,
              const FusedCoef<T>& c, hipStream_t s) {
  const bool hasa = c.dega >= 0;
  if (c.poly == 3)
    return hasa ? launch_variant<T, R, true, 3>(K, rows, in, out, X, Y, Z, TZ, c, s)
                : launch_variant<T, R, false, 3>(K, rows, in, out, X, Y, Z, TZ, c, s);
  return hasa ? launch_variant<T, R, true, 6>(K, rows, in, out, X, Y, Z, TZ, c, s)
              : launch_variant<T, R, false, 6>(K, rows, in, out, X, Y, Z, TZ, c, s);
}

}  // namespace

bool fused3d_shape_ok(int dtype_size, int K, int rows) {
  if (K == 1)
    return rows == 4 || rows == 6 || rows == 8 || rows == 12 || rows == 16;
  if (K == 2)
    return rows == 4 || ((rows == 6 || rows == 8 || rows == 12) && dtype_size == 4);
  return false;
}

template <typename T>
int launch_fused3d(const T* in, T* out, int X, int Y, int Z, int K, int rows,
                   int TZ, const FusedCoef<T>& c, hipStream_t s) {
  if (TZ < 1) {
    set_error("fused 3D step: planes per chunk must be >= 1");
    return SPS_E_INVALID;
  }
  int rc = c.radius == 1
               ? dispatch3<T, 1>(K, rows, in, out, X, Y, Z, TZ, c, s)
               : dispatch3<T, 2>(K, rows, in, out, X, Y, Z, TZ, c, s);
  if (rc != SPS_OK) return rc;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("fused 3D launch failed: %s", hipGetErrorString(e));
    return SPS_E_HIP;
  }
  return SPS_OK;
}

template int launch_fused3d<float>(const float*, float*, int, int, int, int,
                                   int, int, const FusedCoef<float>&, hipStream_t);
template int launch_fused3d<double>(const double*, double*, int, int, int, int,
                                    int, int, const FusedCoef<double>&,
                                    hipStream_t);

}  // namespace sps
