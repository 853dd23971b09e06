// fused3d.hip - fused 3D step: 2.5-D streaming along z.
//
// A workgroup owns a 64 x TY tile of the (x, y) plane and streams the planes
// of one z-chunk.  Planes of psi (tile + 2-cell halo) and of mu (tile +
// 1-cell halo) live in rings of four LDS slots, so each plane needs two
// barriers: psi(i) loaded -> mu(i-1) computed -> psi'(i-2) written.  With the
// 7-point laplacian, lap7∘lap7 is the 25-point bilaplacian of Stencil3d2h
// <7,*,25> exactly, so the canonical form
//   mu = B(psi) + kmu*D7(psi),  psi' = psi + Adt(psi) + kd*D7(mu)
// is the reference's -bilap psi - lap psi + lap psi^3 up to rounding.
#include "fused.hpp"
#include "lanes.hpp"

#include <type_traits>

namespace sps {
namespace {

constexpr int kTX = 64;

template <typename T, int TY, int R, bool HASA, int POLY>
__global__ __launch_bounds__(256) void fused3d_kernel(
    const T* __restrict__ in, T* __restrict__ out, int X, int Y, int Z, int TZ,
    int nbx, int nby, FusedCoef<T> c) {
  constexpr int PX = kTX + 4, PY = TY + 4;
  constexpr int MX = kTX + 2, MY = TY + 2;
  constexpr int NLOAD = (PX * PY + 255) / 256;
  constexpr int NMU = (MX * MY + 255) / 256;
  __shared__ T sp[4][PY][PX];
  __shared__ T sm[R == 2 ? 4 : 1][R == 2 ? MY : 1][R == 2 ? MX : 1];

  const int tid = threadIdx.x;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int bx = b % nbx;
  const int by = (b / nbx) % nby;
  const int bz = b / (nbx * nby);
  const int x0 = bx * kTX, y0 = by * TY, z0 = bz * TZ;
  const int tz = min(TZ, Z - z0);
  const int64_t XY = (int64_t)X * Y;

  // in-plane offsets of the loads this thread does for every plane
  int64_t loff[NLOAD];
#pragma unroll
  for (int q = 0; q < NLOAD; ++q) {
    const int e = tid + 256 * q;
    const int ly = e / PX, lx = e % PX;
    int gx = (x0 + lx - 2) % X;
    if (gx < 0) gx += X;
    int gy = (y0 + ly - 2) % Y;
    if (gy < 0) gy += Y;
    loff[q] = (int64_t)gy * X + gx;
  }

  const int NI = tz + 2 * R;
  int zl = (z0 - R) % Z;
  if (zl < 0) zl += Z;
  for (int i = 0; i < NI; ++i) {
    // A: psi plane of stream index i
    {
      T* dst = &sp[i & 3][0][0];
      const T* src = in + (int64_t)zl * XY;
#pragma unroll
      for (int q = 0; q < NLOAD; ++q) {
        const int e = tid + 256 * q;
        if (e < PX * PY) dst[e] = src[loff[q]];
      }
      zl = zl + 1 == Z ? 0 : zl + 1;
    }
    __syncthreads();
    if constexpr (R == 2) {
      // B: mu of plane i-1 on the tile + 1 halo
      if (i >= 2) {
        const int sc = (i - 1) & 3, sd = (i - 2) & 3, su = i & 3;
#pragma unroll
        for (int q = 0; q < NMU; ++q) {
          const int e = tid + 256 * q;
          if (e < MX * MY) {
            const int my = e / MX, mx = e % MX;
            const int py = my + 1, px = mx + 1;
            const T ctr = sp[sc][py][px];
            const T d = fmad(T(-6), ctr,
                             ((sp[sc][py][px - 1] + sp[sc][py][px + 1]) +
                              (sp[sc][py - 1][px] + sp[sc][py + 1][px])) +
                                 (sp[sd][py][px] + sp[su][py][px]));
            sm[sc][my][mx] = fmad(c.kmu, d, horner<POLY>(c.b, ctr));
          }
        }
      }
      __syncthreads();
      // C: psi' of plane i-2
      if (i >= 4) {
        const int zo = z0 + i - 4;
        const int mc = (i - 2) & 3, md = (i - 3) & 3, mu = (i - 1) & 3;
        const int tx = tid & 63;
#pragma unroll
        for (int m = 0; m < TY / 4; ++m) {
          const int ty = (tid >> 6) + 4 * m;
          const int my = ty + 1, mx = tx + 1;
          const T d = fmad(T(-6), sm[mc][my][mx],
                           ((sm[mc][my][mx - 1] + sm[mc][my][mx + 1]) +
                            (sm[mc][my - 1][mx] + sm[mc][my + 1][mx])) +
                               (sm[md][my][mx] + sm[mu][my][mx]));
          T base = sp[mc][ty + 2][tx + 2];
          if constexpr (HASA) base = base + horner<POLY>(c.adt, base);
          const T v = fmad(c.kd, d, base);
          if (x0 + tx < X && y0 + ty < Y)
            out[(int64_t)zo * XY + (int64_t)(y0 + ty) * X + (x0 + tx)] = v;
        }
      }
    } else {
      // C (radius one): psi' of plane i-1 straight from psi
      if (i >= 2) {
        const int zo = z0 + i - 2;
        const int pc = (i - 1) & 3, pd = (i - 2) & 3, pu = i & 3;
        const int tx = tid & 63;
#pragma unroll
        for (int m = 0; m < TY / 4; ++m) {
          const int ty = (tid >> 6) + 4 * m;
          const int py = ty + 2, px = tx + 2;
          const T ctr = sp[pc][py][px];
          const T d = fmad(T(-6), ctr,
                           ((sp[pc][py][px - 1] + sp[pc][py][px + 1]) +
                            (sp[pc][py - 1][px] + sp[pc][py + 1][px])) +
                               (sp[pd][py][px] + sp[pu][py][px]));
          T base = ctr;
          if constexpr (HASA) base = base + horner<POLY>(c.adt, base);
          const T v = fmad(c.k1, d, base);
          if (x0 + tx < X && y0 + ty < Y)
            out[(int64_t)zo * XY + (int64_t)(y0 + ty) * X + (x0 + tx)] = v;
        }
      }
    }
  }
}

// v2: one wave owns a 64-lane x TYW-row column block and streams the planes
// of one z-chunk through register windows; x neighbours come from adjacent
// lanes (DPP), y and z neighbours are registers.  No LDS, no barriers.
template <typename T, int TYW, int R, bool HASA, int POLY>
__global__ __launch_bounds__(256) void fused3d_stream(
    const T* __restrict__ in, T* __restrict__ out, int X, int Y, int Z, int TZ,
    int nstrips, int nrowblk, int nitems, FusedCoef<T> c) {
  constexpr int HL = R;  // halo lanes (one column per lane)
  constexpr int W = 64 - 2 * HL;
  constexpr int NP = TYW + 2 * R;  // psi rows held per plane
  constexpr int NM = TYW + 2;      // mu rows per plane (radius 2 only)
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
    int y = (y0 - R + r) % Y;
    if (y < 0) y += Y;
    rowoff[r] = __builtin_amdgcn_readfirstlane(y * X);
  }
  const int NI = tz + 2 * R;
  int zl = (z0 - R) % Z;
  if (zl < 0) zl += Z;

  T P[3][NP];
  T M[3][R == 2 ? NM : 1];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
#pragma unroll
    for (int r = 0; r < NP; ++r) P[s][r] = T(0);
#pragma unroll
    for (int r = 0; r < (R == 2 ? NM : 1); ++r) M[s][r] = T(0);
  }

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
  auto phase = [&](auto PHc, int i) {
    constexpr int PH = decltype(PHc)::value;
    constexpr int S0 = PH % 3, S1 = (PH + 1) % 3, S2 = (PH + 2) % 3;
    {
#pragma unroll
      for (int r = 0; r < NP; ++r) P[S2][r] = NX[r];
      if (i + 1 < NI) fetch();
    }
    if constexpr (R == 2) {
      // mu of plane i-1, rows y0-1 .. y0+TYW (psi rows 1 .. TYW+2)
#pragma unroll
      for (int r = 1; r <= TYW + 2; ++r) {
        const T ctr = P[S1][r];
        const T xl = from_prev_lane<false>(ctr), xr = from_next_lane<false>(ctr);
        const T d = fmad(T(-6), ctr,
                         ((xl + xr) + (P[S1][r - 1] + P[S1][r + 1])) +
                             (P[S0][r] + P[S2][r]));
        M[S2][r - 1] = fmad(c.kmu, d, horner<POLY>(c.b, ctr));
      }
      if (i >= 4) {
        const int zo = z0 + i - 4;
        T* dst = out + (int64_t)zo * XY;
#pragma unroll
        for (int m = 1; m <= TYW; ++m) {
          const T ctr = M[S1][m];
          const T xl = from_prev_lane<false>(ctr), xr = from_next_lane<false>(ctr);
          const T d = fmad(T(-6), ctr,
                           ((xl + xr) + (M[S1][m - 1] + M[S1][m + 1])) +
                               (M[S0][m] + M[S2][m]));
          T base = P[S0][m + 1];
          if constexpr (HASA) base = base + horner<POLY>(c.adt, base);
          const T v = fmad(c.kd, d, base);
          if (store_lane && y0 + m - 1 < Y) (dst + rowoff[m + 1])[col] = v;
        }
      }
    } else {
      if (i >= 2) {
        const int zo = z0 + i - 2;
        T* dst = out + (int64_t)zo * XY;
#pragma unroll
        for (int r = 1; r <= TYW; ++r) {
          const T ctr = P[S1][r];
          const T xl = from_prev_lane<false>(ctr), xr = from_next_lane<false>(ctr);
          const T d = fmad(T(-6), ctr,
                           ((xl + xr) + (P[S1][r - 1] + P[S1][r + 1])) +
                               (P[S0][r] + P[S2][r]));
          T base = ctr;
          if constexpr (HASA) base = base + horner<POLY>(c.adt, base);
          const T v = fmad(c.k1, d, base);
          if (store_lane && y0 + r - 1 < Y) (dst + rowoff[r])[col] = v;
        }
      }
    }
  };
  for (int i = 0; i < NI; i += 3) {
    phase(std::integral_constant<int, 0>{}, i);
    if (i + 1 < NI) phase(std::integral_constant<int, 1>{}, i + 1);
    if (i + 2 < NI) phase(std::integral_constant<int, 2>{}, i + 2);
  }
}

template <typename T, int TYW, int R, bool HASA, int POLY>
int launch3s(const T* in, T* out, int X, int Y, int Z, int TZ,
             const FusedCoef<T>& c, hipStream_t s) {
  constexpr int W = 64 - 2 * R;
  const int nstrips = (X + W - 1) / W;
  const int nrowblk = (Y + TYW - 1) / TYW;
  const int nzc = (Z + TZ - 1) / TZ;
  const int nitems = nstrips * nrowblk * nzc;
  fused3d_stream<T, TYW, R, HASA, POLY><<<(nitems + 3) / 4, 256, 0, s>>>(
      in, out, X, Y, Z, TZ, nstrips, nrowblk, nitems, c);
  return SPS_OK;
}

template <typename T, int TY, int R, bool HASA, int POLY>
int launch3(const T* in, T* out, int X, int Y, int Z, int TZ,
            const FusedCoef<T>& c, hipStream_t s) {
  const int nbx = (X + kTX - 1) / kTX;
  const int nby = (Y + TY - 1) / TY;
  const int nbz = (Z + TZ - 1) / TZ;
  fused3d_kernel<T, TY, R, HASA, POLY>
      <<<nbx * nby * nbz, 256, 0, s>>>(in, out, X, Y, Z, TZ, nbx, nby, c);
  return SPS_OK;
}

template <typename T, int TY, int R, bool HASA, int POLY>
int launch_variant(int variant, const T* in, T* out, int X, int Y, int Z,
                   int TZ, const FusedCoef<T>& c, hipStream_t s) {
  switch (variant) {
    case 8: return launch3s<T, 8, R, HASA, POLY>(in, out, X, Y, Z, TZ, c, s);
    case 16: return launch3s<T, 16, R, HASA, POLY>(in, out, X, Y, Z, TZ, c, s);
    default: return launch3<T, TY, R, HASA, POLY>(in, out, X, Y, Z, TZ, c, s);
  }
}

template <typename T, int TY, int R>
int dispatch3(int variant, const T* in, T* out, int X, int Y, int Z, int TZ,
              const FusedCoef<T>& c, hipStream_t s) {
  const bool hasa = c.dega >= 0;
  const int deg = c.dega > c.degb ? c.dega : c.degb;
  if (deg <= 3)
    return hasa ? launch_variant<T, TY, R, true, 3>(variant, in, out, X, Y, Z, TZ, c, s)
                : launch_variant<T, TY, R, false, 3>(variant, in, out, X, Y, Z, TZ, c, s);
  return hasa ? launch_variant<T, TY, R, true, 6>(variant, in, out, X, Y, Z, TZ, c, s)
              : launch_variant<T, TY, R, false, 6>(variant, in, out, X, Y, Z, TZ, c, s);
}

}  // namespace

template <typename T>
int launch_fused3d(const T* in, T* out, int X, int Y, int Z, int rows, int TZ,
                   const FusedCoef<T>& c, hipStream_t s) {
  if (TZ < 1) {
    set_error("fused 3D step: planes per chunk must be >= 1");
    return SPS_E_INVALID;
  }
  constexpr int TY = sizeof(T) == 4 ? 16 : 8;
  int rc = c.radius == 1
               ? dispatch3<T, TY, 1>(rows, in, out, X, Y, Z, TZ, c, s)
               : dispatch3<T, TY, 2>(rows, in, out, X, Y, Z, TZ, c, s);
  if (rc != SPS_OK) return rc;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("fused 3D launch failed: %s", hipGetErrorString(e));
    return SPS_E_HIP;
  }
  return SPS_OK;
}

template int launch_fused3d<float>(const float*, float*, int, int, int, int,
                                   int, const FusedCoef<float>&, hipStream_t);
template int launch_fused3d<double>(const double*, double*, int, int, int, int,
                                    int, const FusedCoef<double>&, hipStream_t);

}  // namespace sps
