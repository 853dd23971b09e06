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

template <typename T, int TY, int R>
int dispatch3(const T* in, T* out, int X, int Y, int Z, int TZ,
              const FusedCoef<T>& c, hipStream_t s) {
  const bool hasa = c.dega >= 0;
  const int deg = c.dega > c.degb ? c.dega : c.degb;
  if (deg <= 3)
    return hasa ? launch3<T, TY, R, true, 3>(in, out, X, Y, Z, TZ, c, s)
                : launch3<T, TY, R, false, 3>(in, out, X, Y, Z, TZ, c, s);
  return hasa ? launch3<T, TY, R, true, 6>(in, out, X, Y, Z, TZ, c, s)
              : launch3<T, TY, R, false, 6>(in, out, X, Y, Z, TZ, c, s);
}

}  // namespace

template <typename T>
int launch_fused3d(const T* in, T* out, int X, int Y, int Z, int TZ,
                   const FusedCoef<T>& c, hipStream_t s) {
  if (TZ < 1) {
    set_error("fused 3D step: planes per chunk must be >= 1");
    return SPS_E_INVALID;
  }
  constexpr int TY = sizeof(T) == 4 ? 16 : 8;
  int rc = c.radius == 1 ? dispatch3<T, TY, 1>(in, out, X, Y, Z, TZ, c, s)
                         : dispatch3<T, TY, 2>(in, out, X, Y, Z, TZ, c, s);
  if (rc != SPS_OK) return rc;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("fused 3D launch failed: %s", hipGetErrorString(e));
    return SPS_E_HIP;
  }
  return SPS_OK;
}

template int launch_fused3d<float>(const float*, float*, int, int, int, int,
                                   const FusedCoef<float>&, hipStream_t);
template int launch_fused3d<double>(const double*, double*, int, int, int, int,
                                    const FusedCoef<double>&, hipStream_t);

}  // namespace sps
