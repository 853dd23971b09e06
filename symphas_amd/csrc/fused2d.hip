// fused2d.hip - single-field fused 2D step (K steps per launch).
#include "fused.hpp"
#include "fused_tile.hpp"

namespace sps {
namespace {

// Halo layout: VEC values per lane (16 or 8 bytes), HL halo lanes per strip
// side.  Halving VEC doubles the strips, so the same number of waves covers
// chunks twice as tall, whose pipeline fill (2*K*R rows) costs less.
// Waves per workgroup.  One-wave workgroups let the dispatcher spread the
// waves evenly over the SIMDs when there are only one or two per SIMD.
constexpr int kHaloWaves = 1;

template <typename T, int VEC, int K, int R, bool HASA, int POLY>
__global__ __launch_bounds__(64 * kHaloWaves) void fused2d_halo(const T* __restrict__ in,
                                                    T* __restrict__ out, int X,
                                                    int Y, int TY, int nstrips,
                                                    int nitems,
                                                    FusedCoef<T> c) {
  constexpr int HL = (K * R + VEC - 1) / VEC;
  constexpr int W = (64 - 2 * HL) * VEC;
  const int lane = threadIdx.x & 63;
  // the work item is wave-uniform: keep it (and the chunk bounds and row
  // loop derived from it) in scalar registers
  const int item = __builtin_amdgcn_readfirstlane(
      xcd_remap(blockIdx.x, gridDim.x) * kHaloWaves + (threadIdx.x >> 6));
  if (item >= nitems) return;
  const int strip = item % nstrips;
  const int chunk = item / nstrips;
  const int cb = strip * W + (lane - HL) * VEC;
  int col = cb % X;
  if (col < 0) col += X;
  const bool store_lane = lane >= HL && lane < 64 - HL && cb < X;
  const int y0 = chunk * TY;
  const int ty = min(TY, Y - y0);
  stream_tile<T, VEC, K, R, HASA, POLY>(in, out, X, Y, y0, ty, col, store_lane,
                                        c);
}

template <typename T, int VEC, int K, int R, bool HASA, int POLY>
int launch_halo(const T* in, T* out, int X, int Y, int TY,
                const FusedCoef<T>& c, hipStream_t s) {
  constexpr int HL = (K * R + VEC - 1) / VEC;
  constexpr int W = (64 - 2 * HL) * VEC;
  const int nstrips = (X + W - 1) / W;
  const int nchunks = (Y + TY - 1) / TY;
  const int nitems = nstrips * nchunks;
  const int nblocks = (nitems + kHaloWaves - 1) / kHaloWaves;
  fused2d_halo<T, VEC, K, R, HASA, POLY>
      <<<nblocks, 64 * kHaloWaves, 0, s>>>(in, out, X, Y, TY, nstrips, nitems, c);
  return SPS_OK;
}

template <typename T, int K, int R, bool HASA, int POLY>
int dispatch_layout(const T* in, T* out, int X, int Y, int TY, int vec,
                    const FusedCoef<T>& c, hipStream_t s) {
  constexpr int WIDE = 16 / sizeof(T), NARROW = 8 / sizeof(T);
  return vec == NARROW
             ? launch_halo<T, NARROW, K, R, HASA, POLY>(in, out, X, Y, TY, c, s)
             : launch_halo<T, WIDE, K, R, HASA, POLY>(in, out, X, Y, TY, c, s);
}

template <typename T, int K, int R>
int dispatch_poly(const T* in, T* out, int X, int Y, int TY, int vec,
                  const FusedCoef<T>& c, hipStream_t s) {
  const bool hasa = c.dega >= 0;
  if (c.poly == 3) {
    return hasa ? dispatch_layout<T, K, R, true, 3>(in, out, X, Y, TY, vec, c, s)
                : dispatch_layout<T, K, R, false, 3>(in, out, X, Y, TY, vec, c, s);
  }
  return hasa ? dispatch_layout<T, K, R, true, 6>(in, out, X, Y, TY, vec, c, s)
              : dispatch_layout<T, K, R, false, 6>(in, out, X, Y, TY, vec, c, s);
}

template <typename T, int K>
int dispatch_radius(const T* in, T* out, int X, int Y, int TY, int vec,
                    const FusedCoef<T>& c, hipStream_t s) {
  return c.radius == 1 ? dispatch_poly<T, K, 1>(in, out, X, Y, TY, vec, c, s)
                       : dispatch_poly<T, K, 2>(in, out, X, Y, TY, vec, c, s);
}

}  // namespace

template <typename T>
int launch_fused2d(const T* in, T* out, int X, int Y, int K, int TY, int vec,
                   const FusedCoef<T>& c, hipStream_t s) {
  const int wide = 16 / (int)sizeof(T), narrow = 8 / (int)sizeof(T);
  if ((vec != wide && vec != narrow) || X % vec != 0 || TY < 1) {
    set_error("fused 2D step: %d values per lane (%d or %d) must divide X=%d",
              vec, wide, narrow, X);
    return SPS_E_UNSUPPORTED;
  }
  int rc;
  switch (K) {
    case 1: rc = dispatch_radius<T, 1>(in, out, X, Y, TY, vec, c, s); break;
    case 2: rc = dispatch_radius<T, 2>(in, out, X, Y, TY, vec, c, s); break;
    case 3: rc = dispatch_radius<T, 3>(in, out, X, Y, TY, vec, c, s); break;
    case 4: rc = dispatch_radius<T, 4>(in, out, X, Y, TY, vec, c, s); break;
    default:
      set_error("fused 2D step: steps per launch %d not in 1..4", K);
      return SPS_E_INVALID;
  }
  if (rc != SPS_OK) return rc;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("fused 2D launch failed: %s", hipGetErrorString(e));
    return SPS_E_HIP;
  }
  return SPS_OK;
}

template int launch_fused2d<float>(const float*, float*, int, int, int, int,
                                   int, const FusedCoef<float>&, hipStream_t);
template int launch_fused2d<double>(const double*, double*, int, int, int, int,
                                    int, const FusedCoef<double>&, hipStream_t);

}  // namespace sps
