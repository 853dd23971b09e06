// nn.hip - BF16 MFMA convolution kernels of the learned surrogate
// (include/sps_nn.h).
//
// Forward (and backward-data, which is the forward kernel with flipped,
// transposed weights) is an implicit GEMM: M = output channels, N = pixels,
// K = 9 taps x input channels, on v_mfma_f32_16x16x32_bf16.  Operand layout
// of that instruction on gfx950 (wave64):
//   A (16 x 32, M x K): lane l holds A[l % 16][8*(l/16) + 0..7]
//   B (32 x 16, K x N): lane l holds B[8*(l/16) + 0..7][l % 16]
//   D (16 x 16):        lane l holds D[4*(l/16) + 0..3][l % 16]
// With NHWC activations and [tap][out][in] weights both A (weights) and B
// (pixels) fragments are single 16-byte loads, so the kernel streams straight
// from L1/L2; each wave owns 16*NT output channels x 64 pixels, reuses every
// weight fragment 4 times and every pixel fragment NT times, and stores 4
// consecutive output channels of a pixel per lane (8 bytes).
//
// The weight gradient reduces over pixels, which are strided in NHWC, so it
// stages a 32-pixel row chunk of dpre and the three shifted input rows in LDS
// and gathers the fragments from there.
#include <cstdint>

#include "common.hpp"
#include "../../include/sps_nn.h"

namespace sps {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4v __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// IEEE round-to-nearest-even FP32 -> BF16 (quiet NaN preserved).
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}

__device__ __forceinline__ float silu(float v) { return v / (1.0f + __expf(-v)); }

__device__ __forceinline__ bf16x8 load_frag(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(p));
}

// 4 consecutive BF16 values as one 8-byte access (callers keep 8-byte
// alignment: channel offsets are multiples of 4)
__device__ __forceinline__ void store4(uint16_t* p, const float (&v)[4]) {
  const uint32_t lo = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
  const uint32_t hi = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
  *reinterpret_cast<uint2*>(p) = make_uint2(lo, hi);
}
__device__ __forceinline__ void load4(const uint16_t* p, float (&v)[4]) {
  const uint2 q = *reinterpret_cast<const uint2*>(p);
  v[0] = bf2f((uint16_t)(q.x & 0xffffu));
  v[1] = bf2f((uint16_t)(q.x >> 16));
  v[2] = bf2f((uint16_t)(q.y & 0xffffu));
  v[3] = bf2f((uint16_t)(q.y >> 16));
}
// 8 consecutive BF16 values as one 16-byte access (16-byte aligned)
__device__ __forceinline__ void store8(uint16_t* p, const float (&v)[8]) {
  uint32_t w[4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
    w[i] = (uint32_t)f2bf(v[2 * i]) | ((uint32_t)f2bf(v[2 * i + 1]) << 16);
  *reinterpret_cast<uint4*>(p) = make_uint4(w[0], w[1], w[2], w[3]);
}
__device__ __forceinline__ void load8(const uint16_t* p, float (&v)[8]) {
  const uint4 q = *reinterpret_cast<const uint4*>(p);
  const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    v[2 * i] = bf2f((uint16_t)(w[i] & 0xffffu));
    v[2 * i + 1] = bf2f((uint16_t)(w[i] >> 16));
  }
}

__device__ __forceinline__ int wrapi(int v, int n) {
  return v < 0 ? v + n : (v >= n ? v - n : v);
}

template <int NT, int EPI>
__global__ __launch_bounds__(256) void conv_fwd(
    const uint16_t* __restrict__ x, int N, int H, int W, int C,
    const uint16_t* __restrict__ wt, const float* __restrict__ bias, int K,
    int act, const uint16_t* __restrict__ skip, uint16_t* __restrict__ y,
    uint16_t* __restrict__ pre, int npix) {
  // D = W (out channels x in channels) * X (in channels x pixels): rows of D
  // are output channels, so lane l ends with 4 consecutive channels of one
  // pixel and stores them as one 8-byte word.
  const int lane = threadIdx.x & 63;
  const int p0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
  if (p0 >= npix) return;
  const int k0 = blockIdx.y * 16 * NT;
  const int HW = H * W;
  const int cq = (lane >> 4) * 8;
  int pn[4], ph[4], pw[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int p = p0 + 16 * m + (lane & 15);
    pn[m] = p / HW;
    const int r = p - pn[m] * HW;
    ph[m] = r / W;
    pw[m] = r - ph[m] * W;
  }
  f32x4v acc[NT][4];
#pragma unroll
  for (int j = 0; j < NT; ++j)
#pragma unroll
    for (int m = 0; m < 4; ++m) acc[j][m] = f32x4v{0.f, 0.f, 0.f, 0.f};

  for (int tap = 0; tap < 9; ++tap) {
    const int dr = tap / 3 - 1, ds = tap % 3 - 1;
    const uint16_t* xrow[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int hh = wrapi(ph[m] + dr, H), ww = wrapi(pw[m] + ds, W);
      xrow[m] = x + ((int64_t)(pn[m] * H + hh) * W + ww) * C + cq;
    }
    const uint16_t* wrow = wt + ((int64_t)tap * K + k0 + (lane & 15)) * C + cq;
#pragma unroll 2
    for (int c0 = 0; c0 < C; c0 += 32) {
      bf16x8 a[NT], b[4];
#pragma unroll
      for (int j = 0; j < NT; ++j) a[j] = load_frag(wrow + (int64_t)j * 16 * C + c0);
#pragma unroll
      for (int m = 0; m < 4; ++m) b[m] = load_frag(xrow[m] + c0);
#pragma unroll
      for (int j = 0; j < NT; ++j)
#pragma unroll
        for (int m = 0; m < 4; ++m)
          acc[j][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[j], b[m],
                                                              acc[j][m], 0, 0, 0);
    }
  }

#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int ch = k0 + 16 * j + 4 * (lane >> 4);  // 4 consecutive channels
    float bv[4] = {0.f, 0.f, 0.f, 0.f};
    if (bias) {
#pragma unroll
      for (int i = 0; i < 4; ++i) bv[i] = bias[ch + i];
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int p = p0 + 16 * m + (lane & 15);
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = acc[j][m][i] + bv[i];
      if constexpr (EPI == SPS_NN_EPI_PLAIN) {
        const int64_t o = (int64_t)p * K + ch;
        if (pre) store4(pre + o, v);
        if (act) {
#pragma unroll
          for (int i = 0; i < 4; ++i) v[i] = silu(v[i]);
        }
        store4(y + o, v);
      } else {
        const int n = pn[m], hh = ph[m], ww = pw[m];
#pragma unroll
        for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
          for (int b2 = 0; b2 < 2; ++b2) {
            const int64_t o =
                ((int64_t)(n * 2 * H + 2 * hh + a2) * (2 * W) + 2 * ww + b2) * K +
                ch;
            float u[4];
            load4(skip + o, u);
#pragma unroll
            for (int i = 0; i < 4; ++i) u[i] += v[i];
            if (pre) store4(pre + o, u);
            if (act) {
#pragma unroll
              for (int i = 0; i < 4; ++i) u[i] = silu(u[i]);
            }
            store4(y + o, u);
          }
      }
    }
  }
}

// Weight gradient: tile KT x CT of [k][c] for all 9 taps, reduction over
// 32-pixel row chunks staged in LDS.  Wave w owns k rows [w%2 * KT/2, +KT/2)
// and c columns [w/2 * CT/2, +CT/2).
// ROWS: chunks are 32-pixel pieces of one row (w % 32 == 0) and the three
// input rows with a 1-column halo are staged once; otherwise chunks are any
// 32 consecutive pixels and the nine shifted pixels are staged per tap.
template <int KT, int CT, bool ROWS>
__global__ __launch_bounds__(256) void conv_wgrad(
    const uint16_t* __restrict__ x, const uint16_t* __restrict__ dy, int N,
    int H, int W, int C, int K, float* __restrict__ dw, int nchunks,
    int chunks_per_block) {
  constexpr int TK = KT / 32, TC = CT / 32;  // MFMA tiles per wave
  __shared__ uint16_t sdy[32][KT + 2];
  __shared__ uint16_t sx[ROWS ? 3 : 9][ROWS ? 34 : 32][CT + 2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c0 = blockIdx.x * CT, k0 = blockIdx.y * KT;
  const int kw = (wave & 1) * (KT / 2), cw = (wave >> 1) * (CT / 2);
  const int rowchunks = W / 32;
  f32x4v acc[9][TK][TC];
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int a = 0; a < TK; ++a)
#pragma unroll
      for (int b = 0; b < TC; ++b) acc[t][a][b] = f32x4v{0.f, 0.f, 0.f, 0.f};

  const int first = blockIdx.z * chunks_per_block;
  const int last = min(nchunks, first + chunks_per_block);
  for (int ch = first; ch < last; ++ch) {
    __syncthreads();
    if constexpr (ROWS) {
      const int n = ch / (H * rowchunks);
      const int rem = ch - n * H * rowchunks;
      const int h = rem / rowchunks;
      const int w0 = (rem - h * rowchunks) * 32;
      // stage dy: 32 pixels x KT channels
      for (int e = threadIdx.x; e < 32 * KT; e += 256) {
        const int px = e / KT, k = e - px * KT;
        sdy[px][k] = dy[((int64_t)(n * H + h) * W + w0 + px) * K + k0 + k];
      }
      // stage x: rows h-1..h+1, columns w0-1..w0+32, CT channels
      for (int e = threadIdx.x; e < 3 * 34 * CT; e += 256) {
        const int r = e / (34 * CT);
        const int e2 = e - r * 34 * CT;
        const int col = e2 / CT, c = e2 - col * CT;
        const int hh = wrapi(h + r - 1, H), ww = wrapi(w0 + col - 1, W);
        sx[r][col][c] = x[((int64_t)(n * H + hh) * W + ww) * C + c0 + c];
      }
    } else {
      const int64_t pbase = (int64_t)ch * 32;
      for (int e = threadIdx.x; e < 32 * KT; e += 256) {
        const int px = e / KT, k = e - px * KT;
        sdy[px][k] = dy[(pbase + px) * K + k0 + k];
      }
      const int HW = H * W;
      for (int e = threadIdx.x; e < 9 * 32 * CT; e += 256) {
        const int t = e / (32 * CT);
        const int e2 = e - t * 32 * CT;
        const int px = e2 / CT, c = e2 - px * CT;
        const int64_t p = pbase + px;
        const int n = (int)(p / HW);
        const int r = (int)(p - (int64_t)n * HW);
        const int h = r / W, w = r - h * W;
        const int hh = wrapi(h + t / 3 - 1, H), ww = wrapi(w + t % 3 - 1, W);
        sx[t][px][c] = x[((int64_t)(n * H + hh) * W + ww) * C + c0 + c];
      }
    }
    __syncthreads();
    const int pq = (lane >> 4) * 8;
    bf16x8 a[TK];
#pragma unroll
    for (int tk = 0; tk < TK; ++tk) {
      uint16_t v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = sdy[pq + i][kw + 16 * tk + (lane & 15)];
      a[tk] = __builtin_bit_cast(bf16x8, v);
    }
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int r = tap / 3, s = tap % 3;
#pragma unroll
      for (int tc = 0; tc < TC; ++tc) {
        uint16_t v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          if constexpr (ROWS)
            v[i] = sx[r][pq + i + s][cw + 16 * tc + (lane & 15)];
          else
            v[i] = sx[tap][pq + i][cw + 16 * tc + (lane & 15)];
        }
        const bf16x8 b = __builtin_bit_cast(bf16x8, v);
#pragma unroll
        for (int tk = 0; tk < TK; ++tk)
          acc[tap][tk][tc] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              a[tk], b, acc[tap][tk][tc], 0, 0, 0);
      }
    }
  }
  if (first >= last) return;
#pragma unroll
  for (int tap = 0; tap < 9; ++tap)
#pragma unroll
    for (int tk = 0; tk < TK; ++tk)
#pragma unroll
      for (int tc = 0; tc < TC; ++tc)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int k = k0 + kw + 16 * tk + 4 * (lane >> 4) + i;
          const int c = c0 + cw + 16 * tc + (lane & 15);
          atomicAdd(&dw[((int64_t)tap * K + k) * C + c], acc[tap][tk][tc][i]);
        }
}

// ------------------------------------------------------------ thin layers
// The first (1 -> K) and last (K -> 1) convolutions have no GEMM shape; they
// run on the vector units with 16-byte accesses of 8 channels.  K % 8 == 0.

__global__ __launch_bounds__(256) void in1_fwd(const float* __restrict__ x,
                                               int N, int H, int W,
                                               const float* __restrict__ wt,
                                               const float* __restrict__ bias,
                                               int K, int act,
                                               uint16_t* __restrict__ y,
                                               uint16_t* __restrict__ pre) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= (int64_t)N * H * W) return;
  const int HW = H * W;
  const int n = p / HW, r = p - (int64_t)n * HW, h = r / W, w = r - h * W;
  float xv[9];
#pragma unroll
  for (int t = 0; t < 9; ++t)
    xv[t] = x[(int64_t)(n * H + wrapi(h + t / 3 - 1, H)) * W +
              wrapi(w + t % 3 - 1, W)];
  for (int k0 = 0; k0 < K; k0 += 8) {
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      float a = bias ? bias[k0 + i] : 0.f;
#pragma unroll
      for (int t = 0; t < 9; ++t) a = fmaf(xv[t], wt[t * K + k0 + i], a);
      v[i] = a;
    }
    if (pre) store8(pre + p * K + k0, v);
    if (act) {
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = silu(v[i]);
    }
    store8(y + p * K + k0, v);
  }
}

// dweight[t][k] += sum_p x[shift_t(p)] * dy[p][k]: 256-pixel chunks staged
// in LDS, then thread (t, k) pairs reduce the chunk.
__global__ __launch_bounds__(256) void in1_wgrad(const float* __restrict__ x,
                                                 const uint16_t* __restrict__ dy,
                                                 int N, int H, int W, int K,
                                                 float* __restrict__ dw,
                                                 int chunks_per_block) {
  __shared__ float sx[9][256];
  __shared__ float sd[256][65];
  const int64_t np = (int64_t)N * H * W;
  const int HW = H * W;
  float acc[3] = {0.f, 0.f, 0.f};  // up to 9*64 = 576 (t, k) pairs / 256
  for (int c = 0; c < chunks_per_block; ++c) {
    const int64_t base = ((int64_t)blockIdx.x * chunks_per_block + c) * 256;
    if (base >= np) break;
    __syncthreads();
    const int64_t p = base + threadIdx.x;
    if (p < np) {
      const int n = p / HW, r = p - (int64_t)n * HW, h = r / W, w = r - h * W;
#pragma unroll
      for (int t = 0; t < 9; ++t)
        sx[t][threadIdx.x] = x[(int64_t)(n * H + wrapi(h + t / 3 - 1, H)) * W +
                               wrapi(w + t % 3 - 1, W)];
      for (int k0 = 0; k0 < K; k0 += 8) {
        float v[8];
        load8(dy + p * K + k0, v);
#pragma unroll
        for (int i = 0; i < 8; ++i) sd[threadIdx.x][k0 + i] = v[i];
      }
    } else {
#pragma unroll
      for (int t = 0; t < 9; ++t) sx[t][threadIdx.x] = 0.f;
      for (int k = 0; k < K; ++k) sd[threadIdx.x][k] = 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int tk = threadIdx.x + 256 * q;
      if (tk < 9 * K) {
        const int t = tk / K, k = tk - t * K;
        float a = acc[q];
        for (int i = 0; i < 256; ++i) a = fmaf(sx[t][i], sd[i][k], a);
        acc[q] = a;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int tk = threadIdx.x + 256 * q;
    if (tk < 9 * K) atomicAdd(&dw[tk], acc[q]);
  }
}

__global__ __launch_bounds__(256) void out1_fwd(const uint16_t* __restrict__ x,
                                                int N, int H, int W, int C,
                                                const float* __restrict__ wt,
                                                const float* __restrict__ bias,
                                                float* __restrict__ y) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= (int64_t)N * H * W) return;
  const int HW = H * W;
  const int n = p / HW, r = p - (int64_t)n * HW, h = r / W, w = r - h * W;
  float v = bias ? bias[0] : 0.f;
  for (int t = 0; t < 9; ++t) {
    const uint16_t* xp =
        x + ((int64_t)(n * H + wrapi(h + t / 3 - 1, H)) * W +
             wrapi(w + t % 3 - 1, W)) * C;
    for (int c0 = 0; c0 < C; c0 += 8) {
      float xv[8];
      load8(xp + c0, xv);
#pragma unroll
      for (int i = 0; i < 8; ++i) v = fmaf(xv[i], wt[t * C + c0 + i], v);
    }
  }
  y[p] = v;
}

// dx[p][c] = sum_t dy[p - shift_t] * w[t][c]
__global__ __launch_bounds__(256) void out1_bwd_data(
    const float* __restrict__ dy, int N, int H, int W, int C,
    const float* __restrict__ wt, uint16_t* __restrict__ dx) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= (int64_t)N * H * W) return;
  const int HW = H * W;
  const int n = p / HW, r = p - (int64_t)n * HW, h = r / W, w = r - h * W;
  float d[9];
#pragma unroll
  for (int t = 0; t < 9; ++t)
    d[t] = dy[(int64_t)(n * H + wrapi(h - (t / 3 - 1), H)) * W +
              wrapi(w - (t % 3 - 1), W)];
  for (int c0 = 0; c0 < C; c0 += 8) {
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      float a = 0.f;
#pragma unroll
      for (int t = 0; t < 9; ++t) a = fmaf(d[t], wt[t * C + c0 + i], a);
      v[i] = a;
    }
    store8(dx + p * C + c0, v);
  }
}

// dweight[t][c] += sum_p x[shift_t(p)][c] * dy[p]: per 256-pixel chunk the
// 9 shifted rows of x are read by thread (t, c) pairs from global memory
// (L1/L2 hits) and dy from LDS.
__global__ __launch_bounds__(256) void out1_wgrad(const uint16_t* __restrict__ x,
                                                  const float* __restrict__ dy,
                                                  int N, int H, int W, int C,
                                                  float* __restrict__ dw,
                                                  int chunks_per_block) {
  __shared__ float sd[256];
  __shared__ int soff[9][256];
  const int64_t np = (int64_t)N * H * W;
  const int HW = H * W;
  float acc[3] = {0.f, 0.f, 0.f};
  for (int c = 0; c < chunks_per_block; ++c) {
    const int64_t base = ((int64_t)blockIdx.x * chunks_per_block + c) * 256;
    if (base >= np) break;
    __syncthreads();
    const int64_t p = base + threadIdx.x;
    if (p < np) {
      const int n = p / HW, r = p - (int64_t)n * HW, h = r / W, w = r - h * W;
      sd[threadIdx.x] = dy[p];
#pragma unroll
      for (int t = 0; t < 9; ++t)
        soff[t][threadIdx.x] =
            (int)((int64_t)(n * H + wrapi(h + t / 3 - 1, H)) * W +
                  wrapi(w + t % 3 - 1, W) - base);
    } else {
      sd[threadIdx.x] = 0.f;
#pragma unroll
      for (int t = 0; t < 9; ++t) soff[t][threadIdx.x] = 0;
    }
    __syncthreads();
    const int64_t lim = np - base;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int tc = threadIdx.x + 256 * q;
      if (tc < 9 * C) {
        const int t = tc / C, ch = tc - t * C;
        float a = acc[q];
        for (int i = 0; i < 256 && i < lim; ++i)
          a = fmaf(bf2f(x[(base + soff[t][i]) * C + ch]), sd[i], a);
        acc[q] = a;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int tc = threadIdx.x + 256 * q;
    if (tc < 9 * C) atomicAdd(&dw[tc], acc[q]);
  }
}

// ------------------------------------------------------------ elementwise

__global__ __launch_bounds__(256) void silu_bwd_k(const uint16_t* __restrict__ dy,
                                                  const uint16_t* __restrict__ pre,
                                                  uint16_t* __restrict__ dx,
                                                  int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * 256) {
    const float z = bf2f(pre[i]);
    const float s = 1.0f / (1.0f + __expf(-z));
    dx[i] = f2bf(bf2f(dy[i]) * s * (1.0f + z * (1.0f - s)));
  }
}

__global__ __launch_bounds__(256) void pool2_k(const uint16_t* __restrict__ x,
                                               int N, int H, int W, int C,
                                               int mode,
                                               uint16_t* __restrict__ y) {
  const int Ho = H / 2, Wo = W / 2;
  const int64_t total = (int64_t)N * Ho * Wo * C;
  const float scale = mode == 0 ? 0.25f : 1.0f;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * 256) {
    const int c = i % C;
    int64_t q = i / C;
    const int wo = q % Wo;
    q /= Wo;
    const int ho = q % Ho;
    const int n = q / Ho;
    const int64_t b = ((int64_t)(n * H + 2 * ho) * W + 2 * wo) * C + c;
    const float s = (bf2f(x[b]) + bf2f(x[b + C])) +
                    (bf2f(x[b + (int64_t)W * C]) + bf2f(x[b + (int64_t)W * C + C]));
    y[i] = f2bf(s * scale);
  }
}

__global__ __launch_bounds__(256) void unpool2_k(const uint16_t* __restrict__ x,
                                                 int N, int H, int W, int C,
                                                 float scale,
                                                 uint16_t* __restrict__ y) {
  const int Ho = 2 * H, Wo = 2 * W;
  const int64_t total = (int64_t)N * Ho * Wo * C;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * 256) {
    const int c = i % C;
    int64_t q = i / C;
    const int wo = q % Wo;
    q /= Wo;
    const int ho = q % Ho;
    const int n = q / Ho;
    y[i] = f2bf(bf2f(x[((int64_t)(n * H + ho / 2) * W + wo / 2) * C + c]) * scale);
  }
}

int grid_for(int64_t count) {
  int64_t g = (count + 255) / 256;
  return (int)(g > 65536 ? 65536 : (g < 1 ? 1 : g));
}

int launch_ok(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s launch failed: %s", what, hipGetErrorString(e));
    return SPS_E_HIP;
  }
  return SPS_OK;
}

}  // namespace
}  // namespace sps

using namespace sps;

extern "C" {

int sps_nn_conv3x3(const void* x, int n, int h, int w, int cin,
                   const void* weight, const float* bias, int cout, int act,
                   int epilogue, const void* skip, void* y, void* pre,
                   void* stream) {
  const int64_t npix = (int64_t)n * h * w;
  if (!x || !weight || !y || n < 1 || h < 1 || w < 1 || cin < 32 ||
      cin % 32 || cout < 32 || cout % 32 || npix % 64 ||
      npix > (int64_t)1 << 30 || (act != 0 && act != 1) ||
      (epilogue != SPS_NN_EPI_PLAIN && epilogue != SPS_NN_EPI_UPSAMPLE_ADD) ||
      (epilogue == SPS_NN_EPI_UPSAMPLE_ADD && !skip)) {
    set_error("sps_nn_conv3x3: need cin, cout multiples of 32, n*h*w a "
              "multiple of 64, act in {0,1}, a skip tensor for UPSAMPLE_ADD");
    return SPS_E_INVALID;
  }
  hipStream_t s = (hipStream_t)stream;
  const int nt = cout % 64 == 0 ? 4 : 2;
  dim3 grid((unsigned)((npix / 64 + 3) / 4), (unsigned)(cout / (16 * nt)));
  const uint16_t* xx = (const uint16_t*)x;
  const uint16_t* ww = (const uint16_t*)weight;
  const uint16_t* sk = (const uint16_t*)skip;
  uint16_t* yy = (uint16_t*)y;
  uint16_t* pp = (uint16_t*)pre;
  if (epilogue == SPS_NN_EPI_PLAIN) {
    if (nt == 4)
      conv_fwd<4, 0><<<grid, 256, 0, s>>>(xx, n, h, w, cin, ww, bias, cout, act, sk, yy, pp, (int)npix);
    else
      conv_fwd<2, 0><<<grid, 256, 0, s>>>(xx, n, h, w, cin, ww, bias, cout, act, sk, yy, pp, (int)npix);
  } else {
    if (nt == 4)
      conv_fwd<4, 1><<<grid, 256, 0, s>>>(xx, n, h, w, cin, ww, bias, cout, act, sk, yy, pp, (int)npix);
    else
      conv_fwd<2, 1><<<grid, 256, 0, s>>>(xx, n, h, w, cin, ww, bias, cout, act, sk, yy, pp, (int)npix);
  }
  return launch_ok("sps_nn_conv3x3");
}

int sps_nn_conv3x3_wgrad(const void* x, const void* dpre, int n, int h, int w,
                         int cin, int cout, float* dweight, void* stream) {
  const int64_t npix = (int64_t)n * h * w;
  if (!x || !dpre || !dweight || n < 1 || h < 1 || w < 1 || npix % 32 ||
      npix > (int64_t)1 << 30 || cin % 32 || cout % 32 || cin < 32 ||
      cout < 32) {
    set_error("sps_nn_conv3x3_wgrad: need n*h*w, cin, cout multiples of 32");
    return SPS_E_INVALID;
  }
  const bool rows = w % 32 == 0;
  const int nchunks = (int)(npix / 32);
  const int per = 32;
  const int kt = cout % 64 == 0 ? 64 : 32;
  const int ct = cin % 64 == 0 ? 64 : 32;
  dim3 grid(cin / ct, cout / kt, (nchunks + per - 1) / per);
  hipStream_t s = (hipStream_t)stream;
  const uint16_t* xx = (const uint16_t*)x;
  const uint16_t* dd = (const uint16_t*)dpre;
#define SPS_WGRAD(KT_, CT_)                                                   \
  (rows ? (conv_wgrad<KT_, CT_, true><<<grid, 256, 0, s>>>(                   \
               xx, dd, n, h, w, cin, cout, dweight, nchunks, per), 0)        \
        : (conv_wgrad<KT_, CT_, false><<<grid, 256, 0, s>>>(                  \
               xx, dd, n, h, w, cin, cout, dweight, nchunks, per), 0))
  if (kt == 64 && ct == 64)
    SPS_WGRAD(64, 64);
  else if (kt == 64)
    SPS_WGRAD(64, 32);
  else if (ct == 64)
    SPS_WGRAD(32, 64);
  else
    SPS_WGRAD(32, 32);
#undef SPS_WGRAD
  return launch_ok("sps_nn_conv3x3_wgrad");
}

int sps_nn_conv3x3_in1(const float* x, int n, int h, int w,
                       const float* weight, const float* bias, int cout,
                       int act, void* y, void* pre, void* stream) {
  const int64_t np = (int64_t)n * h * w;
  if (!x || !weight || !y || np < 1 || cout < 8 || cout > 64 || cout % 8 ||
      (act != 0 && act != 1)) {
    set_error("sps_nn_conv3x3_in1: cout a multiple of 8 in 8..64, act 0/1");
    return SPS_E_INVALID;
  }
  in1_fwd<<<(unsigned)((np + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      x, n, h, w, weight, bias, cout, act, (uint16_t*)y, (uint16_t*)pre);
  return launch_ok("sps_nn_conv3x3_in1");
}

int sps_nn_conv3x3_in1_wgrad(const float* x, const void* dpre, int n, int h,
                             int w, int cout, float* dweight, void* stream) {
  const int64_t np = (int64_t)n * h * w;
  if (!x || !dpre || !dweight || np < 1 || cout < 8 || cout > 64 || cout % 8) {
    set_error("sps_nn_conv3x3_in1_wgrad: cout a multiple of 8 in 8..64");
    return SPS_E_INVALID;
  }
  const int per = 16;  // 256-pixel chunks per block
  const int64_t chunks = (np + 255) / 256;
  in1_wgrad<<<(unsigned)((chunks + per - 1) / per), 256, 0,
              (hipStream_t)stream>>>(x, (const uint16_t*)dpre, n, h, w, cout,
                                     dweight, per);
  return launch_ok("sps_nn_conv3x3_in1_wgrad");
}

int sps_nn_conv3x3_out1(const void* x, int n, int h, int w, int cin,
                        const float* weight, const float* bias, float* y,
                        void* stream) {
  const int64_t np = (int64_t)n * h * w;
  if (!x || !weight || !y || np < 1 || cin < 8 || cin > 64 || cin % 8) {
    set_error("sps_nn_conv3x3_out1: cin a multiple of 8 in 8..64");
    return SPS_E_INVALID;
  }
  out1_fwd<<<(unsigned)((np + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      (const uint16_t*)x, n, h, w, cin, weight, bias, y);
  return launch_ok("sps_nn_conv3x3_out1");
}

int sps_nn_conv3x3_out1_bwd(const void* x, const float* dy, int n, int h,
                            int w, int cin, const float* weight, void* dx,
                            float* dweight, void* stream) {
  const int64_t np = (int64_t)n * h * w;
  if (!x || !dy || !weight || np < 1 || cin < 8 || cin > 64 || cin % 8) {
    set_error("sps_nn_conv3x3_out1_bwd: cin a multiple of 8 in 8..64");
    return SPS_E_INVALID;
  }
  hipStream_t s = (hipStream_t)stream;
  if (dx) {
    out1_bwd_data<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(
        dy, n, h, w, cin, weight, (uint16_t*)dx);
    SPS_TRY(launch_ok("sps_nn_conv3x3_out1_bwd (data)"));
  }
  if (dweight) {
    const int per = 16;
    const int64_t chunks = (np + 255) / 256;
    out1_wgrad<<<(unsigned)((chunks + per - 1) / per), 256, 0, s>>>(
        (const uint16_t*)x, dy, n, h, w, cin, dweight, per);
    SPS_TRY(launch_ok("sps_nn_conv3x3_out1_bwd (weight)"));
  }
  return SPS_OK;
}

int sps_nn_silu_bwd(const void* dy, const void* pre, void* dx, size_t count,
                    void* stream) {
  if (!dy || !pre || !dx) {
    set_error("sps_nn_silu_bwd: NULL pointer");
    return SPS_E_INVALID;
  }
  silu_bwd_k<<<grid_for((int64_t)count), 256, 0, (hipStream_t)stream>>>(
      (const uint16_t*)dy, (const uint16_t*)pre, (uint16_t*)dx, (int64_t)count);
  return launch_ok("sps_nn_silu_bwd");
}

int sps_nn_pool2(const void* x, int n, int h, int w, int c, int mode, void* y,
                 void* stream) {
  if (!x || !y || h % 2 || w % 2 || c < 1 || (mode != 0 && mode != 1)) {
    set_error("sps_nn_pool2: even h and w, mode 0 or 1");
    return SPS_E_INVALID;
  }
  pool2_k<<<grid_for((int64_t)n * h * w * c / 4), 256, 0, (hipStream_t)stream>>>(
      (const uint16_t*)x, n, h, w, c, mode, (uint16_t*)y);
  return launch_ok("sps_nn_pool2");
}

int sps_nn_unpool2(const void* x, int n, int h, int w, int c, float scale,
                   void* y, void* stream) {
  if (!x || !y || c < 1) {
    set_error("sps_nn_unpool2: invalid arguments");
    return SPS_E_INVALID;
  }
  unpool2_k<<<grid_for((int64_t)n * h * w * c * 4), 256, 0, (hipStream_t)stream>>>(
      (const uint16_t*)x, n, h, w, c, scale, (uint16_t*)y);
  return launch_ok("sps_nn_unpool2");
}

}  // extern "C"
