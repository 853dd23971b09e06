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
#include <algorithm>
#include <cstdint>
#include <initializer_list>

#include "common.hpp"
#include "../../include/sps_nn.h"

namespace sps {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4v __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even: gfx950's v_cvt_pk_bf16_f32 (a NaN
// stays a quiet NaN).  pack_bf16 converts two values in one instruction.
typedef float f32x2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint16_t f2bf(float f) {
  return __builtin_bit_cast(uint16_t, (__bf16)f);
}
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2v{lo, hi}, bf16x2v));
}
__device__ __forceinline__ float bf2f(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}

// v * sigmoid(v) with the hardware reciprocal (1 ulp; the result is rounded
// to BF16 or feeds BF16 data)
__device__ __forceinline__ float silu(float v) {
  return v * __builtin_amdgcn_rcpf(1.0f + __expf(-v));
}

__device__ __forceinline__ bf16x8 load_frag(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(p));
}

// 4 consecutive BF16 values as one 8-byte access (callers keep 8-byte
// alignment: channel offsets are multiples of 4)
__device__ __forceinline__ void store4(uint16_t* p, const float (&v)[4]) {
  const uint32_t lo = pack_bf16(v[0], v[1]);
  const uint32_t hi = pack_bf16(v[2], v[3]);
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
    w[i] = pack_bf16(v[2 * i], v[2 * i + 1]);
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

// Forward on spatial tiles staged in LDS.  A workgroup owns TH = NW*WN rows
// x 16 columns of one image and KB = 16*MW*WM output channels; the 4 waves
// split that WM x WN ways, each wave holding MW x NW accumulator tiles
// (16 channels x 16 pixels each, one N tile = 16 columns of one row).  Per
// 32-channel chunk of the input the tile plus its 1-pixel periodic halo,
// (TH+2) x 18 pixels, is copied to LDS once (double-buffered across chunks)
// and the 9 taps read their shifted B fragments from there: each activation
// value leaves HBM/L2 once per chunk instead of once per tap.  A fragments
// (weights, [tap][out][in]) are 16-byte global loads that every workgroup
// shares through L2.  A pixel's 32 channels are padded to 40 in LDS so the
// 16 pixels of one fragment start 20 banks apart.
// NCH > 0 (C = 32 * NCH input channels) stages every chunk at once, one LDS
// buffer per chunk, before the taps: the staging registers are free by then,
// which lets a CU hold more workgroups.  NCH = 0 loops over C / 32 chunks
// with two buffers, fetching the next chunk during the current one's taps.
constexpr int kTileCols = 16;
constexpr int kHaloCols = kTileCols + 2;
constexpr int kPixStride = 40;  // bf16 per staged pixel (32 channels + pad)

template <int MW, int NW, int WM, int WN, int EPI, int NCH>
__global__ __launch_bounds__(256) void conv_fwd_tile(
    const uint16_t* __restrict__ x, int N, int H, int W, int C,
    const uint16_t* __restrict__ wt, const float* __restrict__ bias, int K,
    int act, const uint16_t* __restrict__ skip, uint16_t* __restrict__ y,
    uint16_t* __restrict__ pre) {
  constexpr int TH = NW * WN;
  constexpr int KB = 16 * MW * WM;
  constexpr int TILE = (TH + 2) * kHaloCols;        // staged pixels
  constexpr int PIECES = TILE * 4;                  // 16-byte pieces per chunk
  constexpr int PER_THREAD = (PIECES + 255) / 256;
  // Epilogue through LDS when the block's FP32 results fit 40 KiB: the
  // workgroup then writes whole runs of pixels x channels per store (the
  // accumulator layout gives each lane 4 channels of one pixel, 8 bytes)
  constexpr int SKB = KB + 4;  // staged pixel stride: float4 writes of 16
                               // pixels start 16 bytes x 9 apart (no conflicts)
  constexpr int OUT_FLOATS = TH * kTileCols * SKB;
  constexpr bool STAGED = OUT_FLOATS * 4 <= 40960;
  constexpr int NBUF = NCH == 0 ? 2 : NCH;
  constexpr int TILE_BYTES = NBUF * TILE * kPixStride * 2;
  constexpr int LDS_BYTES =
      STAGED && OUT_FLOATS * 4 > TILE_BYTES ? OUT_FLOATS * 4 : TILE_BYTES;
  __shared__ __align__(16) unsigned char lds[LDS_BYTES];
  uint16_t(*tile)[TILE * kPixStride] =
      reinterpret_cast<uint16_t(*)[TILE * kPixStride]>(lds);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int tiles_w = W / kTileCols, tiles_h = H / TH;
  int t = blockIdx.x;
  const int tw = t % tiles_w;
  t /= tiles_w;
  const int th = t % tiles_h;
  const int n = t / tiles_h;
  const int h0 = th * TH, w0 = tw * kTileCols;
  const int k0 = blockIdx.y * KB + 16 * MW * wm;
  const int cq = (lane >> 4) * 8;

  // global source offsets of this thread's staging pieces (channel 0)
  int64_t src[PER_THREAD];
  int dst[PER_THREAD];
#pragma unroll
  for (int i = 0; i < PER_THREAD; ++i) {
    const int e = threadIdx.x + 256 * i;
    const int pix = e >> 2, q = e & 3;
    const int tr = pix / kHaloCols, tc = pix - tr * kHaloCols;
    const int row = wrapi(h0 + tr - 1, H), col = wrapi(w0 + tc - 1, W);
    src[i] = e < PIECES ? ((int64_t)(n * H + row) * W + col) * C + 8 * q : -1;
    dst[i] = pix * kPixStride + 8 * q;
  }
  constexpr int NST = NCH > 1 ? NCH : 1;  // staging register sets
  u32x4 stage[NST][PER_THREAD];
  auto fetch = [&](int c0, int k) {
#pragma unroll
    for (int i = 0; i < PER_THREAD; ++i)
      if (src[i] >= 0)
        stage[k][i] = *reinterpret_cast<const u32x4*>(x + src[i] + c0);
  };
  auto put = [&](int buf, int k) {
#pragma unroll
    for (int i = 0; i < PER_THREAD; ++i)
      if (src[i] >= 0)
        *reinterpret_cast<u32x4*>(&tile[buf][dst[i]]) = stage[k][i];
  };

  f32x4v acc[MW][NW];
#pragma unroll
  for (int m = 0; m < MW; ++m)
#pragma unroll
    for (int j = 0; j < NW; ++j) acc[m][j] = f32x4v{0.f, 0.f, 0.f, 0.f};

  const int nchunks = NCH ? NCH : C / 32;
  // NCH > 0: every chunk's loads in flight before the first LDS store
#pragma unroll
  for (int c = 0; c < NST; ++c) fetch(32 * c, c);
#pragma unroll
  for (int c = 0; c < NST; ++c) put(c, c);
  for (int ch = 0; ch < nchunks; ++ch) {
    const int buf = NCH ? ch : ch & 1;
    const int c0 = 32 * ch;
    // weight fragments are loaded one tap ahead of their MFMAs (loading all
    // nine taps' at once costs the waves a SIMD can hold)
    bf16x8 apre[2][MW];
    auto load_a = [&](int tap, bf16x8 (&dst)[MW]) {
#pragma unroll
      for (int m = 0; m < MW; ++m)
        dst[m] = load_frag(wt + ((int64_t)tap * K + k0 + 16 * m + (lane & 15)) * C +
                           c0 + cq);
    };
    load_a(0, apre[0]);
    // NCH = 0: tile[buf] written and tile[buf ^ 1] free; NCH > 0: every
    // chunk's tile written
    if (NCH == 0 || ch == 0) __syncthreads();
    if (NCH == 0 && ch + 1 < nchunks) fetch(32 * (ch + 1), 0);
    const uint16_t* tb = tile[buf];
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int dr = tap / 3, ds = tap % 3;  // halo-shifted offsets 0..2
      bf16x8 b[NW];
      if (tap + 1 < 9) load_a(tap + 1, apre[(tap + 1) & 1]);
      const bf16x8 (&a)[MW] = apre[tap & 1];
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        const int tr = wn * NW + j + dr;
        const int tc = (lane & 15) + ds;
        b[j] = __builtin_bit_cast(
            bf16x8, *reinterpret_cast<const u32x4*>(
                        tb + (tr * kHaloCols + tc) * kPixStride + cq));
      }
#pragma unroll
      for (int m = 0; m < MW; ++m)
#pragma unroll
        for (int j = 0; j < NW; ++j)
          acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[j],
                                                              acc[m][j], 0, 0, 0);
    }
    if (NCH == 0 && ch + 1 < nchunks) put(buf ^ 1, 0);
  }

  if constexpr (STAGED) {
    // stage[pixel][channel] in FP32, pixel = row * 16 + column of the tile
    float* stage = reinterpret_cast<float*>(lds);
    const int kb0 = blockIdx.y * KB;
    __syncthreads();  // every wave is done with the input tile
#pragma unroll
    for (int m = 0; m < MW; ++m) {
      const int cl = 16 * MW * wm + 16 * m + 4 * (lane >> 4);
      float bv[4] = {0.f, 0.f, 0.f, 0.f};
      if (bias) {
#pragma unroll
        for (int i = 0; i < 4; ++i) bv[i] = bias[kb0 + cl + i];
      }
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        const int px = (wn * NW + j) * kTileCols + (lane & 15);
        *reinterpret_cast<f32x4v*>(&stage[px * SKB + cl]) =
            f32x4v{acc[m][j][0] + bv[0], acc[m][j][1] + bv[1], acc[m][j][2] + bv[2],
                   acc[m][j][3] + bv[3]};
      }
    }
    __syncthreads();
    constexpr int G = KB / 8;  // 8-channel groups per pixel
    if constexpr (EPI == SPS_NN_EPI_PLAIN) {
      for (int e = threadIdx.x; e < TH * kTileCols * G; e += 256) {
        const int px = e / G, g = e - px * G;
        const int hh = h0 + px / kTileCols, ww = w0 + px % kTileCols;
        float v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = stage[px * SKB + 8 * g + i];
        const int64_t o = ((int64_t)(n * H + hh) * W + ww) * K + kb0 + 8 * g;
        if (pre) store8(pre + o, v);
        if (act) {
#pragma unroll
          for (int i = 0; i < 8; ++i) v[i] = silu(v[i]);
        }
        store8(y + o, v);
      }
    } else {
      // output (2*TH) x 32 pixels, each from tile pixel (r/2, c/2); the
      // skip pieces of U iterations are loaded before any is used, so their
      // memory latencies overlap instead of adding up
      constexpr int ITERS = 4 * TH * kTileCols * G / 256;
      static_assert(4 * TH * kTileCols * G % 256 == 0, "whole iterations");
      constexpr int U = ITERS % 4 == 0 ? 4 : 1;
      for (int it = 0; it < ITERS; it += U) {
        int64_t o[U];
        int sp[U];
        float u[U][8];
#pragma unroll
        for (int q = 0; q < U; ++q) {
          const int e = threadIdx.x + 256 * (it + q);
          const int opx = e / G, g = e - opx * G;
          const int orow = opx / (2 * kTileCols), ocol = opx % (2 * kTileCols);
          sp[q] = ((orow >> 1) * kTileCols + (ocol >> 1)) * SKB + 8 * g;
          o[q] = ((int64_t)(n * 2 * H + 2 * h0 + orow) * (2 * W) + 2 * w0 + ocol) * K +
                 kb0 + 8 * g;
          load8(skip + o[q], u[q]);
        }
#pragma unroll
        for (int q = 0; q < U; ++q) {
#pragma unroll
          for (int i = 0; i < 8; ++i) u[q][i] += stage[sp[q] + i];
          if (pre) store8(pre + o[q], u[q]);
          if (act) {
#pragma unroll
            for (int i = 0; i < 8; ++i) u[q][i] = silu(u[q][i]);
          }
          store8(y + o[q], u[q]);
        }
      }
    }
    return;
  }
#pragma unroll
  for (int m = 0; m < MW; ++m) {
    const int chn = k0 + 16 * m + 4 * (lane >> 4);  // 4 consecutive channels
    float bv[4] = {0.f, 0.f, 0.f, 0.f};
    if (bias) {
#pragma unroll
      for (int i = 0; i < 4; ++i) bv[i] = bias[chn + i];
    }
#pragma unroll
    for (int j = 0; j < NW; ++j) {
      const int hh = h0 + wn * NW + j, ww = w0 + (lane & 15);
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = acc[m][j][i] + bv[i];
      if constexpr (EPI == SPS_NN_EPI_PLAIN) {
        const int64_t o = ((int64_t)(n * H + hh) * W + ww) * K + chn;
        if (pre) store4(pre + o, v);
        if (act) {
#pragma unroll
          for (int i = 0; i < 4; ++i) v[i] = silu(v[i]);
        }
        store4(y + o, v);
      } else {
#pragma unroll
        for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
          for (int b2 = 0; b2 < 2; ++b2) {
            const int64_t o =
                ((int64_t)(n * 2 * H + 2 * hh + a2) * (2 * W) + 2 * ww + b2) * K +
                chn;
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
    int H, int W, int C, int K, float* __restrict__ part, int nchunks,
    int chunks_per_block) {
  constexpr int TK = KT / 32, TC = CT / 32;  // MFMA tiles per wave
  __shared__ __align__(16) uint16_t sdy[32][KT + 2];
  __shared__ __align__(16) uint16_t sx[ROWS ? 3 : 9][ROWS ? 34 : 32][CT + 2];
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
    // staging in pieces of 8 channels: one 16-byte global load, written to
    // the padded LDS row as four 32-bit stores (rows are 4-byte aligned)
    auto put8 = [](uint16_t* dst, const uint16_t* src) {
      const u32x4 q = *reinterpret_cast<const u32x4*>(src);
      uint32_t* d = reinterpret_cast<uint32_t*>(dst);
      d[0] = q.x;
      d[1] = q.y;
      d[2] = q.z;
      d[3] = q.w;
    };
    if constexpr (ROWS) {
      const int n = ch / (H * rowchunks);
      const int rem = ch - n * H * rowchunks;
      const int h = rem / rowchunks;
      const int w0 = (rem - h * rowchunks) * 32;
      // stage dy: 32 pixels x KT channels
      for (int e = threadIdx.x; e < 32 * (KT / 8); e += 256) {
        const int px = e / (KT / 8), k = (e - px * (KT / 8)) * 8;
        put8(&sdy[px][k], dy + ((int64_t)(n * H + h) * W + w0 + px) * K + k0 + k);
      }
      // stage x: rows h-1..h+1, columns w0-1..w0+32, CT channels
      for (int e = threadIdx.x; e < 3 * 34 * (CT / 8); e += 256) {
        const int r = e / (34 * (CT / 8));
        const int e2 = e - r * 34 * (CT / 8);
        const int col = e2 / (CT / 8), c = (e2 - col * (CT / 8)) * 8;
        const int hh = wrapi(h + r - 1, H), ww = wrapi(w0 + col - 1, W);
        put8(&sx[r][col][c], x + ((int64_t)(n * H + hh) * W + ww) * C + c0 + c);
      }
    } else {
      const int64_t pbase = (int64_t)ch * 32;
      for (int e = threadIdx.x; e < 32 * (KT / 8); e += 256) {
        const int px = e / (KT / 8), k = (e - px * (KT / 8)) * 8;
        put8(&sdy[px][k], dy + (pbase + px) * K + k0 + k);
      }
      const int HW = H * W;
      for (int e = threadIdx.x; e < 9 * 32 * (CT / 8); e += 256) {
        const int t = e / (32 * (CT / 8));
        const int e2 = e - t * 32 * (CT / 8);
        const int px = e2 / (CT / 8), c = (e2 - px * (CT / 8)) * 8;
        const int64_t p = pbase + px;
        const int n = (int)(p / HW);
        const int r = (int)(p - (int64_t)n * HW);
        const int h = r / W, w = r - h * W;
        const int hh = wrapi(h + t / 3 - 1, H), ww = wrapi(w + t % 3 - 1, W);
        put8(&sx[t][px][c], x + ((int64_t)(n * H + hh) * W + ww) * C + c0 + c);
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
  // this block's partial sums (zeros if it had no chunk): slice blockIdx.z
  float* dw = part + (int64_t)blockIdx.z * 9 * K * C;
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
          dw[((int64_t)tap * K + k) * C + c] = acc[tap][tk][tc][i];
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

// The same first layer with its stores staged through LDS: each thread
// computes all K channels of one pixel exactly as in1_fwd does (bias, then
// the 9 taps in order, FMA; the weights are uniform across the wave), then
// the workgroup writes its 256 consecutive pixels x K channels, one
// contiguous run, in 16-byte pieces.
template <int K>
__global__ __launch_bounds__(256) void in1_fwd_staged(
    const float* __restrict__ x, int N, int H, int W,
    const float* __restrict__ wt, const float* __restrict__ bias, int act,
    uint16_t* __restrict__ y, uint16_t* __restrict__ pre) {
  __shared__ __align__(16) uint16_t sy[256 * K];
  __shared__ __align__(16) uint16_t sp[256 * K];
  const int64_t np = (int64_t)N * H * W;
  const int64_t p0 = (int64_t)blockIdx.x * 256;
  const int64_t p = p0 + threadIdx.x;
  if (p < np) {
    const int HW = H * W;
    const int n = p / HW, r = p - (int64_t)n * HW, h = r / W, w = r - h * W;
    float xv[9];
#pragma unroll
    for (int t = 0; t < 9; ++t)
      xv[t] = x[(int64_t)(n * H + wrapi(h + t / 3 - 1, H)) * W +
                wrapi(w + t % 3 - 1, W)];
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 8) {
      float v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        float a = bias ? bias[k0 + i] : 0.f;
#pragma unroll
        for (int t = 0; t < 9; ++t) a = fmaf(xv[t], wt[t * K + k0 + i], a);
        v[i] = a;
      }
      if (pre) store8(&sp[threadIdx.x * K + k0], v);
      if (act) {
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = silu(v[i]);
      }
      store8(&sy[threadIdx.x * K + k0], v);
    }
  }
  __syncthreads();
  const int64_t valid = np - p0 < 256 ? np - p0 : 256;
  const int pieces = (int)valid * K / 8;
  for (int e = threadIdx.x; e < pieces; e += 256) {
    *reinterpret_cast<u32x4*>(y + p0 * K + 8 * e) =
        *reinterpret_cast<const u32x4*>(&sy[8 * e]);
    if (pre)
      *reinterpret_cast<u32x4*>(pre + p0 * K + 8 * e) =
          *reinterpret_cast<const u32x4*>(&sp[8 * e]);
  }
}

// The last layer on 16 x 16 pixel tiles: the tile and its periodic halo
// (18 x 18 pixels x C channels) are staged in LDS once, so each activation
// is read from HBM once instead of once per tap; the weights are uniform
// across the wave.  Same arithmetic order as out1_fwd (bias, then taps, then
// channels in order).
template <int C>
__global__ __launch_bounds__(256) void out1_fwd_tile(
    const uint16_t* __restrict__ x, int N, int H, int W,
    const float* __restrict__ wt, const float* __restrict__ bias,
    float* __restrict__ y) {
  constexpr int T = 16, HT = T + 2;
  constexpr int PS = C + 8;  // padded pixel stride
  __shared__ __align__(16) uint16_t tile[HT * HT * PS];
  const int tiles_w = W / T, tiles_h = H / T;
  int b = blockIdx.x;
  const int tw = b % tiles_w;
  b /= tiles_w;
  const int th = b % tiles_h;
  const int n = b / tiles_h;
  const int h0 = th * T, w0 = tw * T;
  constexpr int groups = C / 8;
  for (int e = threadIdx.x; e < HT * HT * groups; e += 256) {
    const int pix = e / groups, q = e - pix * groups;
    const int tr = pix / HT, tc = pix - tr * HT;
    const int row = wrapi(h0 + tr - 1, H), col = wrapi(w0 + tc - 1, W);
    *reinterpret_cast<u32x4*>(&tile[pix * PS + 8 * q]) =
        *reinterpret_cast<const u32x4*>(x + ((int64_t)(n * H + row) * W + col) * C +
                                        8 * q);
  }
  __syncthreads();
  const int ty = threadIdx.x / T, tx = threadIdx.x % T;
  float v = bias ? bias[0] : 0.f;
  for (int t = 0; t < 9; ++t) {
    const uint16_t* xp = &tile[((ty + t / 3) * HT + tx + t % 3) * PS];
#pragma unroll
    for (int c0 = 0; c0 < C; c0 += 8) {
      float xv[8];
      load8(xp + c0, xv);
#pragma unroll
      for (int i = 0; i < 8; ++i) v = fmaf(xv[i], wt[t * C + c0 + i], v);
    }
  }
  y[(int64_t)(n * H + h0 + ty) * W + w0 + tx] = v;
}

// dweight[t][k] += sum_p x[shift_t(p)] * dy[p][k]: 256-pixel chunks staged
// in LDS, then thread (t, k) pairs reduce the chunk.
__global__ __launch_bounds__(256) void in1_wgrad(const float* __restrict__ x,
                                                 const uint16_t* __restrict__ dy,
                                                 int N, int H, int W, int K,
                                                 float* __restrict__ part,
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
  float* dw = part + (int64_t)blockIdx.x * 9 * K;  // this block's partials
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int tk = threadIdx.x + 256 * q;
    if (tk < 9 * K) dw[tk] = acc[q];
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
                                                  float* __restrict__ part,
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
  float* dw = part + (int64_t)blockIdx.x * 9 * C;  // this block's partials
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int tc = threadIdx.x + 256 * q;
    if (tc < 9 * C) dw[tc] = acc[q];
  }
}

// Weight gradient of the last layer on 16 x 16 pixel tiles: the x tile with
// its periodic halo (18 x 18 pixels x C channels, BF16) and the tile's dy
// (FP32) are staged in LDS, and thread (t, c) sums x[shift_t(p)][c] * dy[p]
// over the tile's pixels in a fixed order.  A block walks tiles blockIdx.x,
// + gridDim.x, ... and writes its partial sums to part[blockIdx.x][9][C].
template <int C>
__global__ __launch_bounds__(256) void out1_wgrad_tile(
    const uint16_t* __restrict__ x, const float* __restrict__ dy, int N, int H,
    int W, float* __restrict__ part) {
  constexpr int T = 16, HT = T + 2, PS = C + 8, G = C / 8;
  __shared__ __align__(16) uint16_t sx[HT * HT * PS];
  __shared__ float sd[T * T];
  const int tiles_w = W / T, tiles_h = H / T;
  const int ntiles = N * tiles_h * tiles_w;
  float acc[3] = {0.f, 0.f, 0.f};
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    int b = tile;
    const int tw = b % tiles_w;
    b /= tiles_w;
    const int th = b % tiles_h;
    const int n = b / tiles_h;
    const int h0 = th * T, w0 = tw * T;
    __syncthreads();  // the previous tile's reads are done
    for (int e = threadIdx.x; e < HT * HT * G; e += 256) {
      const int pix = e / G, q = e - pix * G;
      const int tr = pix / HT, tc = pix - tr * HT;
      const int row = wrapi(h0 + tr - 1, H), col = wrapi(w0 + tc - 1, W);
      *reinterpret_cast<u32x4*>(&sx[pix * PS + 8 * q]) =
          *reinterpret_cast<const u32x4*>(x + ((int64_t)(n * H + row) * W + col) * C + 8 * q);
    }
    sd[threadIdx.x] =
        dy[(int64_t)(n * H + h0 + threadIdx.x / T) * W + w0 + threadIdx.x % T];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int tc = threadIdx.x + 256 * q;
      if (tc < 9 * C) {
        const int t = tc / C, ch = tc - t * C;
        const int dr = t / 3, ds = t % 3;
        float a = acc[q];
        for (int py = 0; py < T; ++py)
#pragma unroll 4
          for (int px = 0; px < T; ++px)
            a = fmaf(bf2f(sx[((py + dr) * HT + px + ds) * PS + ch]), sd[py * T + px], a);
        acc[q] = a;
      }
    }
  }
  float* dw = part + (int64_t)blockIdx.x * 9 * C;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int tc = threadIdx.x + 256 * q;
    if (tc < 9 * C) dw[tc] = acc[q];
  }
}

// The deterministic sum of the per-block partial weight gradients, as a
// fixed tree: the live partials sit in slots m*S of the workspace (m < M,
// `count` floats per slot).  sum_partial_groups adds each run of gs of them,
// in order, into the run's first slot (one thread per (run, weight), in
// place); add_partials then adds the last M <= gs, in order, into dw.
__global__ __launch_bounds__(256) void sum_partial_groups(float* __restrict__ part,
                                                          int64_t count, int M,
                                                          int64_t S, int gs) {
  const int64_t ng = (M + gs - 1) / gs;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < ng * count;
       t += (int64_t)gridDim.x * 256) {
    const int64_t g = t / count, i = t - g * count;
    const int m0 = (int)g * gs, m1 = min(M, m0 + gs);
    float s = 0.f;
    for (int m = m0; m < m1; ++m) s += part[(int64_t)m * S * count + i];
    part[(int64_t)m0 * S * count + i] = s;
  }
}

__global__ __launch_bounds__(256) void add_partials(const float* __restrict__ part,
                                                    int M, int64_t S, int64_t count,
                                                    float* __restrict__ dw) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * 256) {
    float s = 0.f;
    for (int m = 0; m < M; ++m) s += part[(int64_t)m * S * count + i];
    dw[i] += s;
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
    const float s = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
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

// The same three passes on 8 consecutive channels per thread: one 16-byte
// access per operand and 32-bit index arithmetic, used when the channel count
// is a multiple of 8, the piece count fits 31 bits and the pointers are
// 16-byte aligned.  Each value goes through the same operations in the same
// order as in the one-value kernels above, so results are identical.
__global__ __launch_bounds__(256) void silu_bwd_v8(const uint16_t* __restrict__ dy,
                                                   const uint16_t* __restrict__ pre,
                                                   uint16_t* __restrict__ dx,
                                                   unsigned pieces) {
  for (unsigned i = blockIdx.x * 256 + threadIdx.x; i < pieces; i += gridDim.x * 256) {
    float z[8], d[8], v[8];
    load8(pre + 8 * (int64_t)i, z);
    load8(dy + 8 * (int64_t)i, d);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float s = __builtin_amdgcn_rcpf(1.0f + __expf(-z[k]));
      v[k] = d[k] * s * (1.0f + z[k] * (1.0f - s));
    }
    store8(dx + 8 * (int64_t)i, v);
  }
}

__global__ __launch_bounds__(256) void pool2_v8(const uint16_t* __restrict__ x,
                                                int H, int W, int C, int mode,
                                                unsigned pieces,
                                                uint16_t* __restrict__ y) {
  const unsigned G = (unsigned)C / 8, Ho = (unsigned)H / 2, Wo = (unsigned)W / 2;
  const float scale = mode == 0 ? 0.25f : 1.0f;
  const int64_t rowc = (int64_t)W * C;
  for (unsigned i = blockIdx.x * 256 + threadIdx.x; i < pieces; i += gridDim.x * 256) {
    const unsigned g = i % G;
    unsigned q = i / G;
    const unsigned wo = q % Wo;
    q /= Wo;
    const unsigned ho = q % Ho;
    const unsigned n = q / Ho;
    const uint16_t* b = x + ((int64_t)(n * H + 2 * ho) * W + 2 * wo) * C + 8 * g;
    float a0[8], a1[8], a2[8], a3[8], v[8];
    load8(b, a0);
    load8(b + C, a1);
    load8(b + rowc, a2);
    load8(b + rowc + C, a3);
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = ((a0[k] + a1[k]) + (a2[k] + a3[k])) * scale;
    store8(y + 8 * (int64_t)i, v);
  }
}

__global__ __launch_bounds__(256) void unpool2_v8(const uint16_t* __restrict__ x,
                                                  int H, int W, int C, float scale,
                                                  unsigned pieces,
                                                  uint16_t* __restrict__ y) {
  const unsigned G = (unsigned)C / 8, Ho = 2u * H, Wo = 2u * W;
  for (unsigned i = blockIdx.x * 256 + threadIdx.x; i < pieces; i += gridDim.x * 256) {
    const unsigned g = i % G;
    unsigned q = i / G;
    const unsigned wo = q % Wo;
    q /= Wo;
    const unsigned ho = q % Ho;
    const unsigned n = q / Ho;
    float v[8];
    load8(x + ((int64_t)(n * H + ho / 2) * W + wo / 2) * C + 8 * g, v);
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] *= scale;
    store8(y + 8 * (int64_t)i, v);
  }
}

// One workgroup per sample: FP64 sum in a fixed order (strided per thread,
// then xor-shuffle and LDS across the 4 waves), then y = x - mean.
__global__ __launch_bounds__(256) void center_k(const float* x, float* y,
                                                int64_t len) {
  const float* xs = x + (int64_t)blockIdx.x * len;
  float* ys = y + (int64_t)blockIdx.x * len;
  // U loads in flight per thread before their adds, which keep the order
  // x[t], x[t + 256], x[t + 512], ... of thread t
  constexpr int U = 16;
  double s = 0.0;
  int64_t i = threadIdx.x;
  for (; i + (U - 1) * 256 < len; i += U * 256) {
    float v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = xs[i + u * 256];
#pragma unroll
    for (int u = 0; u < U; ++u) s += (double)v[u];
  }
  for (; i < len; i += 256) s += (double)xs[i];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
  __shared__ double sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  const float mean = (float)(((sh[0] + sh[1]) + (sh[2] + sh[3])) / (double)len);
  // each thread reads its U values before writing them (x and y may alias)
  for (i = threadIdx.x; i + (U - 1) * 256 < len; i += U * 256) {
    float v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = xs[i + u * 256];
#pragma unroll
    for (int u = 0; u < U; ++u) ys[i + u * 256] = v[u] - mean;
  }
  for (; i < len; i += 256) ys[i] = xs[i] - mean;
}

bool aligned16(std::initializer_list<const void*> ps) {
  for (const void* p : ps)
    if (reinterpret_cast<uintptr_t>(p) % 16) return false;
  return true;
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

// How each weight gradient splits its pixel reduction over workgroups; the
// workspace holds one partial gradient per workgroup (per z slice).
constexpr int64_t kPartFloats = (int64_t)16 << 20;  // partial sums, at most

struct WgradSplit {
  int blocks;    // workgroups that write a partial gradient
  int per;       // chunks / tiles per workgroup
  int64_t count; // floats of one partial gradient
};

WgradSplit conv_split(int64_t npix, int cin, int cout) {
  const int kt = cout % 64 == 0 ? 64 : 32, ct = cin % 64 == 0 ? 64 : 32;
  const int64_t nchunks = npix / 32;
  const int64_t count = (int64_t)9 * cin * cout;
  const int64_t tiles = (int64_t)(cin / ct) * (cout / kt);
  int64_t nz = (2048 + tiles - 1) / tiles;
  nz = std::min<int64_t>(nz, std::max<int64_t>(1, kPartFloats / count));
  nz = std::max<int64_t>(1, std::min<int64_t>(nz, nchunks));
  const int64_t per = (nchunks + nz - 1) / nz;
  return {(int)((nchunks + per - 1) / per), (int)per, count};
}

WgradSplit chunk_split(int64_t np, int64_t count) {
  const int64_t chunks = (np + 255) / 256;
  const int64_t per = std::max<int64_t>(16, (chunks + 2047) / 2048);
  return {(int)((chunks + per - 1) / per), (int)per, count};
}

bool out1_tiled(int h, int w, int c) {
  return h % 16 == 0 && w % 16 == 0 && (c == 32 || c == 64);
}

WgradSplit out1_split(int n, int h, int w, int c) {
  if (!out1_tiled(h, w, c)) return chunk_split((int64_t)n * h * w, 9 * c);
  const int64_t tiles = (int64_t)n * (h / 16) * (w / 16);
  const int64_t per = (tiles + 2047) / 2048;
  return {(int)((tiles + per - 1) / per), (int)per, (int64_t)9 * c};
}

// The workspace's partials are consumed (summed in place).
int add_parts(float* part, const WgradSplit& sp, float* dw, hipStream_t s) {
  constexpr int kGroup = 32;
  int M = sp.blocks;
  int64_t S = 1;
  while (M > kGroup) {
    const int64_t ng = (M + kGroup - 1) / kGroup;
    sum_partial_groups<<<grid_for(ng * sp.count), 256, 0, s>>>(part, sp.count, M, S, kGroup);
    SPS_TRY(launch_ok("weight-gradient partial sum"));
    M = (int)ng;
    S *= kGroup;
  }
  add_partials<<<grid_for(sp.count), 256, 0, s>>>(part, M, S, sp.count, dw);
  return launch_ok("weight-gradient partial sum");
}

int check_ws(const char* who, const WgradSplit& sp, const float* ws,
             size_t ws_floats) {
  if (!ws || (int64_t)ws_floats < (int64_t)sp.blocks * sp.count) {
    set_error("%s: workspace of %zu floats is smaller than the %lld it needs "
              "(sps_nn_wgrad_workspace)", who, ws_floats,
              (long long)sp.blocks * sp.count);
    return SPS_E_INVALID;
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
  const uint16_t* xx = (const uint16_t*)x;
  const uint16_t* ww = (const uint16_t*)weight;
  const uint16_t* sk = (const uint16_t*)skip;
  uint16_t* yy = (uint16_t*)y;
  uint16_t* pp = (uint16_t*)pre;
  // LDS-tiled kernel when the image splits into whole tiles: 16 columns by
  // 8 or 16 rows, 32 / 64 / 128 output channels per workgroup
  if (w % kTileCols == 0) {
    const int tiles_w = w / kTileCols;
#define SPS_CONV_TILE(MW, NW, WM, WN)                                           \
  do {                                                                         \
    constexpr int TH_ = (NW) * (WN), KB_ = 16 * (MW) * (WM);                   \
    /* two chunks staged at once only where registers allow 3 waves/SIMD */   \
    constexpr int C2_ = (MW) == 2 ? 2 : 0;                                     \
    dim3 g((unsigned)(n * (h / TH_) * tiles_w), (unsigned)(cout / KB_));      \
    if (epilogue == SPS_NN_EPI_PLAIN && cin == 32)                             \
      conv_fwd_tile<MW, NW, WM, WN, 0, 1><<<g, 256, 0, s>>>(                    \
          xx, n, h, w, cin, ww, bias, cout, act, sk, yy, pp);                  \
    else if (epilogue == SPS_NN_EPI_PLAIN && cin == 64 && C2_)                 \
      conv_fwd_tile<MW, NW, WM, WN, 0, C2_><<<g, 256, 0, s>>>(                  \
          xx, n, h, w, cin, ww, bias, cout, act, sk, yy, pp);                  \
    else if (epilogue == SPS_NN_EPI_PLAIN)                                     \
      conv_fwd_tile<MW, NW, WM, WN, 0, 0><<<g, 256, 0, s>>>(                    \
          xx, n, h, w, cin, ww, bias, cout, act, sk, yy, pp);                  \
    else if (cin == 64 && C2_)                                                 \
      conv_fwd_tile<MW, NW, WM, WN, 1, C2_><<<g, 256, 0, s>>>(                  \
          xx, n, h, w, cin, ww, bias, cout, act, sk, yy, pp);                  \
    else                                                                       \
      conv_fwd_tile<MW, NW, WM, WN, 1, 0><<<g, 256, 0, s>>>(                    \
          xx, n, h, w, cin, ww, bias, cout, act, sk, yy, pp);                  \
    return launch_ok("sps_nn_conv3x3");                                        \
  } while (0)
    // the upsample-add epilogue stages through LDS: 64 channels x 8 rows
    if (epilogue == SPS_NN_EPI_UPSAMPLE_ADD && cout % 64 == 0 && h % 8 == 0)
      SPS_CONV_TILE(4, 2, 1, 4);
    if (cout % 128 == 0 && h % 8 == 0) SPS_CONV_TILE(4, 4, 2, 2);
    if (cout % 64 == 0 && h % 16 == 0) SPS_CONV_TILE(4, 4, 1, 4);
    if (cout % 64 == 0 && h % 8 == 0) SPS_CONV_TILE(4, 2, 1, 4);
    if (h % 16 == 0) SPS_CONV_TILE(2, 4, 1, 4);
    if (h % 8 == 0) SPS_CONV_TILE(2, 2, 1, 4);
#undef SPS_CONV_TILE
  }
  const int nt = cout % 64 == 0 ? 4 : 2;
  dim3 grid((unsigned)((npix / 64 + 3) / 4), (unsigned)(cout / (16 * nt)));
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

size_t sps_nn_wgrad_workspace(int kind, int n, int h, int w, int cin,
                              int cout) {
  if (n < 1 || h < 1 || w < 1 || cin < 1 || cout < 1) return 0;
  const int64_t np = (int64_t)n * h * w;
  WgradSplit sp{0, 0, 0};
  if (kind == SPS_NN_WGRAD_CONV) sp = conv_split(np, cin, cout);
  else if (kind == SPS_NN_WGRAD_IN1) sp = chunk_split(np, (int64_t)9 * cout);
  else if (kind == SPS_NN_WGRAD_OUT1) sp = out1_split(n, h, w, cin);
  return (size_t)((int64_t)sp.blocks * sp.count);
}

int sps_nn_conv3x3_wgrad(const void* x, const void* dpre, int n, int h, int w,
                         int cin, int cout, float* dweight, float* workspace,
                         size_t ws_floats, void* stream) {
  const int64_t npix = (int64_t)n * h * w;
  if (!x || !dpre || !dweight || n < 1 || h < 1 || w < 1 || npix % 32 ||
      npix > (int64_t)1 << 30 || cin % 32 || cout % 32 || cin < 32 ||
      cout < 32) {
    set_error("sps_nn_conv3x3_wgrad: need n*h*w, cin, cout multiples of 32");
    return SPS_E_INVALID;
  }
  const WgradSplit sp = conv_split(npix, cin, cout);
  SPS_TRY(check_ws("sps_nn_conv3x3_wgrad", sp, workspace, ws_floats));
  const bool rows = w % 32 == 0;
  const int nchunks = (int)(npix / 32);
  const int per = sp.per;
  const int kt = cout % 64 == 0 ? 64 : 32;
  const int ct = cin % 64 == 0 ? 64 : 32;
  dim3 grid(cin / ct, cout / kt, sp.blocks);
  hipStream_t s = (hipStream_t)stream;
  const uint16_t* xx = (const uint16_t*)x;
  const uint16_t* dd = (const uint16_t*)dpre;
#define SPS_WGRAD(KT_, CT_)                                                   \
  (rows ? (conv_wgrad<KT_, CT_, true><<<grid, 256, 0, s>>>(                   \
               xx, dd, n, h, w, cin, cout, workspace, nchunks, per), 0)      \
        : (conv_wgrad<KT_, CT_, false><<<grid, 256, 0, s>>>(                  \
               xx, dd, n, h, w, cin, cout, workspace, nchunks, per), 0))
  if (kt == 64 && ct == 64)
    SPS_WGRAD(64, 64);
  else if (kt == 64)
    SPS_WGRAD(64, 32);
  else if (ct == 64)
    SPS_WGRAD(32, 64);
  else
    SPS_WGRAD(32, 32);
#undef SPS_WGRAD
  SPS_TRY(launch_ok("sps_nn_conv3x3_wgrad"));
  return add_parts(workspace, sp, dweight, s);
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
  const unsigned blocks = (unsigned)((np + 255) / 256);
  if (cout == 32)
    in1_fwd_staged<32><<<blocks, 256, 0, (hipStream_t)stream>>>(
        x, n, h, w, weight, bias, act, (uint16_t*)y, (uint16_t*)pre);
  else if (cout == 64)
    in1_fwd_staged<64><<<blocks, 256, 0, (hipStream_t)stream>>>(
        x, n, h, w, weight, bias, act, (uint16_t*)y, (uint16_t*)pre);
  else
    in1_fwd<<<blocks, 256, 0, (hipStream_t)stream>>>(
        x, n, h, w, weight, bias, cout, act, (uint16_t*)y, (uint16_t*)pre);
  return launch_ok("sps_nn_conv3x3_in1");
}

int sps_nn_conv3x3_in1_wgrad(const float* x, const void* dpre, int n, int h,
                             int w, int cout, float* dweight, float* workspace,
                             size_t ws_floats, void* stream) {
  const int64_t np = (int64_t)n * h * w;
  if (!x || !dpre || !dweight || np < 1 || cout < 8 || cout > 64 || cout % 8) {
    set_error("sps_nn_conv3x3_in1_wgrad: cout a multiple of 8 in 8..64");
    return SPS_E_INVALID;
  }
  const WgradSplit sp = chunk_split(np, (int64_t)9 * cout);
  SPS_TRY(check_ws("sps_nn_conv3x3_in1_wgrad", sp, workspace, ws_floats));
  hipStream_t s = (hipStream_t)stream;
  in1_wgrad<<<(unsigned)sp.blocks, 256, 0, s>>>(x, (const uint16_t*)dpre, n, h,
                                                 w, cout, workspace, sp.per);
  SPS_TRY(launch_ok("sps_nn_conv3x3_in1_wgrad"));
  return add_parts(workspace, sp, dweight, s);
}

int sps_nn_conv3x3_out1(const void* x, int n, int h, int w, int cin,
                        const float* weight, const float* bias, float* y,
                        void* stream) {
  const int64_t np = (int64_t)n * h * w;
  if (!x || !weight || !y || np < 1 || cin < 8 || cin > 64 || cin % 8) {
    set_error("sps_nn_conv3x3_out1: cin a multiple of 8 in 8..64");
    return SPS_E_INVALID;
  }
  const unsigned tiles = (unsigned)(n * (h / 16) * (w / 16));
  if (h % 16 == 0 && w % 16 == 0 && cin == 32)
    out1_fwd_tile<32><<<tiles, 256, 0, (hipStream_t)stream>>>(
        (const uint16_t*)x, n, h, w, weight, bias, y);
  else if (h % 16 == 0 && w % 16 == 0 && cin == 64)
    out1_fwd_tile<64><<<tiles, 256, 0, (hipStream_t)stream>>>(
        (const uint16_t*)x, n, h, w, weight, bias, y);
  else
    out1_fwd<<<(unsigned)((np + 255) / 256), 256, 0, (hipStream_t)stream>>>(
        (const uint16_t*)x, n, h, w, cin, weight, bias, y);
  return launch_ok("sps_nn_conv3x3_out1");
}

int sps_nn_conv3x3_out1_bwd(const void* x, const float* dy, int n, int h,
                            int w, int cin, const float* weight, void* dx,
                            float* dweight, float* workspace, size_t ws_floats,
                            void* stream) {
  const int64_t np = (int64_t)n * h * w;
  if (!x || !dy || !weight || np < 1 || cin < 8 || cin > 64 || cin % 8) {
    set_error("sps_nn_conv3x3_out1_bwd: cin a multiple of 8 in 8..64");
    return SPS_E_INVALID;
  }
  hipStream_t s = (hipStream_t)stream;
  const WgradSplit sp = out1_split(n, h, w, cin);
  if (dweight) SPS_TRY(check_ws("sps_nn_conv3x3_out1_bwd", sp, workspace, ws_floats));
  if (dx) {
    out1_bwd_data<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(
        dy, n, h, w, cin, weight, (uint16_t*)dx);
    SPS_TRY(launch_ok("sps_nn_conv3x3_out1_bwd (data)"));
  }
  if (dweight) {
    const uint16_t* xx = (const uint16_t*)x;
    if (out1_tiled(h, w, cin) && cin == 32)
      out1_wgrad_tile<32><<<(unsigned)sp.blocks, 256, 0, s>>>(xx, dy, n, h, w, workspace);
    else if (out1_tiled(h, w, cin))
      out1_wgrad_tile<64><<<(unsigned)sp.blocks, 256, 0, s>>>(xx, dy, n, h, w, workspace);
    else
      out1_wgrad<<<(unsigned)sp.blocks, 256, 0, s>>>(xx, dy, n, h, w, cin, workspace,
                                                      sp.per);
    SPS_TRY(launch_ok("sps_nn_conv3x3_out1_bwd (weight)"));
    SPS_TRY(add_parts(workspace, sp, dweight, s));
  }
  return SPS_OK;
}

int sps_nn_silu_bwd(const void* dy, const void* pre, void* dx, size_t count,
                    void* stream) {
  if (!dy || !pre || !dx) {
    set_error("sps_nn_silu_bwd: NULL pointer");
    return SPS_E_INVALID;
  }
  if (count % 8 == 0 && count / 8 < ((size_t)1 << 31) &&
      aligned16({dy, pre, dx})) {
    silu_bwd_v8<<<grid_for((int64_t)(count / 8)), 256, 0, (hipStream_t)stream>>>(
        (const uint16_t*)dy, (const uint16_t*)pre, (uint16_t*)dx, (unsigned)(count / 8));
    return launch_ok("sps_nn_silu_bwd");
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
  const int64_t pieces = (int64_t)n * (h / 2) * (w / 2) * (c / 8);
  if (c % 8 == 0 && pieces < ((int64_t)1 << 31) && aligned16({x, y})) {
    pool2_v8<<<grid_for(pieces), 256, 0, (hipStream_t)stream>>>(
        (const uint16_t*)x, h, w, c, mode, (unsigned)pieces, (uint16_t*)y);
    return launch_ok("sps_nn_pool2");
  }
  pool2_k<<<grid_for((int64_t)n * h * w * c / 4), 256, 0, (hipStream_t)stream>>>(
      (const uint16_t*)x, n, h, w, c, mode, (uint16_t*)y);
  return launch_ok("sps_nn_pool2");
}

int sps_nn_center(const float* x, float* y, int n, int64_t len,
                  void* stream) {
  if (!x || !y || n < 1 || len < 1) {
    set_error("sps_nn_center: NULL pointer or empty samples");
    return SPS_E_INVALID;
  }
  center_k<<<(unsigned)n, 256, 0, (hipStream_t)stream>>>(x, y, len);
  return launch_ok("sps_nn_center");
}

int sps_nn_unpool2(const void* x, int n, int h, int w, int c, float scale,
                   void* y, void* stream) {
  if (!x || !y || c < 1) {
    set_error("sps_nn_unpool2: invalid arguments");
    return SPS_E_INVALID;
  }
  const int64_t pieces = (int64_t)n * (2 * h) * (2 * w) * (c / 8);
  if (c % 8 == 0 && pieces < ((int64_t)1 << 31) && aligned16({x, y})) {
    unpool2_v8<<<grid_for(pieces), 256, 0, (hipStream_t)stream>>>(
        (const uint16_t*)x, h, w, c, scale, (unsigned)pieces, (uint16_t*)y);
    return launch_ok("sps_nn_unpool2");
  }
  unpool2_k<<<grid_for((int64_t)n * h * w * c * 4), 256, 0, (hipStream_t)stream>>>(
      (const uint16_t*)x, n, h, w, c, scale, (uint16_t*)y);
  return launch_ok("sps_nn_unpool2");
}

}  // extern "C"
