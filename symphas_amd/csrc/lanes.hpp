// lanes.hpp - wave64 neighbour exchange through DPP, small vector loads, FMA.
#pragma once

#include <hip/hip_runtime.h>

namespace sps {

// DPP controls of the GFX9 wave-wide shifts.
//   lane i receives from lane i-1 (lane 0 receives 0):   wave_shr:1
//   lane i receives from lane i+1 (lane 63 receives 0):  wave_shl:1
constexpr int kDppWaveShl1 = 0x130;
constexpr int kDppWaveShr1 = 0x138;

// bound_ctrl: a lane whose source lies outside the wave reads 0, so no "old"
// value has to be set up before the move
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL,
                                         0xF, 0xF, true));
}

template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL,
                                             0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF,
                                             0xF, true);
  return __builtin_bit_cast(
      double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

// Value held by the previous / next lane (0 at the wave's ends).
template <typename T>
__device__ __forceinline__ T from_prev_lane(T v) {
  return dpp<kDppWaveShr1>(v);
}
template <typename T>
__device__ __forceinline__ T from_next_lane(T v) {
  return dpp<kDppWaveShl1>(v);
}

__device__ __forceinline__ float fmad(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
__device__ __forceinline__ double fmad(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}

// Horner evaluation of sum_k c[k] x^k with a compile-time degree bound.
template <int DEG, typename T>
__device__ __forceinline__ T horner(const T* c, T x) {
  if constexpr (DEG < 0) {
    return T(0);
  } else {
    T r = c[DEG];
#pragma unroll
    for (int k = DEG - 1; k >= 0; --k) r = fmad(r, x, c[k]);
    return r;
  }
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// Load / store VEC consecutive values; 16-byte vector accesses where the
// width allows (callers guarantee the alignment).
template <int VEC>
__device__ __forceinline__ void load_vec(const float* p, float (&v)[VEC]) {
  if constexpr (VEC % 4 == 0) {
#pragma unroll
    for (int q = 0; q < VEC / 4; ++q) {
      f32x4 t = *reinterpret_cast<const f32x4*>(p + 4 * q);
      v[4 * q] = t.x; v[4 * q + 1] = t.y; v[4 * q + 2] = t.z; v[4 * q + 3] = t.w;
    }
  } else if constexpr (VEC % 2 == 0) {
#pragma unroll
    for (int q = 0; q < VEC / 2; ++q) {
      f32x2 t = *reinterpret_cast<const f32x2*>(p + 2 * q);
      v[2 * q] = t.x; v[2 * q + 1] = t.y;
    }
  } else {
#pragma unroll
    for (int q = 0; q < VEC; ++q) v[q] = p[q];
  }
}

template <int VEC>
__device__ __forceinline__ void load_vec(const double* p, double (&v)[VEC]) {
  if constexpr (VEC % 2 == 0) {
#pragma unroll
    for (int q = 0; q < VEC / 2; ++q) {
      f64x2 t = *reinterpret_cast<const f64x2*>(p + 2 * q);
      v[2 * q] = t.x; v[2 * q + 1] = t.y;
    }
  } else {
#pragma unroll
    for (int q = 0; q < VEC; ++q) v[q] = p[q];
  }
}

template <int VEC>
__device__ __forceinline__ void store_vec(float* p, const float (&v)[VEC]) {
  if constexpr (VEC % 4 == 0) {
#pragma unroll
    for (int q = 0; q < VEC / 4; ++q) {
      f32x4 t = {v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]};
      *reinterpret_cast<f32x4*>(p + 4 * q) = t;
    }
  } else if constexpr (VEC % 2 == 0) {
#pragma unroll
    for (int q = 0; q < VEC / 2; ++q) {
      f32x2 t = {v[2 * q], v[2 * q + 1]};
      *reinterpret_cast<f32x2*>(p + 2 * q) = t;
    }
  } else {
#pragma unroll
    for (int q = 0; q < VEC; ++q) p[q] = v[q];
  }
}

template <int VEC>
__device__ __forceinline__ void store_vec(double* p, const double (&v)[VEC]) {
  if constexpr (VEC % 2 == 0) {
#pragma unroll
    for (int q = 0; q < VEC / 2; ++q) {
      f64x2 t = {v[2 * q], v[2 * q + 1]};
      *reinterpret_cast<f64x2*>(p + 2 * q) = t;
    }
  } else {
#pragma unroll
    for (int q = 0; q < VEC; ++q) p[q] = v[q];
  }
}

// Raw buffer access (GFX9 descriptor: stride 0, 32-bit data format) to
// `bytes` bytes from `base`: lane byte offset `off` plus a wave-uniform
// offset `soff`.  A lane whose offset lies past the range reads 0 and its
// store is dropped by the hardware, which masks lanes without a branch.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base,
                                                             int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0,
                                           bytes, 0x00020000);
}

template <typename T, int VEC>
__device__ __forceinline__ void buf_load(__amdgpu_buffer_rsrc_t r, unsigned off,
                                         int soff, T (&v)[VEC]) {
  constexpr int B = VEC * (int)sizeof(T);
  static_assert(B == 4 || B == 8 || B == 16, "4, 8 or 16 bytes per lane");
  if constexpr (B == 4) {
    const auto x = __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0);
    __builtin_memcpy(v, &x, B);
  } else if constexpr (B == 8) {
    const auto x = __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0);
    __builtin_memcpy(v, &x, B);
  } else {
    const auto x = __builtin_amdgcn_raw_buffer_load_b128(r, off, soff, 0);
    __builtin_memcpy(v, &x, B);
  }
}

template <typename T, int VEC>
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r,
                                          unsigned off, int soff,
                                          const T (&v)[VEC]) {
  constexpr int B = VEC * (int)sizeof(T);
  static_assert(B == 4 || B == 8 || B == 16, "4, 8 or 16 bytes per lane");
  if constexpr (B == 4) {
    unsigned x;
    __builtin_memcpy(&x, v, B);
    __builtin_amdgcn_raw_buffer_store_b32(x, r, off, soff, 0);
  } else if constexpr (B == 8) {
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    u32x2 x;
    __builtin_memcpy(&x, v, B);
    __builtin_amdgcn_raw_buffer_store_b64(x, r, off, soff, 0);
  } else {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    u32x4 x;
    __builtin_memcpy(&x, v, B);
    __builtin_amdgcn_raw_buffer_store_b128(x, r, off, soff, 0);
  }
}

// Spread consecutive workgroup ids (which the dispatcher deals round-robin to
// the 8 XCDs) so that each XCD receives a contiguous range of tiles and the
// halo rows neighbouring tiles share stay in one L2.  XCD x receives blocks
// x, x+8, ...: nb/8 of them, one more when x < nb%8; the map is a bijection
// of [0, nb) for any nb.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int x = b & 7, q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

}  // namespace sps
