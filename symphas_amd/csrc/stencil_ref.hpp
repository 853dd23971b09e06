// stencil_ref.hpp - Stencil2d2h / Stencil3d2h point formulas for the
// reference-order kernels.
//
// Every formula keeps the grouping and operand order of the reference
// (sol/inc/stencilh2.h:87-507), so that an FP64 evaluation compiled without
// fused multiply-add rounds exactly as the CPU SolverFT does.  `v(i, j, k)` is
// the operand at offset i*s0 + j*s1 + k*s2, where (s0, s1, s2) are the
// direction vectors of get_stride<axis> (datatypes/inc/grid.h:537-578);
// the accessor resolves them, including the periodic wrap.
#pragma once

#include <hip/hip_runtime.h>

namespace sps::ref {

#define V2(i, j) v(i, j, 0)
#define V3(i, j, k) v(i, j, k)

// ---------------------------------------------------------------- 2D

template <typename T, class A>
__device__ __forceinline__ T lap2_5(const A& v, T divh2) {
  return (V2(-1, 0) + V2(1, 0) + V2(0, 1) + V2(0, -1) - T(4) * V2(0, 0)) *
         divh2;
}

template <typename T, class A>
__device__ __forceinline__ T lap2_9(const A& v, T divh2) {
  return (V2(-1, 1) + V2(1, -1) + V2(1, 1) + V2(-1, -1) + T(4) * V2(1, 0) +
          T(4) * V2(0, 1) + T(4) * V2(0, -1) + T(4) * V2(-1, 0) -
          T(20) * V2(0, 0)) *
         divh2 * T(1.0 / 6);
}

template <typename T, class A>
__device__ __forceinline__ T bilap2_13(const A& v, T divh4) {
  return (V2(0, -2) + V2(0, 2) + V2(-2, 0) + V2(2, 0) +
          T(2) * (V2(-1, -1) + V2(1, -1) + V2(-1, 1) + V2(1, 1)) -
          T(8) * (V2(-1, 0) + V2(1, 0) + V2(0, -1) + V2(0, 1)) +
          T(20) * V2(0, 0)) *
         divh4;
}

template <typename T, class A>
__device__ __forceinline__ T bilap2_17(const A& v, T divh4) {
  return ((V2(2, 2) + V2(-2, 2) + V2(2, -2) + V2(-2, -2)) +
          T(10) * (V2(2, 0) + V2(-2, 0) + V2(0, 2) + V2(0, -2)) +
          T(8) * (V2(1, 1) + V2(-1, 1) + V2(1, -1) + V2(-1, -1)) -
          T(64) * (V2(1, 0) + V2(-1, 0) + V2(0, 1) + V2(0, -1)) +
          T(180) * V2(0, 0)) *
         divh4 * T(1.0 / 12);
}

template <typename T, class A>
__device__ __forceinline__ T bilap2_21(const A& v, T divh4) {
  return ((V2(2, 1) + V2(-2, 1) + V2(2, -1) + V2(-2, -1) + V2(1, 2) +
           V2(-1, 2) + V2(1, -2) + V2(-1, -2)) +
          (V2(2, 0) + V2(-2, 0) + V2(0, 2) + V2(0, -2)) -
          T(2) * (V2(1, 1) + V2(-1, 1) + V2(1, -1) + V2(-1, -1)) -
          T(10) * (V2(1, 0) + V2(-1, 0) + V2(0, 1) + V2(0, -1)) +
          T(36) * V2(0, 0)) *
         divh4 * T(1.0 / 12);
}

template <typename T, class A>
__device__ __forceinline__ T gradlap2_6(const A& v, T divh3, int component) {
  if (component == 0)
    return ((V2(2, 1) - V2(-2, 1) + V2(2, -1) - V2(-2, -1)) -
            T(4) * (V2(1, 0) - V2(-1, 0))) *
           divh3 * T(1.0 / 4);
  return ((V2(1, 2) - V2(1, -2) + V2(-1, 2) - V2(-1, -2)) -
          T(4) * (V2(0, 1) - V2(0, -1))) *
         divh3 * T(1.0 / 4);
}

template <typename T, class A>
__device__ __forceinline__ T gradlap2_8(const A& v, T divh3) {
  return ((V2(2, 0) - V2(-2, 0)) +
          (V2(1, 1) - V2(-1, 1) + V2(1, -1) - V2(-1, -1)) -
          T(4) * (V2(1, 0) - V2(-1, 0))) *
         divh3 * T(1.0 / 2);
}

template <typename T, class A>
__device__ __forceinline__ T gradlap2_12(const A& v, T divh3) {
  return ((V2(2, 2) - V2(-2, 2) + V2(2, -2) - V2(-2, -2)) +
          T(10) * (V2(2, 0) - V2(-2, 0)) +
          T(4) * (V2(1, 1) - V2(-1, 1) + V2(1, -1) - V2(-1, -1)) -
          T(32) * (V2(1, 0) - V2(-1, 0))) *
         divh3 * T(1.0 / 24);
}

template <typename T, class A>
__device__ __forceinline__ T gradlap2_16(const A& v, T divh3) {
  return (T(2) * (V2(2, 1) - V2(-2, 1) + V2(2, -1) - V2(-2, -1)) +
          T(2) * (V2(2, 0) - V2(-2, 0)) +
          (V2(1, 2) - V2(-1, 2) + V2(1, -2) - V2(-1, -2)) -
          T(2) * (V2(1, 1) - V2(-1, 1) + V2(1, -1) - V2(-1, -1)) -
          T(10) * (V2(1, 0) - V2(-1, 0))) *
         divh3 * T(1.0 / 12);
}

// ---------------------------------------------------------------- 3D
// Neighbour groups, each in the order the reference sums it.

#define AXIAL_A                                                          \
  (V3(-1, 0, 0) + V3(1, 0, 0) + V3(0, 1, 0) + V3(0, -1, 0) + V3(0, 0, 1) + \
   V3(0, 0, -1))
#define AXIAL_B                                                          \
  (V3(1, 0, 0) + V3(-1, 0, 0) + V3(0, 1, 0) + V3(0, -1, 0) + V3(0, 0, 1) + \
   V3(0, 0, -1))
#define AXIAL2                                                           \
  (V3(2, 0, 0) + V3(-2, 0, 0) + V3(0, 2, 0) + V3(0, -2, 0) + V3(0, 0, 2) + \
   V3(0, 0, -2))
#define CORNERS_A                                                      \
  (V3(-1, -1, -1) + V3(1, -1, -1) + V3(-1, 1, -1) + V3(1, 1, -1) +     \
   V3(-1, -1, 1) + V3(1, -1, 1) + V3(-1, 1, 1) + V3(1, 1, 1))
#define CORNERS_B                                                      \
  (V3(1, 1, 1) + V3(-1, 1, 1) + V3(1, -1, 1) + V3(-1, -1, 1) +         \
   V3(1, 1, -1) + V3(-1, 1, -1) + V3(1, -1, -1) + V3(-1, -1, -1))
#define EDGES_A                                                            \
  (V3(-1, -1, 0) + V3(1, -1, 0) + V3(-1, 1, 0) + V3(1, 1, 0) +             \
   V3(-1, 0, -1) + V3(1, 0, -1) + V3(-1, 0, 1) + V3(1, 0, 1) +             \
   V3(0, -1, -1) + V3(0, 1, -1) + V3(0, -1, 1) + V3(0, 1, 1))
#define EDGES_B                                                            \
  (V3(1, 1, 0) + V3(-1, 1, 0) + V3(1, -1, 0) + V3(-1, -1, 0) +             \
   V3(1, 0, -1) + V3(-1, 0, -1) + V3(1, 0, 1) + V3(-1, 0, 1) +             \
   V3(0, 1, -1) + V3(0, 1, 1) + V3(0, -1, -1) + V3(0, -1, 1))
#define CORNERS2                                                       \
  (V3(2, 2, 2) + V3(-2, 2, 2) + V3(2, -2, 2) + V3(-2, -2, 2) +         \
   V3(2, 2, -2) + V3(-2, 2, -2) + V3(2, -2, -2) + V3(-2, -2, -2))
#define ODD1 (V3(1, 0, 0) - V3(-1, 0, 0))
#define ODD2 (V3(2, 0, 0) - V3(-2, 0, 0))
#define ODD_EDGES                                                          \
  (V3(1, 1, 0) - V3(-1, 1, 0) + V3(1, -1, 0) - V3(-1, -1, 0) +             \
   V3(1, 0, -1) - V3(-1, 0, -1) + V3(1, 0, 1) - V3(-1, 0, 1))
#define ODD_CORNERS                                                        \
  (V3(1, 1, 1) - V3(-1, 1, 1) + V3(1, -1, 1) - V3(-1, -1, 1) +             \
   V3(1, 1, -1) - V3(-1, 1, -1) + V3(1, -1, -1) - V3(-1, -1, -1))

template <typename T, class A>
__device__ __forceinline__ T lap3_7(const A& v, T divh2) {
  return (V3(-1, 0, 0) + V3(1, 0, 0) + V3(0, 1, 0) + V3(0, -1, 0) +
          V3(0, 0, 1) + V3(0, 0, -1) - T(6) * V3(0, 0, 0)) *
         divh2;
}

template <typename T, class A>
__device__ __forceinline__ T lap3_15(const A& v, T divh2) {
  return (V3(-1, -1, -1) + V3(1, -1, -1) + V3(-1, 1, -1) + V3(1, 1, -1) +
          V3(-1, -1, 1) + V3(1, -1, 1) + V3(-1, 1, 1) + V3(1, 1, 1) +
          T(8) * (V3(-1, 0, 0) + V3(1, 0, 0) + V3(0, 1, 0) + V3(0, -1, 0) +
                  V3(0, 0, 1) + V3(0, 0, -1) - T(7) * V3(0, 0, 0))) *
         divh2 * T(1.0 / 12);
}

// The reference's 19-point laplacian sums "vz + vz" (stencilh2.h:273): the
// -z axial neighbour is replaced by a second +z neighbour.  Reproduced.
template <typename T, class A>
__device__ __forceinline__ T lap3_19(const A& v, T divh2) {
  return (V3(-1, -1, 0) + V3(1, -1, 0) + V3(-1, 1, 0) + V3(1, 1, 0) +
          V3(-1, 0, -1) + V3(1, 0, -1) + V3(-1, 0, 1) + V3(1, 0, 1) +
          V3(0, -1, -1) + V3(0, 1, -1) + V3(0, -1, 1) + V3(0, 1, 1) +
          T(2) * (V3(-1, 0, 0) + V3(1, 0, 0) + V3(0, 1, 0) + V3(0, -1, 0) +
                  V3(0, 0, 1) + V3(0, 0, 1)) -
          T(24) * V3(0, 0, 0)) *
         divh2 * T(1.0 / 6);
}

template <typename T, class A>
__device__ __forceinline__ T lap3_21(const A& v, T divh2) {
  return (-CORNERS_A + T(4) * EDGES_A + T(-40) * V3(0, 0, 0)) * divh2 *
         T(1.0 / 12);
}

template <typename T, class A>
__device__ __forceinline__ T lap3_27(const A& v, T divh2) {
  return (CORNERS_A + T(3) * EDGES_A + T(14) * AXIAL_A -
          T(128) * V3(0, 0, 0)) *
         divh2 * T(1.0 / 30);
}

template <typename T, class A>
__device__ __forceinline__ T bilap3_21(const A& v, T divh4) {
  return ((V3(2, 2, 0) + V3(-2, 2, 0) + V3(2, -2, 0) + V3(-2, -2, 0) +
           V3(2, 0, -2) + V3(-2, 0, -2) + V3(2, 0, 2) + V3(-2, 0, 2) +
           V3(0, 2, -2) + V3(0, 2, 2) + V3(0, -2, -2) + V3(0, -2, 2)) -
          T(4) * CORNERS_B + T(20) * V3(0, 0, 0)) *
         divh4 * T(1.0 / 4);
}

template <typename T, class A>
__device__ __forceinline__ T bilap3_25(const A& v, T divh4) {
  return (T(2) * EDGES_B + AXIAL2 - T(12) * AXIAL_B + T(42) * V3(0, 0, 0)) *
         divh4;
}

template <typename T, class A>
__device__ __forceinline__ T bilap3_41(const A& v, T divh4) {
  return (CORNERS2 - T(40) * CORNERS_B + T(96) * EDGES_B + T(20) * AXIAL2 -
          T(320) * AXIAL_B + T(960) * V3(0, 0, 0)) *
         divh4 * T(1.0 / 24);
}

template <typename T, class A>
__device__ __forceinline__ T bilap3_52(const A& v, T divh4) {
  return (-CORNERS2 +
          T(10) * (V3(1, 1, 2) + V3(-1, 1, 2) + V3(1, -1, 2) + V3(-1, -1, 2) +
                   V3(1, 1, -2) + V3(-1, 1, -2) + V3(1, -1, -2) +
                   V3(-1, -1, -2) + V3(1, 2, 1) + V3(-1, 2, 1) +
                   V3(1, -2, 1) + V3(-1, -2, 1) + V3(1, 2, -1) +
                   V3(-1, 2, -1) + V3(1, -2, -1) + V3(-1, -2, -1) +
                   V3(2, 1, 1) + V3(-2, 1, 1) + V3(2, -1, 1) +
                   V3(-2, -1, 1) + V3(2, 1, -1) + V3(-2, 1, -1) +
                   V3(2, -1, -1) + V3(-2, -1, -1)) -
          T(20) * CORNERS_B - T(36) * EDGES_B + T(360) * V3(0, 0, 0)) *
         divh4 * T(1.0 / 36);
}

template <typename T, class A>
__device__ __forceinline__ T bilap3_57(const A& v, T divh4) {
  return ((V3(0, 1, 2) + V3(0, -1, 2) + V3(0, 1, -2) + V3(0, -1, -2) +
           V3(0, 2, 1) + V3(0, -2, 1) + V3(0, 2, -1) + V3(0, -2, -1) +
           V3(2, 0, 1) + V3(-2, 0, 1) + V3(2, 0, -1) + V3(-2, 0, -1) +
           V3(1, 0, 2) + V3(-1, 0, 2) + V3(1, 0, -2) + V3(-1, 0, -2) +
           V3(1, 2, 0) + V3(-1, 2, 0) + V3(1, -2, 0) + V3(-1, -2, 0) +
           V3(2, 1, 0) + V3(-2, 1, 0) + V3(2, -1, 0) + V3(-2, -1, 0)) +
          T(3) * CORNERS_B - T(8) * EDGES_B - AXIAL2 + T(4) * AXIAL_B +
          T(30) * V3(0, 0, 0)) *
         divh4 * T(1.0 / 3);
}

template <typename T, class A>
__device__ __forceinline__ T gradlap3_10(const A& v, T divh3) {
  return ((V3(2, 1, 1) - V3(-2, 1, 1) + V3(2, -1, 1) - V3(-2, -1, 1) +
           V3(2, 1, -1) - V3(-2, 1, -1) + V3(2, -1, -1) - V3(-2, -1, -1)) -
          T(8) * ODD1) *
         divh3 * T(1.0 / 8);
}

template <typename T, class A>
__device__ __forceinline__ T gradlap3_12(const A& v, T divh3) {
  return (ODD2 + ODD_EDGES - T(6) * ODD1) * divh3 * T(1.0 / 2);
}

template <typename T, class A>
__device__ __forceinline__ T gradlap3_28(const A& v, T divh3) {
  return ((V3(2, 2, 0) - V3(-2, 2, 0) + V3(2, -2, 0) - V3(-2, -2, 0) +
           V3(2, 0, -2) - V3(-2, 0, -2) + V3(2, 0, 2) - V3(-2, 0, 2)) +
          T(8) * ODD2 + T(6) * ODD_CORNERS - T(8) * ODD_EDGES -
          T(16) * ODD1) *
         divh3 * T(1.0 / 24);
}

template <typename T, class A>
__device__ __forceinline__ T gradlap3_36(const A& v, T divh3) {
  return (T(2) * (V3(2, 0, 1) - V3(-2, 0, 1) + V3(2, 0, -1) - V3(-2, 0, -1) +
                  V3(2, 1, 0) - V3(-2, 1, 0) + V3(2, -1, 0) - V3(-2, -1, 0)) -
          T(2) * ODD2 +
          (V3(1, 0, 2) - V3(-1, 0, 2) + V3(1, 0, -2) - V3(-1, 0, -2) +
           V3(1, 2, 0) - V3(-1, 2, 0) + V3(1, -2, 0) - V3(-1, -2, 0)) +
          T(3) * ODD_CORNERS - T(8) * ODD_EDGES + T(4) * ODD1) *
         divh3 * T(1.0 / 12);
}

template <typename T, class A>
__device__ __forceinline__ T gradlap3_40(const A& v, T divh3) {
  return (T(16) * (V3(2, 0, 1) - V3(-2, 0, 1) + V3(2, 0, -1) -
                   V3(-2, 0, -1) + V3(2, 1, 0) - V3(-2, 1, 0) +
                   V3(2, -1, 0) - V3(-2, -1, 0)) -
          T(16) * ODD2 +
          T(3) * (V3(1, 2, 1) - V3(-1, 2, 1) + V3(1, -2, 1) - V3(-1, -2, 1) +
                  V3(1, 2, -1) - V3(-1, 2, -1) + V3(1, -2, -1) -
                  V3(-1, -2, -1) + V3(1, 1, 2) - V3(-1, 1, 2) +
                  V3(1, -1, 2) - V3(-1, -1, 2) + V3(1, 1, -2) -
                  V3(-1, 1, -2) + V3(1, -1, -2) - V3(-1, -1, -2)) +
          T(2) * (V3(1, 0, 2) - V3(-1, 0, 2) + V3(1, 0, -2) - V3(-1, 0, -2) +
                  V3(1, 2, 0) - V3(-1, 2, 0) + V3(1, -2, 0) - V3(-1, -2, 0)) -
          T(22) * ODD_EDGES - T(40) * ODD1) *
         divh3 * T(1.0 / 96);
}

#undef AXIAL_A
#undef AXIAL_B
#undef AXIAL2
#undef CORNERS_A
#undef CORNERS_B
#undef EDGES_A
#undef EDGES_B
#undef CORNERS2
#undef ODD1
#undef ODD2
#undef ODD_EDGES
#undef ODD_CORNERS
#undef V2
#undef V3

}  // namespace sps::ref
