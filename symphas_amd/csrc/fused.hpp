// fused.hpp - launchers of the fused stencil kernels and the reductions.
#pragma once

#include "common.hpp"

namespace sps {

// 2D: K steps of the canonical form per launch; X must be a multiple of
// 16/sizeof(T).  TY = rows per streamed chunk.
template <typename T>
int launch_fused2d(const T* in, T* out, int X, int Y, int K, int TY,
                   const FusedCoef<T>& c, hipStream_t s);

// 3D: one step per launch; TZ = planes per streamed chunk; rows = y rows per
// wave of the register-streaming kernel (8 or 16), any other value selects
// the LDS plane-ring kernel.
template <typename T>
int launch_fused3d(const T* in, T* out, int X, int Y, int Z, int rows, int TZ,
                   const FusedCoef<T>& c, hipStream_t s);

// FP64 mass and free energy of an unpadded periodic field.  partials holds
// 2*kReduceBlocks doubles; sums receives {mass, energy}.
constexpr int kReduceBlocks = 1024;
template <typename T>
int launch_reduce(const T* f, int dim, const int* n, double h,
                  double* partials, double* sums, hipStream_t s);

}  // namespace sps
