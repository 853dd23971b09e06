// fused.hpp - launchers of the fused stencil kernels and the reductions.
#pragma once

#include "common.hpp"

namespace sps {

// 2D: K steps of the canonical form per launch; TY = rows per streamed
// chunk; vec = values per lane, 16/sizeof(T) or 8/sizeof(T), dividing X.
template <typename T>
int launch_fused2d(const T* in, T* out, int X, int Y, int K, int TY, int vec,
                   const FusedCoef<T>& c, hipStream_t s);

// 3D: K steps per launch; rows = y rows per wave, TZ = planes per streamed
// chunk.  fused3d_shape_ok lists the (K, rows) pairs built for each element
// size (those whose register rings fit without spilling).
bool fused3d_shape_ok(int dtype_size, int K, int rows);
template <typename T>
int launch_fused3d(const T* in, T* out, int X, int Y, int Z, int K, int rows,
                   int TZ, const FusedCoef<T>& c, hipStream_t s);

// FP64 mass and free energy (functional ep) of an unpadded periodic field.
// partials holds 2*kReduceBlocks doubles; sums receives {mass, energy}.
constexpr int kReduceBlocks = 1024;
struct EnergyParams;
template <typename T>
int launch_reduce(const T* f, int dim, const int* n, double h,
                  const EnergyParams& ep, double* partials, double* sums,
                  hipStream_t s);

}  // namespace sps
