// reduce.hip - FP64 mass and free-energy sums of a periodic field.
//
// Pass 1: a fixed grid of at most kReduceBlocks workgroups walks the rows
// (y, z) in a block-strided order; every site contributes psi (mass) and the
// Landau density of energy.hpp (energy), accumulated in FP64 even for FP32
// fields.  Waves reduce with xor shuffles, the four waves of a block through
// LDS.  Pass 2: one workgroup adds the partials in a fixed order and applies
// the functional's scale (h^D, or 1/sites for the mean).  The result is
// deterministic for a given field and size.
#include "energy.hpp"
#include "fused.hpp"

namespace sps {
namespace {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename T>
__global__ __launch_bounds__(256) void reduce_pass1(const T* __restrict__ f,
                                                    int dim, int X, int Y,
                                                    int Z, EnergyParams ep,
                                                    double* partials) {
  double m = 0.0, e = 0.0;
  const int rows = Y * Z;
  const int64_t XY = (int64_t)X * Y;
  const bool central = ep.central != 0;
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    const int y = r % Y, z = r / Y;
    const int yn = y + 1 == Y ? 0 : y + 1;
    const int zn = z + 1 == Z ? 0 : z + 1;
    // the backward neighbour is the site itself for the forward difference
    const int yp = central ? (y == 0 ? Y - 1 : y - 1) : y;
    const int zp = central ? (z == 0 ? Z - 1 : z - 1) : z;
    const T* row = f + (int64_t)z * XY + (int64_t)y * X;
    const T* rowyn = f + (int64_t)z * XY + (int64_t)yn * X;
    const T* rowyp = f + (int64_t)z * XY + (int64_t)yp * X;
    const T* rowzn = f + (int64_t)zn * XY + (int64_t)y * X;
    const T* rowzp = f + (int64_t)zp * XY + (int64_t)y * X;
    for (int x = threadIdx.x; x < X; x += blockDim.x) {
      const int xn = x + 1 == X ? 0 : x + 1;
      const int xp = central ? (x == 0 ? X - 1 : x - 1) : x;
      const double p = (double)row[x];
      const double gx = (double)row[xn] - (double)row[xp];
      const double gy = (double)rowyn[x] - (double)rowyp[x];
      double g2 = gx * gx + gy * gy;
      if (dim == 3) {
        const double gz = (double)rowzn[x] - (double)rowzp[x];
        g2 += gz * gz;
      }
      m += p;
      e += landau_density(ep, p, g2);
    }
  }
  __shared__ double sh[2][4];
  m = wave_sum(m);
  e = wave_sum(e);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[0][w] = m;
    sh[1][w] = e;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
    partials[2 * blockIdx.x + 1] =
        (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
  }
}

__global__ __launch_bounds__(256) void reduce_pass2(const double* partials,
                                                    int nb, double escale,
                                                    double* sums) {
  double m = 0.0, e = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) {
    m += partials[2 * i];
    e += partials[2 * i + 1];
  }
  __shared__ double sh[2][4];
  m = wave_sum(m);
  e = wave_sum(e);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[0][w] = m;
    sh[1][w] = e;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    sums[0] = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
    sums[1] = ((sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3])) * escale;
  }
}

}  // namespace

template <typename T>
int launch_reduce(const T* f, int dim, const int* n, double h,
                  const EnergyParams& ep, double* partials, double* sums,
                  hipStream_t s) {
  const int X = n[0], Y = n[1], Z = dim == 3 ? n[2] : 1;
  const int rows = Y * Z;
  const int nb = rows < kReduceBlocks ? rows : kReduceBlocks;
  reduce_pass1<T><<<nb, 256, 0, s>>>(f, dim, X, Y, Z, ep, partials);
  const double sites = (double)X * Y * Z;
  reduce_pass2<<<1, 256, 0, s>>>(partials, nb, energy_scale(ep, dim, h, sites),
                                 sums);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("reduction launch failed: %s", hipGetErrorString(e));
    return SPS_E_HIP;
  }
  return SPS_OK;
}

template int launch_reduce<float>(const float*, int, const int*, double,
                                  const EnergyParams&, double*, double*,
                                  hipStream_t);
template int launch_reduce<double>(const double*, int, const int*, double,
                                   const EnergyParams&, double*, double*,
                                   hipStream_t);

}  // namespace sps
