// reduce.hip - FP64 mass and free-energy sums of a periodic field.
//
// Pass 1: a fixed grid of at most kReduceBlocks workgroups walks the rows
// (y, z) in a block-strided order; every site contributes psi (mass) and the
// Landau density of energy.hpp (energy), accumulated in FP64 even for FP32
// fields.  Waves reduce with xor shuffles, the four waves of a block through
// LDS.  Pass 2: one workgroup adds the partials in a fixed order and applies
// the functional's scale (h^D, or 1/sites for the mean).  The result is
// deterministic for a given field and size.
#include <mutex>

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

// ---------------------------------------------------- allocation-free sums
// sps_fd_reduce_device runs the two passes above with their partials in a
// file-scope device array instead of plan memory, so it calls no allocator.
// Pass 1 never uses more than kReduceBlocks workgroups, so this fixed 16 KiB
// holds the partials of a field of any size.  Each GPU holds its own copy of
// the array; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (the array's name is a
// host-side stand-in, not a device address).
namespace sps {
namespace {

__device__ __align__(256) double scratch[2 * kReduceBlocks];

constexpr int kScratchDevices = 64;
std::mutex scratch_mu;
double* scratch_dev[kScratchDevices] = {};

// Address of the current device's scratch array, looked up once per device.
int scratch_address(double** out) {
  int dev = -1;
  SPS_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices) {
    set_error("device %d beyond the %d devices the scratch table covers", dev,
              kScratchDevices);
    return SPS_E_HIP;
  }
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error("hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                hipGetErrorString(e));
      return SPS_E_HIP;
    }
    scratch_dev[dev] = static_cast<double*>(ptr);
  }
  *out = scratch_dev[dev];
  return SPS_OK;
}

}  // namespace
}  // namespace sps

extern "C" int sps_fd_reduce_device(int dim, int dtype, const int32_t* n,
                                    double h, const sps_fd_energy* energy,
                                    const void* field, double* sums,
                                    void* hip_stream) {
  using namespace sps;
  // every argument is checked before the first HIP call
  if (dim != 2 && dim != 3) {
    set_error("dim must be 2 or 3 (got %d)", dim);
    return SPS_E_INVALID;
  }
  if (dtype != SPS_F32 && dtype != SPS_F64) {
    set_error("dtype must be SPS_F32 or SPS_F64 (got %d)", dtype);
    return SPS_E_INVALID;
  }
  if (!n || !field || !sums) {
    set_error("n, field and sums must not be NULL");
    return SPS_E_INVALID;
  }
  double sites = 1.0;
  for (int a = 0; a < dim; ++a) {
    if (n[a] < 1 || n[a] > (1 << 20)) {
      set_error("n[%d] = %d out of range [1, 2^20]", a, n[a]);
      return SPS_E_INVALID;
    }
    sites *= n[a];
  }
  if (sites > (double)SPS_REDUCE_DEVICE_MAX_SITES) {
    set_error("field of %.0f sites exceeds the maximum of %d sites", sites,
              SPS_REDUCE_DEVICE_MAX_SITES);
    return SPS_E_INVALID;
  }
  if (!(h > 0) || !std::isfinite(h)) {
    set_error("h must be positive and finite");
    return SPS_E_INVALID;
  }
  EnergyParams ep;
  SPS_TRY(make_energy_params(energy, h, ep));
  double* partials = nullptr;
  SPS_TRY(scratch_address(&partials));
  const int nn[3] = {n[0], n[1], dim == 3 ? n[2] : 1};
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  return dtype == SPS_F32
             ? launch_reduce<float>(static_cast<const float*>(field), dim, nn,
                                    h, ep, partials, sums, s)
             : launch_reduce<double>(static_cast<const double*>(field), dim,
                                     nn, h, ep, partials, sums, s);
}
