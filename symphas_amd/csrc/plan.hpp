// plan.hpp - the single-field plan behind sps_fd_plan*.
#pragma once

#include <algorithm>

#include "common.hpp"
#include "energy.hpp"
#include "fused.hpp"

namespace sps {
void set_error(const char* fmt, ...);
const char* last_error();
int validate_desc(const sps_fd_desc* d);
Lowered lower(const sps_fd_desc& d);
template <typename T>
void make_fused_coef(const sps_fd_desc& d, const Lowered& lw, FusedCoef<T>& c);
// Reference-order parameters of term group `group` (0: the right-hand side,
// k >= 1: the definition of var(k)).
template <typename T>
void make_ref_params(const sps_fd_desc& d, RefParams<T>& p, int group = 0);
void interior_to_padded(const sps_fd_desc& d, const char* src, char* dst,
                        size_t es, bool fill_ghosts);
void padded_to_interior(const sps_fd_desc& d, const char* src, char* dst,
                        size_t es);

// Launches replayed per captured graph (even, so buffer parity is kept).
constexpr int kGraphLaunches = 16;

struct GraphSlot {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};
}  // namespace sps

struct sps_fd_plan {
  sps_fd_desc d{};
  hipStream_t stream = nullptr;      // stream work is enqueued on
  hipStream_t own_stream = nullptr;  // the plan's own stream
  size_t nsites = 0, esize = 0;
  void* buf[2] = {nullptr, nullptr};  // ping-pong fields (unpadded)
  int cur = 0;                        // buf[cur] holds the current field
  void* work = nullptr;               // dframe / derivative output
  double* partials = nullptr;
  double* sums = nullptr;
  sps::EnergyParams energy{};  // functional of sps_fd_reduce
  bool has_data = false;
  int path = SPS_PATH_REFERENCE;
  sps::Lowered low;
  sps::RefParams<float> rp32;   // right-hand side
  sps::RefParams<double> rp64;
  void* vars[SPS_MAX_PROV] = {nullptr, nullptr};  // provisional grids
  sps::RefParams<float> prov32[SPS_MAX_PROV];     // their definitions
  sps::RefParams<double> prov64[SPS_MAX_PROV];
  sps::FusedCoef<float> fc32;
  sps::FusedCoef<double> fc64;
  int K = 1;    // fused: steps per launch
  int TY = 32;  // fused 2D: rows per chunk; fused 3D: y rows per wave
  int TZ = 64;  // fused 3D: planes per chunk
  int VEC = 4;  // fused 2D: values per lane (16 or 8 bytes)
  sps::GraphSlot graphs[2];  // indexed by the buffer parity at capture

  int launch_one(int k);
  // reference-order passes that fill the provisional grids from buf[cur]
  int launch_provisionals();
  int steps_per_launch() const;
  int step(int nsteps);
  void drop_graphs();

  // Launch shape before any tuning.  2D: the best of a sweep of steps per
  // launch {1..4} x rows per chunk x lane width on Model B 4096^2 (FP32:
  // 4 x 32 rows, 2 values per lane; FP64: 4 x 64).  3D: two steps per launch
  // (FP32: 12 rows per wave, two waves per SIMD at 254 VGPRs; FP64: 4 rows,
  // one wave per SIMD) with the z-chunk height that splits the grid into as
  // few rounds of resident waves as possible.
  void choose_default_shape() {
    if (d.dim == 3) {
      K = 2;
      TY = d.dtype == SPS_F32 ? 12 : 4;
      const int slots = d.dtype == SPS_F32 ? 2048 : 1024;  // resident waves
      const int halo = 2 * K;  // lanes per side at radius two
      const int64_t per_chunk = (int64_t)((d.n[0] + 63 - 2 * halo) / (64 - 2 * halo)) *
                                ((d.n[1] + TY - 1) / TY);
      const int64_t m = std::max<int64_t>(1, slots / per_chunk);
      TZ = (int)std::min<int64_t>(1024, std::max<int64_t>(16, (d.n[2] + m - 1) / m));
    } else if (d.dtype == SPS_F32) {
      K = 4;
      TY = 32;
    } else {
      K = 4;
      TY = 64;
    }
    // FP32 2D: 8-byte lanes (2 values) won the shape sweep; FP64: 16-byte
    VEC = (d.dim == 2 && d.dtype == SPS_F32) ? 8 / (int)esize : 16 / (int)esize;
    if (d.dim == 2 && d.n[0] % VEC != 0) VEC = 8 / (int)esize;
  }

  int allocate() {
    SPS_HIP_TRY(hipStreamCreateWithFlags(&own_stream, hipStreamNonBlocking));
    stream = own_stream;
    for (int i = 0; i < 2; ++i) {
      hipError_t e = hipMalloc(&buf[i], nsites * esize);
      if (e != hipSuccess) {
        buf[i] = nullptr;
        sps::set_error("hipMalloc of %zu bytes failed: %s", nsites * esize,
                       hipGetErrorString(e));
        return SPS_E_NOMEM;
      }
    }
    SPS_HIP_TRY(hipMalloc(&partials, 2 * sps::kReduceBlocks * sizeof(double)));
    SPS_HIP_TRY(hipMalloc(&sums, 2 * sizeof(double)));
    for (int k = 0; k < d.nprov; ++k) {
      hipError_t e = hipMalloc(&vars[k], nsites * esize);
      if (e != hipSuccess) {
        vars[k] = nullptr;
        sps::set_error("hipMalloc of a provisional grid failed: %s",
                       hipGetErrorString(e));
        return SPS_E_NOMEM;
      }
    }
    // every term list reads var(k) from vars[k - 1]
    for (int k = 1; k <= d.nprov; ++k) {
      rp32.vars[k] = (const float*)vars[k - 1];
      rp64.vars[k] = (const double*)vars[k - 1];
      for (int j = 0; j < d.nprov; ++j) {
        prov32[j].vars[k] = (const float*)vars[k - 1];
        prov64[j].vars[k] = (const double*)vars[k - 1];
      }
    }
    return SPS_OK;
  }
};
