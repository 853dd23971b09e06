// batch.hip - many independent 2D fields in one launch per step group (the
// batched counterpart of find_solution over an array of models,
// modules-io.h:368-474, which steps the models one after another).
//
// Every instance is stored unpadded in one device pool with two buffers
// (ping-pong shared by all instances).  A work item is one strip (64 lanes x 2
// values, halo lanes at both sides, as the single-field kernel) by a chunk of
// rows of one instance, so every instance runs the same register-light stream
// whatever its width; a persistent grid of resident waves takes the items
// with a static stride.  Per-instance coefficients
// (mobility, gradient coefficient, dt, h, model) live in a descriptor table
// read with scalar loads.  Captured once in a hipGraph per parity and
// replayed.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"
#include "energy.hpp"
#include "fused_tile.hpp"
#include "plan.hpp"

namespace sps {
namespace {

struct BatchInst {
  float* buf[2];
  int X, Y, pad, kind;  // kind: 0 radius 1 with A, 1 radius 2 without A,
                        //       2 radius 2 with A, 3 radius 1 without A
  double h, divh;
  FusedCoef<float> c;
};

struct BatchItem {
  int inst, y0, ty, strip;
};

// Strip geometry of the 2-values-per-lane stream (fused_tile.hpp): a wave
// covers (64 - 2*HL)*2 columns and HL halo lanes per side carry the K*R
// columns the K fused steps reach.
constexpr int kBatchVec = 2;
__host__ __device__ constexpr int batch_halo_lanes(int K, int R) {
  return (K * R + kBatchVec - 1) / kBatchVec;
}
__host__ __device__ constexpr int batch_strip_width(int K, int R) {
  return (64 - 2 * batch_halo_lanes(K, R)) * kBatchVec;
}

template <int K, int R, bool HASA>
__device__ __forceinline__ void run_strip(const BatchInst& I, int parity,
                                          const BatchItem& it, int lane) {
  constexpr int HL = batch_halo_lanes(K, R);
  constexpr int W = batch_strip_width(K, R);
  // the descriptor is read once into registers: the stream's stores could
  // alias it, which would otherwise force a reload per row
  const FusedCoef<float> c = I.c;
  const int X = I.X, Y = I.Y;
  const float* in = I.buf[parity];
  float* out = I.buf[1 - parity];
  const int cb = it.strip * W + (lane - HL) * kBatchVec;
  int col = cb % X;
  if (col < 0) col += X;
  const bool store_lane = lane >= HL && lane < 64 - HL && cb < X;
  stream_tile<float, kBatchVec, K, R, HASA, 3>(in, out, X, Y, it.y0, it.ty, col,
                                               store_lane, c);
}

// Persistent: a grid of resident waves walks the item list (strip x row
// chunk, near-uniform in cost) with a static stride, so no counter or atomic
// is needed; consecutive blocks go to one XCD.  Every wave's loop bound is
// nitems, which it reaches.
template <int K>
__global__ __launch_bounds__(256) void batch_kernel(
    const BatchInst* __restrict__ insts, const BatchItem* __restrict__ items,
    int nitems, int parity) {
  const int lane = threadIdx.x & 63;
  const int first = __builtin_amdgcn_readfirstlane(
      xcd_remap(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6));
  const int stride = (int)gridDim.x * 4;
  for (int it = first; it < nitems; it += stride) {
    const BatchItem item = items[it];
    const BatchInst& I = insts[item.inst];
    switch (I.kind) {
      case 0: run_strip<K, 1, true>(I, parity, item, lane); break;
      case 1: run_strip<K, 2, false>(I, parity, item, lane); break;
      case 2: run_strip<K, 2, true>(I, parity, item, lane); break;
      default: run_strip<K, 1, false>(I, parity, item, lane); break;
    }
  }
}

// FP64 sums per instance: one workgroup per instance, each with the free-
// energy functional set for it (energy.hpp).
__global__ __launch_bounds__(256) void batch_reduce(
    const BatchInst* __restrict__ insts, const EnergyParams* __restrict__ eps,
    int parity, double* __restrict__ out) {
  const BatchInst& I = insts[blockIdx.x];
  const EnergyParams ep = eps[blockIdx.x];
  const float* f = I.buf[parity];
  const int X = I.X, Y = I.Y;
  const bool central = ep.central != 0;
  double m = 0.0, e = 0.0;
  for (int y = 0; y < Y; ++y) {
    const float* row = f + (int64_t)y * X;
    const float* rowyn = f + (int64_t)(y + 1 == Y ? 0 : y + 1) * X;
    const float* rowyp =
        central ? f + (int64_t)(y == 0 ? Y - 1 : y - 1) * X : row;
    for (int x = threadIdx.x; x < X; x += 256) {
      const double p = row[x];
      const int xp = central ? (x == 0 ? X - 1 : x - 1) : x;
      const double gx = (double)row[x + 1 == X ? 0 : x + 1] - (double)row[xp];
      const double gy = (double)rowyn[x] - (double)rowyp[x];
      m += p;
      e += landau_density(ep, p, gx * gx + gy * gy);
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    m += __shfl_xor(m, o, 64);
    e += __shfl_xor(e, o, 64);
  }
  __shared__ double sh[2][4];
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = m;
    sh[1][threadIdx.x >> 6] = e;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
    out[2 * blockIdx.x + 1] = ((sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3])) *
                              (ep.mean ? 1.0 / ((double)X * Y) : I.h * I.h);
  }
}

int kind_of(const FusedCoef<float>& c) {
  const bool hasa = c.dega >= 0;
  if (c.radius == 1) return hasa ? 0 : 3;
  return hasa ? 2 : 1;
}

}  // namespace
}  // namespace sps

using namespace sps;

struct sps_fd_batch {
  int count = 0;
  std::vector<sps_fd_desc> descs;
  std::vector<size_t> offset;  // element offset of each instance in a buffer
  size_t total = 0;            // elements per buffer
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  float* pool = nullptr;  // 2 * total floats
  BatchInst* d_insts = nullptr;
  BatchItem* d_items[4] = {};  // item table per steps-per-launch (strip widths differ)
  double* d_red = nullptr;
  EnergyParams* d_energy = nullptr;  // per-instance free-energy functional
  int nitems[4] = {};
  int parity = 0;
  // 4 steps x 64-row chunks: best on 256 mixed instances (1.61 T site-updates/s;
  // 32-row chunks give more items than resident waves and a second round)
  int K = 4, TY = 64;
  bool has_data = false;
  GraphSlot graphs[2];
  std::vector<BatchInst> h_insts;

  void drop_graphs() {
    for (auto& g : graphs) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      if (g.graph) (void)hipGraphDestroy(g.graph);
      g = GraphSlot{};
    }
  }

  // work items of the K-step kernel at item height ty (strips x chunks)
  long count_items(int k, int ty) const {
    long n = 0;
    for (int i = 0; i < count; ++i) {
      const int R = (h_insts[i].kind == 1 || h_insts[i].kind == 2) ? 2 : 1;
      const int W = batch_strip_width(k, R);
      n += (long)((descs[i].n[0] + W - 1) / W) * ((descs[i].n[1] + ty - 1) / ty);
    }
    return n;
  }

  int build_items() {
    for (int k = 1; k <= 4; ++k) {
      std::vector<BatchItem> items;
      for (int i = 0; i < count; ++i) {
        const int R = (h_insts[i].kind == 1 || h_insts[i].kind == 2) ? 2 : 1;
        const int W = batch_strip_width(k, R);
        const int X = descs[i].n[0], Y = descs[i].n[1];
        for (int st = 0; st * W < X; ++st)
          for (int y0 = 0; y0 < Y; y0 += TY)
            items.push_back(BatchItem{i, y0, std::min(TY, Y - y0), st});
      }
      // largest first: cost ~ streamed rows (every strip is equally wide)
      std::stable_sort(items.begin(), items.end(),
                       [](const BatchItem& a, const BatchItem& b) {
                         return a.ty > b.ty;
                       });
      if (d_items[k - 1]) (void)hipFree(d_items[k - 1]);
      d_items[k - 1] = nullptr;
      nitems[k - 1] = (int)items.size();
      SPS_HIP_TRY(hipMalloc(&d_items[k - 1], sizeof(BatchItem) * items.size()));
      SPS_HIP_TRY(hipMemcpy(d_items[k - 1], items.data(),
                            sizeof(BatchItem) * items.size(),
                            hipMemcpyHostToDevice));
    }
    return SPS_OK;
  }

  int resident[4] = {1, 1, 1, 1};  // blocks resident at once, per kernel
  unsigned blocks(int k) const {
    return (unsigned)std::max(1, std::min((nitems[k] + 3) / 4, resident[k]));
  }

  int launch_one(int k) {
    switch (k) {
      case 1: batch_kernel<1><<<blocks(0), 256, 0, stream>>>(d_insts, d_items[0], nitems[0], parity); break;
      case 2: batch_kernel<2><<<blocks(1), 256, 0, stream>>>(d_insts, d_items[1], nitems[1], parity); break;
      case 3: batch_kernel<3><<<blocks(2), 256, 0, stream>>>(d_insts, d_items[2], nitems[2], parity); break;
      default: batch_kernel<4><<<blocks(3), 256, 0, stream>>>(d_insts, d_items[3], nitems[3], parity); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_error("batch launch failed: %s", hipGetErrorString(e));
      return SPS_E_HIP;
    }
    parity = 1 - parity;
    return SPS_OK;
  }

  int step(int nsteps) {
    int launches = nsteps / K;
    const int rem = nsteps % K;
    // the null stream cannot be captured: launch one by one there
    if (launches >= 2 * kGraphLaunches && stream != nullptr) {
      GraphSlot& g = graphs[parity];
      if (!g.exec) {
        const int start = parity;
        SPS_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
        int rc = SPS_OK;
        for (int i = 0; i < kGraphLaunches && rc == SPS_OK; ++i) rc = launch_one(K);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(stream, &graph);
        parity = start;
        if (rc != SPS_OK) {
          if (graph) (void)hipGraphDestroy(graph);
          return rc;
        }
        if (e != hipSuccess) {
          set_error("batch graph capture failed: %s", hipGetErrorString(e));
          return SPS_E_HIP;
        }
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
          (void)hipGraphDestroy(graph);
          set_error("batch graph instantiate failed: %s", hipGetErrorString(e));
          return SPS_E_HIP;
        }
        g.graph = graph;
        g.exec = exec;
      }
      while (launches >= kGraphLaunches) {
        SPS_HIP_TRY(hipGraphLaunch(g.exec, stream));
        launches -= kGraphLaunches;
      }
    }
    for (int i = 0; i < launches; ++i) SPS_TRY(launch_one(K));
    if (rem) SPS_TRY(launch_one(rem));
    return SPS_OK;
  }
};

static int batch_width_ok(int n) { return n >= 2 && n % kBatchVec == 0; }

extern "C" {

int sps_fd_batch_create(const sps_fd_desc* descs, int count,
                        sps_fd_batch** out) {
  if (!out || !descs || count < 1) {
    set_error("batch: descriptors, count >= 1 and out are required");
    return SPS_E_INVALID;
  }
  *out = nullptr;
  for (int i = 0; i < count; ++i) {
    int rc = validate_desc(&descs[i]);
    if (rc != SPS_OK) return rc;
    const sps_fd_desc& d = descs[i];
    if (d.dim != 2 || d.dtype != SPS_F32) {
      set_error("batch instance %d: the batched kernel advances 2D FP32 fields",
                i);
      return SPS_E_UNSUPPORTED;
    }
    if (!batch_width_ok(d.n[0])) {
      set_error("batch instance %d: x extent %d is odd (lanes hold 2 values)",
                i, d.n[0]);
      return SPS_E_UNSUPPORTED;
    }
    Lowered lw = lower(d);
    if (!lw.fusable) {
      set_error("batch instance %d: %s", i, lw.why_not);
      return SPS_E_UNSUPPORTED;
    }
    for (int k = 4; k <= kMaxDeg; ++k)
      if (lw.a[k] != 0.0 || lw.b[k] != 0.0) {
        set_error("batch instance %d: polynomial degree above 3", i);
        return SPS_E_UNSUPPORTED;
      }
    FusedCoef<float> fc;
    make_fused_coef(d, lw, fc);
    if (fc.poly != 3) {
      set_error("batch instance %d: laplacian of a non-linear operand without "
                "a bilaplacian term",
                i);
      return SPS_E_UNSUPPORTED;
    }
  }
  if (sps_fd_device_count() < 1) {
    set_error("no HIP device: this library runs the step on an MI355X only");
    return SPS_E_HIP;
  }
  sps_fd_batch* b = new (std::nothrow) sps_fd_batch();
  if (!b) {
    set_error("host allocation failed");
    return SPS_E_NOMEM;
  }
  b->count = count;
  b->descs.assign(descs, descs + count);
  size_t off = 0;
  for (int i = 0; i < count; ++i) {
    b->offset.push_back(off);
    const size_t n = (size_t)descs[i].n[0] * descs[i].n[1];
    off += (n + 63) & ~(size_t)63;  // 256-byte aligned instances
  }
  b->total = off;
  auto fail = [&](int rc) {
    sps_fd_batch_destroy(b);
    return rc;
  };
  if (hipStreamCreateWithFlags(&b->own_stream, hipStreamNonBlocking) !=
      hipSuccess) {
    set_error("stream creation failed");
    return fail(SPS_E_HIP);
  }
  b->stream = b->own_stream;
  if (hipMalloc(&b->pool, 2 * b->total * sizeof(float)) != hipSuccess) {
    set_error("hipMalloc of the batch pool (%zu bytes) failed",
              2 * b->total * sizeof(float));
    return fail(SPS_E_NOMEM);
  }
  b->h_insts.resize(count);
  for (int i = 0; i < count; ++i) {
    BatchInst& I = b->h_insts[i];
    I.buf[0] = b->pool + b->offset[i];
    I.buf[1] = b->pool + b->total + b->offset[i];
    I.X = descs[i].n[0];
    I.Y = descs[i].n[1];
    I.pad = 0;
    I.h = descs[i].h;
    I.divh = 1.0 / descs[i].h;
    make_fused_coef(descs[i], lower(descs[i]), I.c);
    I.kind = kind_of(I.c);
  }
  if (hipMalloc(&b->d_insts, sizeof(BatchInst) * count) != hipSuccess ||
      hipMemcpy(b->d_insts, b->h_insts.data(), sizeof(BatchInst) * count,
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMalloc(&b->d_red, 2 * sizeof(double) * count) != hipSuccess ||
      hipMalloc(&b->d_energy, sizeof(EnergyParams) * count) != hipSuccess) {
    set_error("batch device set-up failed");
    return fail(SPS_E_HIP);
  }
  if (sps_fd_batch_set_energy(b, nullptr) != SPS_OK) return fail(SPS_E_HIP);
  int rc = b->build_items();
  if (rc != SPS_OK) return fail(rc);
  // persistent grid: as many workgroups as are resident on the device at once
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
          hipSuccess) {
    set_error("device query failed");
    return fail(SPS_E_HIP);
  }
  const void* kernels[4] = {(const void*)batch_kernel<1>, (const void*)batch_kernel<2>,
                            (const void*)batch_kernel<3>, (const void*)batch_kernel<4>};
  for (int k = 0; k < 4; ++k) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernels[k], 256,
                                                     0) != hipSuccess)
      per_cu = 1;
    b->resident[k] = std::max(1, cus * std::max(1, per_cu));
  }
  *out = b;
  return SPS_OK;
}

int sps_fd_batch_destroy(sps_fd_batch* b) {
  if (!b) return SPS_OK;
  b->drop_graphs();
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->pool) (void)hipFree(b->pool);
  if (b->d_insts) (void)hipFree(b->d_insts);
  for (BatchItem* p : b->d_items)
    if (p) (void)hipFree(p);
  if (b->d_red) (void)hipFree(b->d_red);
  if (b->d_energy) (void)hipFree(b->d_energy);
  if (b->own_stream) (void)hipStreamDestroy(b->own_stream);
  delete b;
  return SPS_OK;
}

static int batch_index_ok(sps_fd_batch* b, int index, const void* host) {
  if (!b || !host) {
    set_error("batch or host pointer is NULL");
    return SPS_E_INVALID;
  }
  if (index < 0 || index >= b->count) {
    set_error("batch index %d out of range [0, %d)", index, b->count);
    return SPS_E_INVALID;
  }
  return SPS_OK;
}

int sps_fd_batch_upload(sps_fd_batch* b, int index, const void* host,
                        int layout) {
  SPS_TRY(batch_index_ok(b, index, host));
  const sps_fd_desc& d = b->descs[index];
  const size_t n = (size_t)d.n[0] * d.n[1];
  std::vector<float> tmp;
  const void* src = host;
  if (layout == SPS_LAYOUT_PADDED) {
    tmp.resize(n);
    padded_to_interior(d, (const char*)host, (char*)tmp.data(), sizeof(float));
    src = tmp.data();
  } else if (layout != SPS_LAYOUT_INTERIOR) {
    set_error("unknown layout %d", layout);
    return SPS_E_INVALID;
  }
  SPS_HIP_TRY(hipMemcpyAsync(b->h_insts[index].buf[b->parity], src,
                             n * sizeof(float), hipMemcpyHostToDevice,
                             b->stream));
  SPS_HIP_TRY(hipStreamSynchronize(b->stream));
  b->has_data = true;
  return SPS_OK;
}

int sps_fd_batch_download(sps_fd_batch* b, int index, void* host, int layout) {
  SPS_TRY(batch_index_ok(b, index, host));
  const sps_fd_desc& d = b->descs[index];
  const size_t n = (size_t)d.n[0] * d.n[1];
  std::vector<float> tmp(n);
  SPS_HIP_TRY(hipMemcpyAsync(tmp.data(), b->h_insts[index].buf[b->parity],
                             n * sizeof(float), hipMemcpyDeviceToHost,
                             b->stream));
  SPS_HIP_TRY(hipStreamSynchronize(b->stream));
  if (layout == SPS_LAYOUT_INTERIOR) {
    std::memcpy(host, tmp.data(), n * sizeof(float));
  } else if (layout == SPS_LAYOUT_PADDED) {
    interior_to_padded(d, (const char*)tmp.data(), (char*)host, sizeof(float),
                       true);
  } else {
    set_error("unknown layout %d", layout);
    return SPS_E_INVALID;
  }
  return SPS_OK;
}

int sps_fd_batch_step(sps_fd_batch* b, int nsteps) {
  if (!b || nsteps < 0) {
    set_error("batch is NULL or nsteps < 0");
    return SPS_E_INVALID;
  }
  if (!b->has_data) {
    set_error("no field uploaded");
    return SPS_E_STATE;
  }
  return b->step(nsteps);
}

int sps_fd_batch_reduce(sps_fd_batch* b, double* mass, double* energy) {
  if (!b) {
    set_error("batch is NULL");
    return SPS_E_INVALID;
  }
  batch_reduce<<<b->count, 256, 0, b->stream>>>(b->d_insts, b->d_energy,
                                                 b->parity, b->d_red);
  SPS_HIP_TRY(hipGetLastError());
  std::vector<double> h(2 * (size_t)b->count);
  SPS_HIP_TRY(hipMemcpyAsync(h.data(), b->d_red, h.size() * sizeof(double),
                             hipMemcpyDeviceToHost, b->stream));
  SPS_HIP_TRY(hipStreamSynchronize(b->stream));
  for (int i = 0; i < b->count; ++i) {
    if (mass) mass[i] = h[2 * i];
    if (energy) energy[i] = h[2 * i + 1];
  }
  return SPS_OK;
}

int sps_fd_batch_set_energy(sps_fd_batch* b, const sps_fd_energy* energy) {
  if (!b) {
    set_error("batch is NULL");
    return SPS_E_INVALID;
  }
  std::vector<EnergyParams> eps((size_t)b->count);
  for (int i = 0; i < b->count; ++i) {
    const int rc = make_energy_params(energy ? &energy[i] : nullptr,
                                      b->descs[i].h, eps[i]);
    if (rc != SPS_OK) {
      const std::string why = last_error();  // set_error overwrites it
      set_error("batch instance %d: %s", i, why.c_str());
      return rc;
    }
  }
  // ordered after the batch's earlier reductions on its stream
  SPS_HIP_TRY(hipMemcpyAsync(b->d_energy, eps.data(),
                             sizeof(EnergyParams) * eps.size(),
                             hipMemcpyHostToDevice, b->stream));
  SPS_HIP_TRY(hipStreamSynchronize(b->stream));
  return SPS_OK;
}

int sps_fd_batch_sync(sps_fd_batch* b) {
  if (!b) {
    set_error("batch is NULL");
    return SPS_E_INVALID;
  }
  SPS_HIP_TRY(hipStreamSynchronize(b->stream));
  return SPS_OK;
}

int sps_fd_batch_tune(sps_fd_batch* b, int steps_per_launch,
                      int rows_per_item) {
  if (!b || steps_per_launch < 0 || steps_per_launch > 4 ||
      (rows_per_item != 0 && (rows_per_item < 4 || rows_per_item > 512))) {
    set_error("batch tune: steps_per_launch in 0..4, rows_per_item 0 or "
              "4..512");
    return SPS_E_INVALID;
  }
  SPS_HIP_TRY(hipStreamSynchronize(b->stream));
  if (steps_per_launch) b->K = steps_per_launch;
  if (rows_per_item && rows_per_item != b->TY) {
    b->TY = rows_per_item;
    SPS_TRY(b->build_items());
  }
  b->drop_graphs();
  return SPS_OK;
}

int sps_fd_batch_copy_device(sps_fd_batch* b, int index, void* dev,
                             int direction) {
  SPS_TRY(batch_index_ok(b, index, dev));
  if (direction != 0 && direction != 1) {
    set_error("batch copy_device: direction must be 0 or 1");
    return SPS_E_INVALID;
  }
  const sps_fd_desc& d = b->descs[index];
  const size_t bytes = (size_t)d.n[0] * d.n[1] * sizeof(float);
  float* buf = b->h_insts[index].buf[b->parity];
  if (direction == 0)
    SPS_HIP_TRY(hipMemcpyAsync(dev, buf, bytes, hipMemcpyDeviceToDevice,
                               b->stream));
  else {
    SPS_HIP_TRY(hipMemcpyAsync(buf, dev, bytes, hipMemcpyDeviceToDevice,
                               b->stream));
    b->has_data = true;
  }
  return SPS_OK;
}

int sps_fd_batch_autotune(sps_fd_batch* b, int* out_k, int* out_rows) {
  if (!b) {
    set_error("batch is NULL");
    return SPS_E_INVALID;
  }
  if (!b->has_data) {
    set_error("autotune needs uploaded fields");
    return SPS_E_STATE;
  }
  const size_t bytes = b->total * sizeof(float);
  float* saved = nullptr;
  SPS_HIP_TRY(hipMalloc(&saved, bytes));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = SPS_OK;
  const int oldK = b->K, oldTY = b->TY;
  int bestK = oldK, bestTY = oldTY;
  double best = 1e300;
  float* cur = b->pool + (size_t)b->parity * b->total;
  if (hipMemcpyAsync(saved, cur, bytes, hipMemcpyDeviceToDevice, b->stream) !=
          hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = SPS_E_HIP;
  // item heights: fixed ones, plus per m = 2..6 the lowest height whose
  // 4-step item count fits m items per SIMD of the device (a SIMD with one
  // item more than the others sets the launch's length)
  std::vector<int> cand_rows = {16, 32, 64, 128};
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
          hipSuccess)
    rc = SPS_E_HIP;
  for (int m = 2; m <= 6 && rc == SPS_OK; ++m) {
    const long target = (long)m * cus * 4;
    if (b->count_items(4, 512) > target) continue;
    int lo = 4, hi = 512;  // count_items is non-increasing in the height
    while (lo < hi) {
      const int mid = (lo + hi) / 2;
      if (b->count_items(4, mid) <= target)
        hi = mid;
      else
        lo = mid + 1;
    }
    if (std::find(cand_rows.begin(), cand_rows.end(), lo) == cand_rows.end())
      cand_rows.push_back(lo);
  }
  for (int rows : cand_rows) {
    if (rc != SPS_OK) break;
    b->TY = rows;
    if ((rc = b->build_items()) != SPS_OK) break;
    for (int k = 1; k <= 4 && rc == SPS_OK; ++k) {
      double ms_best = 1e300;
      for (int rep = 0; rep < 3 && rc == SPS_OK; ++rep) {
        if ((rc = b->launch_one(k)) != SPS_OK) break;
        if (hipEventRecord(e0, b->stream) != hipSuccess) rc = SPS_E_HIP;
        for (int i = 0; i < 12 / k && rc == SPS_OK; ++i) rc = b->launch_one(k);
        if (rc != SPS_OK || hipEventRecord(e1, b->stream) != hipSuccess ||
            hipEventSynchronize(e1) != hipSuccess) {
          if (rc == SPS_OK) rc = SPS_E_HIP;
          break;
        }
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          rc = SPS_E_HIP;
          break;
        }
        if (rep > 0) ms_best = std::min(ms_best, (double)ms);
      }
      if (rc == SPS_OK && ms_best < best) {
        best = ms_best;
        bestK = k;
        bestTY = rows;
      }
    }
  }
  cur = b->pool + (size_t)b->parity * b->total;
  if (hipMemcpyAsync(cur, saved, bytes, hipMemcpyDeviceToDevice, b->stream) !=
          hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess) {
    if (rc == SPS_OK) rc = SPS_E_HIP;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(saved);
  b->K = rc == SPS_OK ? bestK : oldK;
  b->TY = rc == SPS_OK ? bestTY : oldTY;
  int rc2 = b->build_items();
  b->drop_graphs();
  if (rc != SPS_OK) {
    set_error("batch autotune failed");
    return rc;
  }
  if (rc2 != SPS_OK) return rc2;
  if (out_k) *out_k = b->K;
  if (out_rows) *out_rows = b->TY;
  return SPS_OK;
}

int sps_fd_batch_set_stream(sps_fd_batch* b, void* hip_stream) {
  if (!b) {
    set_error("batch is NULL");
    return SPS_E_INVALID;
  }
  SPS_HIP_TRY(hipStreamSynchronize(b->stream));
  b->stream = (hipStream_t)hip_stream;
  return SPS_OK;
}

}  // extern "C"
