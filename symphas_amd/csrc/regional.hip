// regional.hip - SolverFT on a regional grid (SolverSystemFDwSD), one GPU.
//
// The field lives on a box of the periodic domain with a boundary layer of
// SPS_BOUNDARY_DEPTH cells per side (RegionalGrid, datatypes/inc/grid.h:
// 2038-2157); every site outside the box reads as the grid's `empty` value.
// A step evaluates the right-hand side and forward Euler over the box
// interior with the reference-order kernel in its box form, so its FP64
// arithmetic is that of the CPU SolverFT.  The box follows the sites above a
// cutoff (SolverSystemFDwSD::update, sol/inc/solversystem.h:62-77;
// update_regional_system, sol/inc/boundarysystem.h:437-479):
//   - first update: cutoff = min + eps, empty = min, resize by the padding
//     factor (resize_adjust_region, datatypes/inc/gridfunctions.h:2194-2238);
//   - then, whenever time / resize_time reaches the next resize count:
//     re-centre a box of the same dims (adjust_region, :2166-2176) or, when
//     the resize is not fixed, resize with the relative dims tolerance.
// The active box along an axis is the shortest periodic interval covering
// every coordinate of an interior site above the cutoff.  The device marks
// the occupied coordinates of each axis; the host finds the interval.
// Along an axis the box spans, its layer is the periodic image of its
// interior, refreshed before every right-hand side
// (regional_update_boundary, datatypes/inc/boundaryupdatedata.h:1398-1433);
// along the others the layer keeps what it was given when the box moved.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

#include "fused_tile.hpp"
#include "lanes.hpp"
#include "plan.hpp"
#include "ref_kernels.hpp"

namespace sps {
namespace {

constexpr int kB = SPS_BOUNDARY_DEPTH;
constexpr int kMinBlocks = 256;

__device__ __forceinline__ int modn(int v, int n) {
  int r = v % n;
  return r < 0 ? r + n : r;
}

struct BoxGeom {
  int n[3];  // domain interior counts (1 beyond dim)
  int o[3];  // domain coordinate of the first interior site of the box
  int b[3];  // box dims including the layer (1 beyond dim)
};

// Per-block minimum of the box interior, FP64.
template <typename T>
__global__ __launch_bounds__(256) void box_min_pass1(const T* __restrict__ box,
                                                     BoxGeom g, int dim,
                                                     double* partial) {
  const int mx = g.b[0] - 2 * kB, my = g.b[1] - 2 * kB;
  const int mz = dim == 3 ? g.b[2] - 2 * kB : 1;
  const int zo = dim == 3 ? kB : 0;
  const int rows = my * mz;
  double m = INFINITY;
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    const int y = r % my, z = r / my;
    const T* row = box + ((int64_t)(z + zo) * g.b[1] + (y + kB)) * g.b[0] + kB;
    for (int x = threadIdx.x; x < mx; x += blockDim.x) m = fmin(m, (double)row[x]);
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) m = fmin(m, __shfl_xor(m, o, 64));
  __shared__ double sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    partial[blockIdx.x] = fmin(fmin(sh[0], sh[1]), fmin(sh[2], sh[3]));
}

__global__ __launch_bounds__(256) void box_min_pass2(const double* partial,
                                                     int nb, double* out) {
  double m = INFINITY;
  for (int i = threadIdx.x; i < nb; i += 256) m = fmin(m, partial[i]);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) m = fmin(m, __shfl_xor(m, o, 64));
  __shared__ double sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) *out = fmin(fmin(sh[0], sh[1]), fmin(sh[2], sh[3]));
}

__global__ __launch_bounds__(256) void fill_int(int* p, int count) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < count) p[i] = 0;
}

// occ[a][coordinate] = 1 for every axis-a coordinate of an interior site
// whose value exceeds the cutoff (occ holds n0 + n1 + n2 flags).  Every
// writer of a flag writes the same 1.
template <typename T>
__global__ __launch_bounds__(256) void box_occupancy(const T* __restrict__ box,
                                                     BoxGeom g, int dim,
                                                     double cutoff, int* occ) {
  const int mx = g.b[0] - 2 * kB, my = g.b[1] - 2 * kB;
  const int mz = dim == 3 ? g.b[2] - 2 * kB : 1;
  const int zo = dim == 3 ? kB : 0;
  const int64_t total = (int64_t)mx * my * mz;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * 256) {
    const int x = (int)(i % mx);
    const int64_t r = i / mx;
    const int y = (int)(r % my), z = (int)(r / my);
    const T v = box[((int64_t)(z + zo) * g.b[1] + (y + kB)) * g.b[0] + x + kB];
    if ((double)v > cutoff) {
      occ[modn(g.o[0] + x, g.n[0])] = 1;
      occ[g.n[0] + modn(g.o[1] + y, g.n[1])] = 1;
      if (dim == 3) occ[g.n[0] + g.n[1] + modn(g.o[2] + z, g.n[2])] = 1;
    }
  }
}

// New box from the old: a new site takes the old interior value at the same
// domain site, else `empty`.
template <typename T>
__global__ __launch_bounds__(256) void box_move(const T* __restrict__ old,
                                                BoxGeom go, T* __restrict__ nw,
                                                BoxGeom gn, int dim, T empty) {
  const int64_t total = (int64_t)gn.b[0] * gn.b[1] * gn.b[2];
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * 256) {
    int l[3];
    l[0] = (int)(i % gn.b[0]);
    const int64_t r = i / gn.b[0];
    l[1] = (int)(r % gn.b[1]);
    l[2] = (int)(r / gn.b[1]);
    int64_t src = 0, stride = 1;
    bool inside = true;
    for (int a = 0; a < dim; ++a) {
      const int gc = modn(gn.o[a] + l[a] - kB, gn.n[a]);
      const int k = modn(gc - go.o[a], gn.n[a]);
      if (k >= go.b[a] - 2 * kB) inside = false;
      src += (int64_t)(k + kB) * stride;
      stride *= go.b[a];
    }
    nw[i] = inside ? old[src] : empty;
  }
}

// Periodic layer along axis `ax` of a box that spans it.
template <typename T>
__global__ __launch_bounds__(256) void box_refresh_axis(T* __restrict__ box,
                                                        BoxGeom g, int ax) {
  const int64_t total = (int64_t)g.b[0] * g.b[1] * g.b[2];
  int64_t sa = 1;
  for (int j = 0; j < ax; ++j) sa *= g.b[j];
  const int64_t shift = (int64_t)g.n[ax] * sa;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * 256) {
    const int la = (int)((i / sa) % g.b[ax]);
    if (la < kB)
      box[i] = box[i + shift];
    else if (la >= g.b[ax] - kB)
      box[i] = box[i - shift];
  }
}

// Whole-domain field (interior layout) <-> box.
template <typename T>
__global__ __launch_bounds__(256) void box_gather(const T* __restrict__ box,
                                                  BoxGeom g, int dim,
                                                  int has_box, T empty,
                                                  T* __restrict__ out) {
  const int64_t total = (int64_t)g.n[0] * g.n[1] * g.n[2];
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * 256) {
    int c[3];
    c[0] = (int)(i % g.n[0]);
    const int64_t r = i / g.n[0];
    c[1] = (int)(r % g.n[1]);
    c[2] = (int)(r / g.n[1]);
    int64_t src = 0, stride = 1;
    bool inside = has_box != 0;
    for (int a = 0; a < dim; ++a) {
      const int k = modn(c[a] - g.o[a], g.n[a]);
      if (k >= g.b[a] - 2 * kB) inside = false;
      src += (int64_t)(k + kB) * stride;
      stride *= g.b[a];
    }
    out[i] = inside ? box[src] : empty;
  }
}

// The padded layout of the whole domain (the box spanning every axis at
// origin 0) from an interior-layout field.
template <typename T>
__global__ __launch_bounds__(256) void box_scatter_full(
    const T* __restrict__ in, BoxGeom g, int dim, T* __restrict__ box) {
  const int64_t total = (int64_t)g.b[0] * g.b[1] * g.b[2];
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * 256) {
    int l[3];
    l[0] = (int)(i % g.b[0]);
    const int64_t r = i / g.b[0];
    l[1] = (int)(r % g.b[1]);
    l[2] = (int)(r / g.b[1]);
    int64_t src = 0, stride = 1;
    for (int a = 0; a < dim; ++a) {
      src += (int64_t)modn(l[a] - kB, g.n[a]) * stride;
      stride *= g.n[a];
    }
    box[i] = in[src];
  }
}

// Fused form on a box (the canonical form of fused.hpp's kernels, one step per
// pass pair): pass A evaluates mu~ = S + psi*p(psi) at every box site within
// one site of the interior, pass B the update psi' = psi + q(psi) +
// k2*(S[mu~] - w*mu~) over the interior; the radius-one form is pass B alone
// on psi.  Each expression, its operand order and its FMAs are those of the
// streamed kernels (fused_tile.hpp, fused3d.hip), so a box that spans the
// domain gives their one-step-per-launch results bit for bit.
template <typename T, int DIM>
__device__ __forceinline__ T box_nsum(const T* f, int64_t i, int64_t sy,
                                      int64_t sz) {
  if (DIM == 2) return (f[i - 1] + f[i + 1]) + (f[i + sy] + f[i - sy]);
  return ((f[i - 1] + f[i - sy]) + (f[i + 1] + f[i + sy])) + (f[i - sz] + f[i + sz]);
}

// One thread per box site, (64 x 4) threads per workgroup over (x, y) and
// one grid layer per z: no index division.
template <typename T, int DIM, int POLY>
__global__ __launch_bounds__(256) void box_mu(const T* __restrict__ psi,
                                              T* __restrict__ mu, BoxGeom g,
                                              FusedCoef<T> c) {
  const int lx = blockIdx.x * 64 + (threadIdx.x & 63);
  const int ly = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int lz = blockIdx.z;
  // sites within one of the interior: their neighbours are in the box
  if (lx < kB - 1 || lx >= g.b[0] - kB + 1 || ly < kB - 1 || ly >= g.b[1] - kB + 1)
    return;
  if (DIM == 3 && (lz < kB - 1 || lz >= g.b[2] - kB + 1)) return;
  const int64_t sy = g.b[0], sz = (int64_t)g.b[0] * g.b[1];
  const int64_t i = lx + sy * ly + sz * lz;
  const T sn = box_nsum<T, DIM>(psi, i, sy, sz);
  const T x = psi[i];
  const T t = horner<POLY - 1>(c.p, x);
  mu[i] = fmad(t, x, POLY == 3 ? sn : c.ks * sn);
}

template <typename T, int DIM, int R, bool HASA, int POLY>
__global__ __launch_bounds__(256) void box_update(const T* __restrict__ psi,
                                                  const T* __restrict__ mu,
                                                  T* __restrict__ out,
                                                  BoxGeom g, FusedCoef<T> c) {
  const int lx = blockIdx.x * 64 + (threadIdx.x & 63);
  const int ly = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int lz = blockIdx.z;
  if (lx >= g.b[0] || ly >= g.b[1]) return;
  const int64_t sy = g.b[0], sz = (int64_t)g.b[0] * g.b[1];
  const int64_t i = lx + sy * ly + sz * lz;
  constexpr T w = T(2 * DIM);
  const bool interior = lx >= kB && lx < g.b[0] - kB && ly >= kB &&
                        ly < g.b[1] - kB &&
                        (DIM == 2 || (lz >= kB && lz < g.b[2] - kB));
  if (!interior) {
    out[i] = psi[i];
    return;
  }
  if constexpr (R == 2) {
    const T sn = box_nsum<T, DIM>(mu, i, sy, sz);
    T base = psi[i];
    if constexpr (HASA) base = base + horner<POLY>(c.q, base);
    out[i] = fmad(c.k2, fmad(-w, mu[i], sn), base);
  } else {
    const T sn = box_nsum<T, DIM>(psi, i, sy, sz);
    const T x = psi[i];
    T base = x;
    if constexpr (HASA) base = base + horner<POLY>(c.q, x);
    out[i] = fmad(c.k1, fmad(-w, x, sn), base);
  }
}

// The 2D fused box step in one pass: a workgroup stages psi on its 64 x 16
// output tile plus a 2-site halo in LDS, evaluates mu~ on the tile plus a
// 1-site halo there, then writes the update (layer sites copied).  Same
// expressions as box_mu / box_update, so the same bits, with one launch and
// one read and write of the box per step.
constexpr int kBTX = 64, kBTY = 16;

template <typename T, int R, bool HASA, int POLY>
__global__ __launch_bounds__(256) void box_fused2d_tile(const T* __restrict__ psi,
                                                        T* __restrict__ out,
                                                        BoxGeom g, FusedCoef<T> c) {
  constexpr int PX = kBTX + 4, PY = kBTY + 4;  // psi tile with a 2-site halo
  constexpr int MX = kBTX + 2, MY = kBTY + 2;  // mu~ tile with a 1-site halo
  __shared__ T sp[PY * PX];
  __shared__ T sm[R == 2 ? MY * MX : 1];
  const int x0 = blockIdx.x * kBTX, y0 = blockIdx.y * kBTY;
  const int bx = g.b[0], by = g.b[1];
  for (int e = threadIdx.x; e < PY * PX; e += 256) {
    const int ty = e / PX, tx = e - ty * PX;
    const int lx = x0 + tx - 2, ly = y0 + ty - 2;
    sp[e] = (lx >= 0 && lx < bx && ly >= 0 && ly < by) ? psi[(int64_t)ly * bx + lx] : T(0);
  }
  __syncthreads();
  if constexpr (R == 2) {
    for (int e = threadIdx.x; e < MY * MX; e += 256) {
      const int ty = e / MX, tx = e - ty * MX;
      const int lx = x0 + tx - 1, ly = y0 + ty - 1;
      if (lx < kB - 1 || lx >= bx - kB + 1 || ly < kB - 1 || ly >= by - kB + 1) continue;
      const int i = (ty + 1) * PX + tx + 1;  // the site in sp
      const T sn = (sp[i - 1] + sp[i + 1]) + (sp[i + PX] + sp[i - PX]);
      const T x = sp[i];
      const T t = horner<POLY - 1>(c.p, x);
      sm[e] = fmad(t, x, POLY == 3 ? sn : c.ks * sn);
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < kBTY * kBTX; e += 256) {
    const int ty = e / kBTX, tx = e - ty * kBTX;
    const int lx = x0 + tx, ly = y0 + ty;
    if (lx >= bx || ly >= by) continue;
    const int ip = (ty + 2) * PX + tx + 2;
    const int64_t o = (int64_t)ly * bx + lx;
    const bool interior = lx >= kB && lx < bx - kB && ly >= kB && ly < by - kB;
    if (!interior) {
      out[o] = sp[ip];
      continue;
    }
    if constexpr (R == 2) {
      const int im = (ty + 1) * MX + tx + 1;
      const T sn = (sm[im - 1] + sm[im + 1]) + (sm[im + MX] + sm[im - MX]);
      T base = sp[ip];
      if constexpr (HASA) base = base + horner<POLY>(c.q, base);
      out[o] = fmad(c.k2, fmad(T(-4), sm[im], sn), base);
    } else {
      const T sn = (sp[ip - 1] + sp[ip + 1]) + (sp[ip + PX] + sp[ip - PX]);
      const T x = sp[ip];
      T base = x;
      if constexpr (HASA) base = base + horner<POLY>(c.q, x);
      out[o] = fmad(c.k1, fmad(T(-4), x, sn), base);
    }
  }
}

template <typename T, int DIM, int R, bool HASA, int POLY>
int launch_box_fused(const T* psi, T* mu, T* out, const BoxGeom& g,
                     const FusedCoef<T>& c, int64_t sites, hipStream_t s);

int grid_for(int64_t work) {
  const int64_t blocks = (work + 255) / 256;
  return (int)std::max<int64_t>(1, std::min<int64_t>(blocks, 8192));
}

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s launch failed: %s", what, hipGetErrorString(e));
    return SPS_E_HIP;
  }
  return SPS_OK;
}

// Shortest periodic interval of [0, n) covering the set flags: the
// complement of the largest run of clear flags on the circle, the
// lowest-starting among equals.  false when no flag is set.
bool cover_interval(const int* occ, int n, int& start, int& length) {
  int first = -1;
  for (int i = 0; i < n; ++i)
    if (occ[i]) {
      first = i;
      break;
    }
  if (first < 0) {
    start = 0;
    length = 0;
    return false;
  }
  int best = 0, best_start = 0, run = 0, run_start = 0;
  for (int k = 1; k <= n; ++k) {
    const int i = (first + k) % n;
    if (!occ[i]) {
      if (run == 0) run_start = i;
      ++run;
      continue;
    }
    if (run > best || (run > 0 && run == best && run_start < best_start)) {
      best = run;
      best_start = run_start;
    }
    run = 0;
  }
  if (best == 0) {
    start = 0;
    length = n;
  } else {
    start = (best_start + best) % n;
    length = n - best;
  }
  return true;
}

// Four fused box steps per launch, for a 2D box whose layer is frozen along
// both axes (no axis spans the domain): the row stream of the whole-domain
// kernel (fused_tile.hpp) in box mode.  A wave owns a strip of 64 lanes x VEC
// columns and streams a chunk of rows through the four stages; rows and
// columns outside the box read 0, and after each stage a layer site keeps
// its value.  Each site sees box_mu / box_update's expressions, so the
// result equals four one-step launches bit for bit.
constexpr int kBoxSteps = 4;

template <typename T, int VEC, int R, bool HASA, int POLY>
__global__ __launch_bounds__(64) void box_stream2d(const T* __restrict__ in,
                                                   T* __restrict__ out, int bx, int by,
                                                   int TY, int nstrips, int nitems,
                                                   FusedCoef<T> c) {
  constexpr int HL = (kBoxSteps * R + VEC - 1) / VEC;
  constexpr int W = (64 - 2 * HL) * VEC;
  const int lane = threadIdx.x;
  const int item = __builtin_amdgcn_readfirstlane(xcd_remap(blockIdx.x, gridDim.x));
  if (item >= nitems) return;
  const int strip = item % nstrips;
  const int chunk = item / nstrips;
  const int cb = strip * W + (lane - HL) * VEC;
  const bool store_lane = lane >= HL && lane < 64 - HL && cb < bx;
  BoxMask bm;
  bm.depth = kB;
#pragma unroll
  for (int j = 0; j < VEC; ++j)
    if (cb + j >= kB && cb + j < bx - kB) bm.cols |= 1u << j;
  const int y0 = chunk * TY;
  const int ty = min(TY, by - y0);
  stream_tile<T, VEC, kBoxSteps, R, HASA, POLY, true>(in, out, bx, by, y0, ty, cb,
                                                       store_lane, c, bm);
}

// Chunk height of the box stream: about two waves per SIMD, at least 16 rows
// (shorter chunks spend most of their rows on the pipeline fill).
inline int box_stream_rows(int by, int nstrips, int cus) {
  const int64_t target = 8LL * cus;  // waves
  const int64_t nch = std::max<int64_t>(1, target / std::max(1, nstrips));
  return (int)std::max<int64_t>(16, (by + nch - 1) / nch);
}

template <typename T, int VEC, int R, bool HASA, int POLY>
int launch_box_stream(const T* psi, T* out, const BoxGeom& g, const FusedCoef<T>& c,
                      int cus, hipStream_t s) {
  constexpr int HL = (kBoxSteps * R + VEC - 1) / VEC;
  constexpr int W = (64 - 2 * HL) * VEC;
  const int bx = g.b[0], by = g.b[1];
  const int nstrips = (bx + W - 1) / W;
  const int TY = box_stream_rows(by, nstrips, cus);
  const int nitems = nstrips * ((by + TY - 1) / TY);
  box_stream2d<T, VEC, R, HASA, POLY><<<nitems, 64, 0, s>>>(psi, out, bx, by, TY, nstrips,
                                                           nitems, c);
  return check_launch("regional fused 4-step launch");
}

// Four fused steps of a frozen-layer 2D box.  FP32 boxes of even width take
// two values per lane (8-byte aligned rows), the others one.
template <typename T>
int box_multi_step(const T* psi, T* out, const BoxGeom& g, const FusedCoef<T>& c,
                   int cus, hipStream_t s) {
  const bool hasa = c.dega >= 0;
  const bool two = sizeof(T) == 4 && g.b[0] % 2 == 0;
#define SPS_BOXS(V_, R_, A_, P_) launch_box_stream<T, V_, R_, A_, P_>(psi, out, g, c, cus, s)
#define SPS_BOXR(V_)                                                                   \
  (c.radius == 1                                                                       \
       ? (c.poly == 3 ? (hasa ? SPS_BOXS(V_, 1, true, 3) : SPS_BOXS(V_, 1, false, 3))  \
                      : (hasa ? SPS_BOXS(V_, 1, true, 6) : SPS_BOXS(V_, 1, false, 6))) \
       : (c.poly == 3 ? (hasa ? SPS_BOXS(V_, 2, true, 3) : SPS_BOXS(V_, 2, false, 3))  \
                      : (hasa ? SPS_BOXS(V_, 2, true, 6) : SPS_BOXS(V_, 2, false, 6))))
  if constexpr (sizeof(T) == 4) {
    if (two) return SPS_BOXR(2);
  }
  return SPS_BOXR(1);
#undef SPS_BOXR
#undef SPS_BOXS
}

template <typename T, int DIM, int R, bool HASA, int POLY>
int launch_box_fused(const T* psi, T* mu, T* out, const BoxGeom& g,
                     const FusedCoef<T>& c, int64_t sites, hipStream_t s) {
  (void)sites;
  if constexpr (DIM == 2) {
    const dim3 tiles((g.b[0] + kBTX - 1) / kBTX, (g.b[1] + kBTY - 1) / kBTY);
    box_fused2d_tile<T, R, HASA, POLY><<<tiles, 256, 0, s>>>(psi, out, g, c);
    return check_launch("regional fused step");
  }
  const dim3 grid((g.b[0] + 63) / 64, (g.b[1] + 3) / 4, DIM == 3 ? g.b[2] : 1);
  if constexpr (R == 2) {
    box_mu<T, DIM, POLY><<<grid, 256, 0, s>>>(psi, mu, g, c);
    SPS_TRY(check_launch("regional fused mu"));
  }
  box_update<T, DIM, R, HASA, POLY><<<grid, 256, 0, s>>>(psi, mu, out, g, c);
  return check_launch("regional fused update");
}

// The instance of the fused box step for a lowered form (radius, A present,
// polynomial bound), as the streamed kernels dispatch it.
template <typename T, int DIM>
int box_fused_step(const T* psi, T* mu, T* out, const BoxGeom& g,
                   const FusedCoef<T>& c, int64_t sites, hipStream_t s) {
  const bool hasa = c.dega >= 0;
#define SPS_BOX(R_, A_, P_) launch_box_fused<T, DIM, R_, A_, P_>(psi, mu, out, g, c, sites, s)
  if (c.radius == 1)
    return c.poly == 3 ? (hasa ? SPS_BOX(1, true, 3) : SPS_BOX(1, false, 3))
                       : (hasa ? SPS_BOX(1, true, 6) : SPS_BOX(1, false, 6));
  return c.poly == 3 ? (hasa ? SPS_BOX(2, true, 3) : SPS_BOX(2, false, 3))
                     : (hasa ? SPS_BOX(2, true, 6) : SPS_BOX(2, false, 6));
#undef SPS_BOX
}

int mod_host(int64_t v, int n) {
  const int64_t r = v % n;
  return (int)(r < 0 ? r + n : r);
}

}  // namespace
}  // namespace sps

struct sps_fdr_plan {
  sps_fd_desc d{};
  sps_fd_regional rp{};
  hipStream_t stream = nullptr;
  size_t esize = 0, nsites = 0, cap = 0;  // cap: sites of the whole padded box
  void* box[2] = {nullptr, nullptr};
  int cur = 0;
  void* staging = nullptr;  // whole-domain field, interior layout
  int* occ = nullptr;       // n0 + n1 + n2 flags
  double* partial = nullptr;
  double* minval = nullptr;
  sps::RefParams<float> rp32;
  sps::RefParams<double> rp64;
  bool fused = false;      // canonical form: the fused box step
  void* mu = nullptr;      // its mu~ box (radius two)
  sps::FusedCoef<float> fc32;
  sps::FusedCoef<double> fc64;
  int o[3] = {0, 0, 0}, b[3] = {1, 1, 1};
  bool has_data = false, has_box = true;
  int next_resize = 0, index = 0;
  int cus = 256;  // compute units of the plan's device (box stream chunking)
  double empty = 0.0, cutoff = 0.0, time = 0.0;

  sps::BoxGeom geom() const {
    sps::BoxGeom g;
    for (int a = 0; a < 3; ++a) {
      g.n[a] = a < d.dim ? d.n[a] : 1;
      g.o[a] = a < d.dim ? o[a] : 0;
      g.b[a] = a < d.dim ? b[a] : 1;
    }
    return g;
  }

  int64_t box_sites() const {
    return (int64_t)b[0] * b[1] * (d.dim == 3 ? b[2] : 1);
  }

  int find_min(double& out) {
    const sps::BoxGeom g = geom();
    const int rows = (b[1] - 2 * sps::kB) * (d.dim == 3 ? b[2] - 2 * sps::kB : 1);
    const int nb = std::max(1, std::min(rows, sps::kMinBlocks));
    if (d.dtype == SPS_F32)
      sps::box_min_pass1<float><<<nb, 256, 0, stream>>>((const float*)box[cur], g,
                                                         d.dim, partial);
    else
      sps::box_min_pass1<double><<<nb, 256, 0, stream>>>((const double*)box[cur],
                                                          g, d.dim, partial);
    sps::box_min_pass2<<<1, 256, 0, stream>>>(partial, nb, minval);
    SPS_TRY(sps::check_launch("minimum"));
    SPS_HIP_TRY(hipMemcpyAsync(&out, minval, sizeof(double),
                               hipMemcpyDeviceToHost, stream));
    SPS_HIP_TRY(hipStreamSynchronize(stream));
    return SPS_OK;
  }

  // Active box (start s, length L) per axis; found = false when no interior
  // site exceeds the cutoff.
  int active_box(int* s, int* L, bool& found) {
    const sps::BoxGeom g = geom();
    const int nflags = d.n[0] + d.n[1] + (d.dim == 3 ? d.n[2] : 0);
    sps::fill_int<<<(nflags + 255) / 256, 256, 0, stream>>>(occ, nflags);
    const int64_t work = (int64_t)(b[0] - 2 * sps::kB) * (b[1] - 2 * sps::kB) *
                         (d.dim == 3 ? b[2] - 2 * sps::kB : 1);
    if (d.dtype == SPS_F32)
      sps::box_occupancy<float><<<sps::grid_for(work), 256, 0, stream>>>(
          (const float*)box[cur], g, d.dim, cutoff, occ);
    else
      sps::box_occupancy<double><<<sps::grid_for(work), 256, 0, stream>>>(
          (const double*)box[cur], g, d.dim, cutoff, occ);
    SPS_TRY(sps::check_launch("occupancy"));
    std::vector<int> h((size_t)nflags);
    SPS_HIP_TRY(hipMemcpyAsync(h.data(), occ, sizeof(int) * nflags,
                               hipMemcpyDeviceToHost, stream));
    SPS_HIP_TRY(hipStreamSynchronize(stream));
    found = true;
    int off = 0;
    for (int a = 0; a < d.dim; ++a) {
      if (!sps::cover_interval(h.data() + off, d.n[a], s[a], L[a])) found = false;
      off += d.n[a];
    }
    return SPS_OK;
  }

  // New box from the active box: resize (resize_adjust_region) or re-centre
  // with the same dims (adjust_region).  Integer division truncates toward
  // zero, as in the reference.
  void place(const int* s, const int* L, bool resize, double rel_eps, int* no,
             int* nb) const {
    constexpr int B2 = 2 * sps::kB;
    bool same = true;
    for (int a = 0; a < d.dim; ++a) {
      const int dims = L[a] + B2;
      if (!resize) {
        no[a] = sps::mod_host((int64_t)s[a] - (b[a] - dims) / 2, d.n[a]);
        nb[a] = b[a];
        continue;
      }
      const int grid = d.n[a] + B2;
      int dim0 = std::min(grid, (int)(dims * rp.resize_factor));
      if (dim0 < 3 + B2) dim0 = std::min(grid, 3 + B2);
      no[a] = sps::mod_host((int64_t)s[a] - (dim0 - dims) / 2, d.n[a]);
      nb[a] = dim0;
      if (std::fabs((double)(dim0 - b[a])) / (double)b[a] > rel_eps) same = false;
    }
    if (resize && same) {
      for (int a = 0; a < d.dim; ++a) {
        no[a] = sps::mod_host((int64_t)s[a] - (b[a] - (L[a] + B2)) / 2, d.n[a]);
        nb[a] = b[a];
      }
    }
    for (int a = d.dim; a < 3; ++a) {
      no[a] = 0;
      nb[a] = 1;
    }
  }

  int move_to(const int* no, const int* nb) {
    const sps::BoxGeom go = geom();
    for (int a = 0; a < 3; ++a) {
      o[a] = no[a];
      b[a] = nb[a];
    }
    const sps::BoxGeom gn = geom();
    const int nxt = cur ^ 1;
    if (d.dtype == SPS_F32)
      sps::box_move<float><<<sps::grid_for(box_sites()), 256, 0, stream>>>(
          (const float*)box[cur], go, (float*)box[nxt], gn, d.dim, (float)empty);
    else
      sps::box_move<double><<<sps::grid_for(box_sites()), 256, 0, stream>>>(
          (const double*)box[cur], go, (double*)box[nxt], gn, d.dim, empty);
    SPS_TRY(sps::check_launch("box move"));
    cur = nxt;
    return SPS_OK;
  }

  // SolverSystemFDwSD::update(index, time): resize when due, then the
  // periodic layer of every axis the box spans.
  int update() {
    if (next_resize == 0) {
      double mn = 0.0;
      SPS_TRY(find_min(mn));
      cutoff = mn + rp.cutoff_eps;
      empty = mn;
      int s[3], L[3];
      bool found = false;
      SPS_TRY(active_box(s, L, found));
      if (!found) {
        has_box = false;
      } else {
        int no[3], nb[3];
        place(s, L, true, 0.0, no, nb);
        SPS_TRY(move_to(no, nb));
      }
      next_resize = 1;
    } else if (has_box) {
      const int current = (int)(time / rp.resize_time);
      if (current >= next_resize) {
        int s[3], L[3];
        bool found = false;
        SPS_TRY(active_box(s, L, found));
        if (found) {
          int no[3], nb[3];
          place(s, L, rp.fixed_resize == 0, rp.dims_relative_eps, no, nb);
          SPS_TRY(move_to(no, nb));
        }
        next_resize = current + 1;
      }
    }
    if (!has_box) return SPS_OK;
    const sps::BoxGeom g = geom();
    for (int a = 0; a < d.dim; ++a) {
      if (b[a] - 2 * sps::kB != d.n[a]) continue;
      if (d.dtype == SPS_F32)
        sps::box_refresh_axis<float><<<sps::grid_for(box_sites()), 256, 0, stream>>>(
            (float*)box[cur], g, a);
      else
        sps::box_refresh_axis<double><<<sps::grid_for(box_sites()), 256, 0, stream>>>(
            (double*)box[cur], g, a);
      SPS_TRY(sps::check_launch("layer refresh"));
    }
    return SPS_OK;
  }

  int step_once() {
    SPS_TRY(update());
    SPS_TRY(box_step());
    ++index;
    time += d.dt;  // run_model's time += dt (modules-models.h:489)
    return SPS_OK;
  }

  // one step of the box (no update, no time advance)
  int box_step() {
    if (has_box && fused) {
      const int nxt = cur ^ 1;
      const sps::BoxGeom g = geom();
      int rc;
      if (d.dtype == SPS_F32)
        rc = d.dim == 2 ? sps::box_fused_step<float, 2>((const float*)box[cur], (float*)mu,
                                                       (float*)box[nxt], g, fc32,
                                                       box_sites(), stream)
                        : sps::box_fused_step<float, 3>((const float*)box[cur], (float*)mu,
                                                       (float*)box[nxt], g, fc32,
                                                       box_sites(), stream);
      else
        rc = d.dim == 2 ? sps::box_fused_step<double, 2>((const double*)box[cur], (double*)mu,
                                                        (double*)box[nxt], g, fc64,
                                                        box_sites(), stream)
                        : sps::box_fused_step<double, 3>((const double*)box[cur], (double*)mu,
                                                        (double*)box[nxt], g, fc64,
                                                        box_sites(), stream);
      if (rc != SPS_OK) return rc;
      cur = nxt;
    } else if (has_box) {
      const int nxt = cur ^ 1;
      int rc;
      if (d.dtype == SPS_F32) {
        for (int a = 0; a < 3; ++a) rp32.n[a] = a < d.dim ? b[a] : 1;
        rc = sps::launch_ref_box<float>((const float*)box[cur], (float*)box[nxt],
                                        rp32, d.dim, stream);
      } else {
        for (int a = 0; a < 3; ++a) rp64.n[a] = a < d.dim ? b[a] : 1;
        rc = sps::launch_ref_box<double>((const double*)box[cur],
                                         (double*)box[nxt], rp64, d.dim, stream);
      }
      if (rc != SPS_OK) return rc;
      cur = nxt;
    }
    return SPS_OK;
  }

  // Steps per run of the fused 2D box: a box whose layer is frozen along both
  // axes (one spanning the domain refreshes its layer every step) runs 4
  // steps per launch; rp.steps_per_launch below 4 keeps one step per launch.
  int fused_run_limit() const {
    if (!fused || !has_box || d.dim != 2) return 1;
    if (rp.steps_per_launch != 0 && rp.steps_per_launch < sps::kBoxSteps) return 1;
    for (int a = 0; a < 2; ++a)
      if (b[a] - 2 * sps::kB == d.n[a]) return 1;
    return sps::kBoxSteps;
  }

  // How many of the next steps, from this one on and at most `limit`, find
  // nothing to do in update(): the first is the one whose update() just ran,
  // each later one only if its time (accumulated as step_once does) has not
  // reached the next resize count.
  int quiet_steps(int limit) const {
    int m = 1;
    double t = time;
    while (m < limit) {
      t += d.dt;
      if ((int)(t / rp.resize_time) >= next_resize) break;
      ++m;
    }
    return m;
  }

  // n steps: runs of steps with no update work between them go through the
  // multi-step box kernel (4 or 2 per launch), the rest one at a time.  Each
  // site sees the same expressions either way.
  int advance(int n) {
    while (n > 0) {
      SPS_TRY(update());
      // decided after update(): a resize may have changed the box
      const int limit = std::min(n, fused_run_limit());
      const int m = limit >= 2 ? quiet_steps(limit) : 1;
      int done = 0;
      while (m - done >= sps::kBoxSteps) {
        const int nxt = cur ^ 1;
        const sps::BoxGeom g = geom();
        const int rc =
            d.dtype == SPS_F32
                ? sps::box_multi_step<float>((const float*)box[cur], (float*)box[nxt], g,
                                             fc32, cus, stream)
                : sps::box_multi_step<double>((const double*)box[cur], (double*)box[nxt],
                                              g, fc64, cus, stream);
        if (rc != SPS_OK) return rc;
        cur = nxt;
        for (int j = 0; j < sps::kBoxSteps; ++j) {
          ++index;
          time += d.dt;
        }
        done += sps::kBoxSteps;
      }
      // the rest of the run one step per launch (their updates find nothing
      // to do, and the first step's update already ran)
      for (; done < m; ++done) {
        SPS_TRY(box_step());
        ++index;
        time += d.dt;
      }
      n -= done;
    }
    return SPS_OK;
  }

  void release() {
    for (int i = 0; i < 2; ++i)
      if (box[i]) (void)hipFree(box[i]);
    if (staging) (void)hipFree(staging);
    if (occ) (void)hipFree(occ);
    if (partial) (void)hipFree(partial);
    if (minval) (void)hipFree(minval);
    if (mu) (void)hipFree(mu);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int validate_regional(const sps_fd_regional& r) {
  if (!std::isfinite(r.cutoff_eps) || !(r.resize_factor >= 1.0) ||
      !std::isfinite(r.resize_factor) || !(r.resize_time > 0) ||
      !std::isfinite(r.resize_time) || !(r.dims_relative_eps >= 0) ||
      !std::isfinite(r.dims_relative_eps) ||
      (r.fixed_resize != 0 && r.fixed_resize != 1) || r.steps_per_launch < 0 ||
      r.steps_per_launch > 4) {
    sps::set_error("regional parameters: cutoff_eps finite, resize_factor >= 1, "
                   "resize_time > 0, dims_relative_eps >= 0, fixed_resize 0 or 1, "
                   "steps_per_launch 0..4");
    return SPS_E_INVALID;
  }
  return SPS_OK;
}

}  // namespace

extern "C" {

int sps_fdr_create(const sps_fd_desc* desc, const sps_fd_regional* params,
                   sps_fdr_plan** out) {
  using namespace sps;
  if (!out) {
    set_error("out is NULL");
    return SPS_E_INVALID;
  }
  *out = nullptr;
  SPS_TRY(validate_desc(desc));
  if (desc->nprov != 0) {
    set_error("regional plans take no provisional variables");
    return SPS_E_UNSUPPORTED;
  }
  sps_fd_regional rp{1e-6, 2.0, 3.0, 1, 0, 0.05};
  if (params) rp = *params;
  SPS_TRY(validate_regional(rp));
  if (sps_fd_device_count() < 1) {
    set_error("no HIP device: this library runs the step on an MI355X only");
    return SPS_E_HIP;
  }
  sps_fdr_plan* p = new (std::nothrow) sps_fdr_plan();
  if (!p) {
    set_error("host allocation failed");
    return SPS_E_NOMEM;
  }
  p->d = *desc;
  p->rp = rp;
  p->esize = desc->dtype == SPS_F32 ? 4 : 8;
  p->nsites = 1;
  p->cap = 1;
  for (int a = 0; a < desc->dim; ++a) {
    p->nsites *= (size_t)desc->n[a];
    p->cap *= (size_t)desc->n[a] + 2 * SPS_BOUNDARY_DEPTH;
  }
  make_ref_params(*desc, p->rp32);
  make_ref_params(*desc, p->rp64);
  // the canonical form takes the fused box step unless the reference order
  // is asked for
  const Lowered low = lower(*desc);
  if (desc->path == SPS_PATH_FUSED && !low.fusable) {
    set_error("fused path unavailable: %s", low.why_not);
    delete p;
    return SPS_E_UNSUPPORTED;
  }
  p->fused = low.fusable && desc->path != SPS_PATH_REFERENCE;
  if (p->fused) {
    make_fused_coef(*desc, low, p->fc32);
    make_fused_coef(*desc, low, p->fc64);
  }
  int rc = SPS_OK;
  auto alloc = [&](void** ptr, size_t bytes) {
    if (rc != SPS_OK) return;
    const hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
      *ptr = nullptr;
      set_error("hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
      rc = SPS_E_NOMEM;
    }
  };
  const hipError_t se = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
  if (se != hipSuccess) {
    p->stream = nullptr;
    set_error("hipStreamCreate failed: %s", hipGetErrorString(se));
    rc = SPS_E_HIP;
  }
  if (rc == SPS_OK) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, dev) !=
            hipSuccess ||
        p->cus < 1) {
      set_error("device query failed");
      rc = SPS_E_HIP;
    }
  }
  alloc(&p->box[0], p->cap * p->esize);
  alloc(&p->box[1], p->cap * p->esize);
  alloc(&p->staging, p->nsites * p->esize);
  alloc((void**)&p->occ,
        sizeof(int) * (desc->n[0] + desc->n[1] + (desc->dim == 3 ? desc->n[2] : 0)));
  alloc((void**)&p->partial, sizeof(double) * kMinBlocks);
  alloc((void**)&p->minval, sizeof(double));
  if (p->fused) alloc(&p->mu, p->cap * p->esize);
  if (rc != SPS_OK) {
    p->release();
    delete p;
    return rc;
  }
  *out = p;
  return SPS_OK;
}

int sps_fdr_destroy(sps_fdr_plan* p) {
  if (!p) return SPS_OK;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  p->release();
  delete p;
  return SPS_OK;
}

int sps_fdr_upload(sps_fdr_plan* p, const void* host, int layout) {
  using namespace sps;
  if (!p || !host) {
    set_error("plan or host pointer is NULL");
    return SPS_E_INVALID;
  }
  if (layout != SPS_LAYOUT_INTERIOR && layout != SPS_LAYOUT_PADDED) {
    set_error("layout must be SPS_LAYOUT_INTERIOR or SPS_LAYOUT_PADDED");
    return SPS_E_INVALID;
  }
  std::vector<char> tmp;
  const void* src = host;
  if (layout == SPS_LAYOUT_PADDED) {
    tmp.resize(p->nsites * p->esize);
    padded_to_interior(p->d, (const char*)host, tmp.data(), p->esize);
    src = tmp.data();
  }
  SPS_HIP_TRY(hipMemcpyAsync(p->staging, src, p->nsites * p->esize,
                             hipMemcpyHostToDevice, p->stream));
  // the box starts as the whole domain at origin 0 (the padded layout)
  for (int a = 0; a < 3; ++a) {
    p->o[a] = 0;
    p->b[a] = a < p->d.dim ? p->d.n[a] + 2 * SPS_BOUNDARY_DEPTH : 1;
  }
  p->cur = 0;
  const BoxGeom g = p->geom();
  if (p->d.dtype == SPS_F32)
    box_scatter_full<float><<<grid_for(p->box_sites()), 256, 0, p->stream>>>(
        (const float*)p->staging, g, p->d.dim, (float*)p->box[0]);
  else
    box_scatter_full<double><<<grid_for(p->box_sites()), 256, 0, p->stream>>>(
        (const double*)p->staging, g, p->d.dim, (double*)p->box[0]);
  SPS_TRY(check_launch("upload scatter"));
  SPS_HIP_TRY(hipStreamSynchronize(p->stream));
  p->has_data = true;
  p->has_box = true;
  p->next_resize = 0;
  p->index = 0;
  p->time = 0.0;
  p->empty = 0.0;
  p->cutoff = 0.0;
  return SPS_OK;
}

int sps_fdr_step(sps_fdr_plan* p, int nsteps) {
  using namespace sps;
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  if (nsteps < 0) {
    set_error("nsteps must be >= 0");
    return SPS_E_INVALID;
  }
  if (!p->has_data) {
    set_error("no field uploaded");
    return SPS_E_STATE;
  }
  return p->advance(nsteps);
}

int sps_fdr_download(sps_fdr_plan* p, void* host, int layout) {
  using namespace sps;
  if (!p || !host) {
    set_error("plan or host pointer is NULL");
    return SPS_E_INVALID;
  }
  if (layout != SPS_LAYOUT_INTERIOR && layout != SPS_LAYOUT_PADDED) {
    set_error("layout must be SPS_LAYOUT_INTERIOR or SPS_LAYOUT_PADDED");
    return SPS_E_INVALID;
  }
  if (!p->has_data) {
    set_error("no field uploaded");
    return SPS_E_STATE;
  }
  const BoxGeom g = p->geom();
  // before the first update every site is still stored
  const int has_box = p->has_box ? 1 : 0;
  if (p->d.dtype == SPS_F32)
    box_gather<float><<<grid_for((int64_t)p->nsites), 256, 0, p->stream>>>(
        (const float*)p->box[p->cur], g, p->d.dim, has_box, (float)p->empty,
        (float*)p->staging);
  else
    box_gather<double><<<grid_for((int64_t)p->nsites), 256, 0, p->stream>>>(
        (const double*)p->box[p->cur], g, p->d.dim, has_box, p->empty,
        (double*)p->staging);
  SPS_TRY(check_launch("download gather"));
  if (layout == SPS_LAYOUT_INTERIOR) {
    SPS_HIP_TRY(hipMemcpyAsync(host, p->staging, p->nsites * p->esize,
                               hipMemcpyDeviceToHost, p->stream));
    SPS_HIP_TRY(hipStreamSynchronize(p->stream));
    return SPS_OK;
  }
  std::vector<char> tmp(p->nsites * p->esize);
  SPS_HIP_TRY(hipMemcpyAsync(tmp.data(), p->staging, tmp.size(),
                             hipMemcpyDeviceToHost, p->stream));
  SPS_HIP_TRY(hipStreamSynchronize(p->stream));
  interior_to_padded(p->d, tmp.data(), (char*)host, p->esize, true);
  return SPS_OK;
}

int sps_fdr_sync(sps_fdr_plan* p) {
  using namespace sps;
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  SPS_HIP_TRY(hipStreamSynchronize(p->stream));
  return SPS_OK;
}

int sps_fdr_region(const sps_fdr_plan* p, sps_fdr_region_info* out) {
  using namespace sps;
  if (!p || !out) {
    set_error("plan or out is NULL");
    return SPS_E_INVALID;
  }
  for (int a = 0; a < 3; ++a) {
    out->origin[a] = p->o[a];
    out->dims[a] = p->has_box ? p->b[a] : 0;
  }
  out->empty = p->empty;
  out->cutoff = p->cutoff;
  out->time = p->time;
  out->index = p->index;
  out->reserved = 0;
  return SPS_OK;
}

}  // extern "C"
