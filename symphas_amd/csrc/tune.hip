// tune.hip - launch-shape autotuner of the fused single-field step.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "fused.hpp"
#include "plan.hpp"

using namespace sps;

namespace {

std::string device_tag() {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return "unknown";
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%s/%dcu", prop.gcnArchName,
                prop.multiProcessorCount);
  for (char* c = buf; *c; ++c)
    if (*c == ' ') *c = '_';
  return buf;
}

std::string plan_key(const sps_fd_plan* p) {
  const FusedCoef<double>& c = p->fc64;
  char buf[512];
  std::snprintf(buf, sizeof(buf),
                "dim=%d n=%dx%dx%d dtype=%s radius=%d a=%d deg=%d dev=%s",
                p->d.dim, p->d.n[0], p->d.n[1], p->d.dim == 3 ? p->d.n[2] : 1,
                p->d.dtype == SPS_F32 ? "f32" : "f64", c.radius,
                c.dega >= 0 ? 1 : 0,
                (c.dega > c.degb ? c.dega : c.degb) <= 3 ? 3 : 6,
                device_tag().c_str());
  return buf;
}

bool cache_lookup(const std::string& key, int* k, int* r, int* z, int* v) {
  const char* path = std::getenv("SPS_FD_TUNE_CACHE");
  if (!path || !*path) return false;
  FILE* f = std::fopen(path, "r");
  if (!f) return false;
  char line[1024];
  bool found = false;
  while (std::fgets(line, sizeof(line), f)) {
    char* bar = std::strstr(line, " | ");
    if (!bar) continue;
    if (std::string(line, bar - line) != key) continue;
    int a, b, c, e;
    if (std::sscanf(bar + 3, "%d %d %d %d", &a, &b, &c, &e) == 4) {
      *k = a;
      *r = b;
      *z = c;
      *v = e;
      found = true;  // the last entry for a key wins
    }
  }
  std::fclose(f);
  return found;
}

void cache_store(const std::string& key, int k, int r, int z, int v,
                 double us) {
  const char* path = std::getenv("SPS_FD_TUNE_CACHE");
  if (!path || !*path) return;
  FILE* f = std::fopen(path, "a");
  if (!f) return;
  std::fprintf(f, "%s | %d %d %d %d | %.3f us/step\n", key.c_str(), k, r, z, v,
               us);
  std::fclose(f);
}

}  // namespace

extern "C" int sps_fd_autotune(sps_fd_plan* p, int use_cache, int* out_k,
                               int* out_rows, int* out_planes, int* out_vec) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  auto report = [&]() {
    if (out_k) *out_k = p->steps_per_launch();
    if (out_rows) *out_rows = p->TY;
    if (out_planes) *out_planes = p->d.dim == 3 ? p->TZ : 0;
    if (out_vec) *out_vec = p->d.dim == 2 ? p->VEC : 1;
    return SPS_OK;
  };
  if (p->path != SPS_PATH_FUSED) return report();
  if (!p->has_data) {
    set_error("autotune needs an uploaded field");
    return SPS_E_STATE;
  }
  const std::string key = plan_key(p);
  int ck, cr, cz, cv;
  if (use_cache && cache_lookup(key, &ck, &cr, &cz, &cv)) {
    // a cached shape is applied only if sps_fd_tune accepts it as an explicit
    // shape (no zero meaning "keep"); a stale or corrupt entry is ignored and
    // the candidates are timed again
    const bool explicit_shape =
        ck >= 1 && cr >= 1 && (p->d.dim == 2 ? (cv >= 1 && cz == 0) : cz >= 1);
    if (explicit_shape &&
        sps_fd_tune(p, ck, cr, p->d.dim == 3 ? cz : 0,
                    p->d.dim == 2 ? cv : 0) == SPS_OK)
      return report();
    set_error("");
  }
  struct Cand {
    int k, rows, planes, vec;
  };
  std::vector<Cand> cands;
  if (p->d.dim == 2) {
    // wide lanes (16 bytes) with short chunks, narrow lanes (8 bytes, twice
    // the strips) with tall ones; plus, per lane width and K, the chunk
    // heights that give every CU the same number m of 4-wave workgroups (a
    // CU with one workgroup more than the others sets the launch's length)
    const int wide = 16 / (int)p->esize, narrow = 8 / (int)p->esize;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                              dev) != hipSuccess) {
      set_error("device query failed");
      return SPS_E_HIP;
    }
    const int X = p->d.n[0], Y = p->d.n[1], R = p->fc64.radius;
    for (int k = 1; k <= 4; ++k) {
      for (int vec : {wide, narrow}) {
        if (X % vec != 0) continue;
        std::vector<int> rows = vec == wide ? std::vector<int>{8, 16, 32, 64}
                                            : std::vector<int>{16, 32, 64, 128};
        const int hl = (k * R + vec - 1) / vec;
        const int nstrips = (X + (64 - 2 * hl) * vec - 1) / ((64 - 2 * hl) * vec);
        for (int m = 1; m <= 6; ++m) {
          const long nch = (long)m * cus * 4 / nstrips;
          if (nch < 1) continue;
          const long ty = (Y + nch - 1) / nch;
          if (ty >= 4 && ty <= 512 &&
              std::find(rows.begin(), rows.end(), (int)ty) == rows.end())
            rows.push_back((int)ty);
        }
        for (int r : rows) cands.push_back({k, r, 0, vec});
      }
    }
  } else {
    // chunk heights: short chunks, and Z split evenly into m chunks, which
    // puts the whole grid in one or a few rounds of resident waves
    std::vector<int> planes = {8, 16, 32};
    const int Z = p->d.n[2];
    for (int m = 1; m <= 8; ++m) {
      const int z = (Z + m - 1) / m;
      if (z > 32 && z <= 1024 &&
          std::find(planes.begin(), planes.end(), z) == planes.end())
        planes.push_back(z);
    }
    for (int k : {1, 2})
      for (int r : {4, 6, 8, 12, 16})
        for (int z : planes)
          if ((k == 1 && r > 4 && sps::fused3d_shape_ok((int)p->esize, k, r)) ||
              (k == 2 && sps::fused3d_shape_ok((int)p->esize, k, r)))
            cands.push_back({k, r, z, 1});
  }
  const size_t bytes = p->nsites * p->esize;
  void* saved = nullptr;
  SPS_HIP_TRY(hipMalloc(&saved, bytes));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = SPS_OK;
  const int oldK = p->K, oldTY = p->TY, oldTZ = p->TZ, oldVEC = p->VEC;
  Cand best{oldK, oldTY, oldTZ, oldVEC};
  double best_us = 1e300;
  do {
    if (hipMemcpyAsync(saved, p->buf[p->cur], bytes, hipMemcpyDeviceToDevice,
                       p->stream) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
      set_error("autotune set-up failed");
      rc = SPS_E_HIP;
      break;
    }
    for (const Cand& c : cands) {
      p->K = c.k;
      p->TY = c.rows;
      if (p->d.dim == 3) p->TZ = c.planes;
      if (p->d.dim == 2) p->VEC = c.vec;
      const int spl = p->steps_per_launch();
      // about 12 steps (a multiple of every K) per trial, twice
      const int launches = p->d.dim == 3 ? 4 : 12 / spl;
      double us = 1e300;
      for (int rep = 0; rep < 3 && rc == SPS_OK; ++rep) {
        if ((rc = p->launch_one(spl)) != SPS_OK) break;  // warm
        if (hipEventRecord(e0, p->stream) != hipSuccess) rc = SPS_E_HIP;
        for (int i = 0; i < launches && rc == SPS_OK; ++i)
          rc = p->launch_one(spl);
        if (rc != SPS_OK || hipEventRecord(e1, p->stream) != hipSuccess ||
            hipEventSynchronize(e1) != hipSuccess) {
          if (rc == SPS_OK) rc = SPS_E_HIP;
          break;
        }
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          rc = SPS_E_HIP;
          break;
        }
        if (rep > 0) us = std::min(us, 1e3 * ms / (launches * spl));
      }
      if (rc != SPS_OK) break;
      if (us < best_us) {
        best_us = us;
        best = c;
      }
    }
    // restore the field the caller had
    if (hipMemcpyAsync(p->buf[p->cur], saved, bytes, hipMemcpyDeviceToDevice,
                       p->stream) != hipSuccess ||
        hipStreamSynchronize(p->stream) != hipSuccess) {
      if (rc == SPS_OK) {
        set_error("autotune restore failed");
        rc = SPS_E_HIP;
      }
    }
  } while (false);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(saved);
  if (rc != SPS_OK) {
    p->K = oldK;
    p->TY = oldTY;
    p->TZ = oldTZ;
    p->VEC = oldVEC;
    p->drop_graphs();
    return rc;
  }
  p->K = best.k;
  p->TY = best.rows;
  if (p->d.dim == 3) p->TZ = best.planes;
  if (p->d.dim == 2) p->VEC = best.vec;
  p->drop_graphs();
  cache_store(key, p->K, p->TY, p->d.dim == 3 ? p->TZ : 0,
              p->d.dim == 2 ? p->VEC : 1, best_us);
  return report();
}
