// api.hip - the C ABI declared in include/sps_fd.h (single-field plans).
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"
#include "fused.hpp"
#include "plan.hpp"
#include "ref_kernels.hpp"

namespace sps {

namespace {
thread_local char g_err[1024] = "";
}

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char* last_error() { return g_err; }

static bool valid_points(const sps_fd_desc& d) {
  const int L = d.lap_points, G = d.gradlap_points, B = d.bilap_points;
  if (d.dim == 2)
    return (L == 5 || L == 9) && (G == 6 || G == 8 || G == 12 || G == 16) &&
           (B == 13 || B == 17 || B == 21);
  return (L == 7 || L == 15 || L == 19 || L == 21 || L == 27) &&
         (G == 10 || G == 12 || G == 28 || G == 36 || G == 40) &&
         (B == 21 || B == 25 || B == 41 || B == 52 || B == 57);
}

int validate_desc(const sps_fd_desc* d) {
  if (!d) {
    set_error("descriptor is NULL");
    return SPS_E_INVALID;
  }
  if (d->dim != 2 && d->dim != 3) {
    set_error("dim must be 2 or 3 (got %d)", d->dim);
    return SPS_E_INVALID;
  }
  if (d->dtype != SPS_F32 && d->dtype != SPS_F64) {
    set_error("dtype must be SPS_F32 or SPS_F64 (got %d)", d->dtype);
    return SPS_E_INVALID;
  }
  for (int a = 0; a < d->dim; ++a) {
    // the reference's periodic ghost copy needs at least BOUNDARY_DEPTH
    // interior cells per axis (datatypes/inc/boundaryupdate.h:650-708)
    if (d->n[a] < SPS_BOUNDARY_DEPTH || d->n[a] > (1 << 20)) {
      set_error("n[%d] = %d out of range [%d, 2^20]", a, d->n[a],
                SPS_BOUNDARY_DEPTH);
      return SPS_E_INVALID;
    }
  }
  double sites = (double)d->n[0] * d->n[1] * (d->dim == 3 ? d->n[2] : 1);
  if (sites > 2147483647.0) {
    set_error("field of %.0f sites exceeds 2^31-1", sites);
    return SPS_E_INVALID;
  }
  if (!(d->h > 0) || !std::isfinite(d->h) || !std::isfinite(d->dt)) {
    set_error("h must be positive and finite, dt finite");
    return SPS_E_INVALID;
  }
  if (!valid_points(*d)) {
    set_error("no Stencil%dd2h<%d,%d,%d> in the reference's point lists",
              d->dim, d->lap_points, d->gradlap_points, d->bilap_points);
    return SPS_E_INVALID;
  }
  if (d->nterms < 0 || d->nterms > SPS_MAX_TERMS) {
    set_error("nterms %d not in [0, %d]", d->nterms, SPS_MAX_TERMS);
    return SPS_E_INVALID;
  }
  if (d->nprov < 0 || d->nprov > SPS_MAX_PROV) {
    set_error("nprov %d not in [0, %d]", d->nprov, SPS_MAX_PROV);
    return SPS_E_INVALID;
  }
  // group 0: the right-hand side (operands psi or any var); group g >= 1:
  // the definition of var(g) (operands psi or var(j), j < g)
  for (int g = 0; g <= d->nprov; ++g) {
    const sps_fd_term* terms = g == 0 ? d->terms : d->prov[g - 1];
    const int nt = g == 0 ? d->nterms : d->prov_nterms[g - 1];
    const char* what = g == 0 ? "term" : "provisional term";
    if (g > 0 && (nt < 1 || nt > SPS_MAX_TERMS)) {
      set_error("var(%d): %d terms, not in [1, %d]", g, nt, SPS_MAX_TERMS);
      return SPS_E_INVALID;
    }
    for (int t = 0; t < nt; ++t) {
      const sps_fd_term& tm = terms[t];
      if (tm.op < SPS_OP_ID || tm.op > SPS_OP_GRADLAP) {
        set_error("%s %d: unknown op %d", what, t, tm.op);
        return SPS_E_INVALID;
      }
      const int max_operand = g == 0 ? d->nprov : g - 1;
      if (tm.operand < 0 || tm.operand > max_operand) {
        set_error("%s %d: operand %d not in [0, %d]", what, t, tm.operand,
                  max_operand);
        return SPS_E_INVALID;
      }
      if (tm.op == SPS_OP_GRADLAP) {
        if (tm.axis < 0 || tm.axis >= d->dim) {
          set_error("%s %d: gradlaplacian axis %d", what, t, tm.axis);
          return SPS_E_INVALID;
        }
        const int ncomp = (d->dim == 2 && d->gradlap_points == 6) ? 2 : 1;
        if (tm.component < 0 || tm.component >= ncomp) {
          set_error("%s %d: component %d (stencil has %d)", what, t,
                    tm.component, ncomp);
          return SPS_E_INVALID;
        }
      }
      if (!std::isfinite(tm.coeff)) {
        set_error("%s %d: coefficient not finite", what, t);
        return SPS_E_INVALID;
      }
      for (int k = 0; k <= SPS_MAX_DEGREE; ++k)
        if (!std::isfinite(tm.poly[k])) {
          set_error("%s %d: polynomial coefficient %d not finite", what, t, k);
          return SPS_E_INVALID;
        }
    }
  }
  if (d->path < SPS_PATH_AUTO || d->path > SPS_PATH_FUSED) {
    set_error("path %d unknown", d->path);
    return SPS_E_INVALID;
  }
  return SPS_OK;
}

// Lower the descriptor's terms into dpsi/dt = A + lap(B) + kc*bilap(psi).
// Provisional variables are substituted first.  A term c*D[var(k)] (D the
// identity or the laplacian, operand exactly var(k)) expands into c times
// D applied to each term of var(k)'s definition, whose operands must be psi:
//   D = id:  v*E[P(psi)]  ->  c*v*E[P(psi)]
//   D = lap: v*P(psi)     ->  c*v*lap[P(psi)]
//            v*lap[a*psi] ->  c*v*a*bilap[psi]   (lap5∘lap5 = bilap13 and
//                             lap7∘lap7 = bilap25, so the fused kernels'
//                             two 5/7-point passes are the reference's
//                             evaluation of lap(var) with var = lap psi + ...)
// Anything else keeps the reference-order path.
Lowered lower(const sps_fd_desc& d) {
  Lowered lw;
  const int lapL = d.dim == 2 ? 5 : 7;
  const int bilB = d.dim == 2 ? 13 : 25;
  bool any_lap = false, any_bilap = false;
  std::vector<sps_fd_term> flat;
  for (int t = 0; t < d.nterms; ++t) {
    const sps_fd_term& tm = d.terms[t];
    if (tm.operand == 0) {
      flat.push_back(tm);
      continue;
    }
    bool linear = tm.poly[1] == 1.0;
    for (int k = 0; k <= kMaxDeg; ++k)
      if (k != 1 && tm.poly[k] != 0.0) linear = false;
    if (!linear || (tm.op != SPS_OP_ID && tm.op != SPS_OP_LAP)) {
      lw.why_not = "a provisional variable enters other than as c*var or "
                   "c*lap(var)";
      return lw;
    }
    const int k = tm.operand;
    for (int j = 0; j < d.prov_nterms[k - 1]; ++j) {
      sps_fd_term v = d.prov[k - 1][j];
      if (v.operand != 0) {
        lw.why_not = "a provisional variable defined through another";
        return lw;
      }
      v.coeff *= tm.coeff;
      if (tm.op == SPS_OP_LAP) {
        if (v.op == SPS_OP_ID) {
          v.op = SPS_OP_LAP;
        } else if (v.op == SPS_OP_LAP) {
          for (int q = 2; q <= kMaxDeg; ++q)
            if (v.poly[q] != 0.0) {
              lw.why_not = "laplacian of a laplacian of a non-linear operand";
              return lw;
            }
          any_lap = true;  // needs the 5/7-point laplacian, not bilap_points
          lw.kc += v.coeff * v.poly[1];
          continue;
        } else {
          lw.why_not = "derivative of order above four";
          return lw;
        }
      }
      flat.push_back(v);
    }
  }
  for (const sps_fd_term& tm : flat) {
    switch (tm.op) {
      case SPS_OP_ID:
        for (int k = 0; k <= kMaxDeg; ++k) lw.a[k] += tm.coeff * tm.poly[k];
        break;
      case SPS_OP_LAP:
        any_lap = true;
        for (int k = 1; k <= kMaxDeg; ++k) lw.b[k] += tm.coeff * tm.poly[k];
        break;
      case SPS_OP_BILAP:
        any_bilap = true;
        for (int k = 2; k <= kMaxDeg; ++k)
          if (tm.poly[k] != 0.0) {
            lw.why_not = "bilaplacian of a non-linear operand";
            return lw;
          }
        lw.kc += tm.coeff * tm.poly[1];
        break;
      default:
        lw.why_not = "gradlaplacian terms";
        return lw;
    }
  }
  if (any_lap && d.lap_points != lapL) {
    lw.why_not = d.dim == 2 ? "laplacian stencil is not the 5-point one"
                            : "laplacian stencil is not the 7-point one";
    return lw;
  }
  if (any_bilap && d.bilap_points != bilB) {
    lw.why_not = d.dim == 2 ? "bilaplacian stencil is not the 13-point one"
                            : "bilaplacian stencil is not the 25-point one";
    return lw;
  }
  lw.fusable = true;
  return lw;
}

template <typename T>
void make_fused_coef(const sps_fd_desc& d, const Lowered& lw, FusedCoef<T>& c) {
  const double divh = 1.0 / d.h;
  const double divh2 = divh * divh;
  const double w = 2.0 * d.dim;  // centre weight of the 5/7-point difference
  int dega = -1, degb = -1;
  for (int k = 0; k <= kMaxDeg; ++k) {
    if (lw.a[k] != 0.0) dega = k;
    if (lw.b[k] != 0.0) degb = k;
  }
  const double kmu = lw.kc * divh2, kd = d.dt * divh2, k1 = kd * lw.b[1];
  c.dega = dega;
  c.degb = degb;
  c.radius = (lw.kc == 0.0 && degb <= 1) ? 1 : 2;
  // mu~ = sigma*mu: sigma = 1/kmu, or 1 when mu has no laplacian term
  const double sigma = kmu != 0.0 ? 1.0 / kmu : 1.0;
  for (int k = 0; k < kMaxDeg; ++k) c.p[k] = T(sigma * lw.b[k + 1]);
  c.p[0] = T(sigma * (lw.b[1] - w * kmu));
  for (int k = 0; k <= kMaxDeg; ++k) c.q[k] = T(d.dt * lw.a[k]);
  c.ks = T(kmu != 0.0 ? 1.0 : 0.0);
  c.k1 = T(k1);
  c.k2 = T(kd / sigma);
  const int deg = dega > degb ? dega : degb;
  c.poly = (deg <= 3 && (c.radius == 1 || kmu != 0.0)) ? 3 : 6;
}

// Reference-order parameters for one term list: group 0 is the right-hand
// side, group k >= 1 the definition of var(k).
template <typename T>
void make_ref_params(const sps_fd_desc& d, RefParams<T>& p, int group) {
  std::memset(&p, 0, sizeof(p));
  for (int a = 0; a < 3; ++a) p.n[a] = a < d.dim ? d.n[a] : 1;
  p.L = d.lap_points;
  p.G = d.gradlap_points;
  p.B = d.bilap_points;
  const sps_fd_term* terms = group == 0 ? d.terms : d.prov[group - 1];
  p.nterms = group == 0 ? d.nterms : d.prov_nterms[group - 1];
  // divh = 1/h; divh2..4 as StencilBase{2,3}d2h builds them
  // (sol/inc/stencilincludes.h:100-104,123-127)
  const double divh = 1.0 / d.h;
  const double divh2 = divh * divh;
  const double divh3 = divh * divh2;
  const double divh4 = divh2 * divh2;
  p.divh2 = T(divh2);
  p.divh3 = T(divh3);
  p.divh4 = T(divh4);
  p.dt = T(d.dt);
  for (int t = 0; t < p.nterms; ++t) {
    const sps_fd_term& tm = terms[t];
    DevTerm<T>& dt = p.terms[t];
    dt.op = tm.op;
    dt.axis = tm.axis;
    dt.component = tm.component;
    dt.operand = tm.operand;
    bool ident = tm.poly[1] == 1.0;
    for (int k = 0; k <= kMaxDeg; ++k) {
      dt.poly[k] = T(tm.poly[k]);
      if (k != 1 && tm.poly[k] != 0.0) ident = false;
    }
    dt.identity = ident ? 1 : 0;
    dt.coeff = T(tm.coeff);
  }
}

template void make_ref_params<float>(const sps_fd_desc&, RefParams<float>&,
                                     int);
template void make_ref_params<double>(const sps_fd_desc&, RefParams<double>&,
                                      int);
template void make_fused_coef<float>(const sps_fd_desc&, const Lowered&,
                                     FusedCoef<float>&);
template void make_fused_coef<double>(const sps_fd_desc&, const Lowered&,
                                      FusedCoef<double>&);

}  // namespace sps

using namespace sps;

// ----------------------------------------------------------------- plan

static size_t sites_of(const sps_fd_desc& d) {
  size_t n = (size_t)d.n[0] * d.n[1];
  if (d.dim == 3) n *= d.n[2];
  return n;
}

static size_t padded_sites(const sps_fd_desc& d) {
  size_t n = (size_t)(d.n[0] + 2 * SPS_BOUNDARY_DEPTH) *
             (d.n[1] + 2 * SPS_BOUNDARY_DEPTH);
  if (d.dim == 3) n *= d.n[2] + 2 * SPS_BOUNDARY_DEPTH;
  return n;
}

namespace sps {

// interior (unpadded) <-> SymPhas padded layout, element size es bytes
void interior_to_padded(const sps_fd_desc& d, const char* src, char* dst,
                        size_t es, bool fill_ghosts) {
  const int D = SPS_BOUNDARY_DEPTH;
  const int64_t X = d.n[0] + 2 * D, Y = d.n[1] + 2 * D;
  const int64_t Z = d.dim == 3 ? d.n[2] + 2 * D : 1;
  const int64_t nx = d.n[0], ny = d.n[1], nz = d.dim == 3 ? d.n[2] : 1;
  for (int64_t z = 0; z < Z; ++z)
    for (int64_t y = 0; y < Y; ++y)
      for (int64_t x = 0; x < X; ++x) {
        const bool ghost = x < D || x >= X - D || y < D || y >= Y - D ||
                           (d.dim == 3 && (z < D || z >= Z - D));
        char* out = dst + es * (x + X * (y + Y * z));
        if (ghost && !fill_ghosts) {
          std::memset(out, 0, es);
          continue;
        }
        int64_t ix = ((x - D) % nx + nx) % nx;
        int64_t iy = ((y - D) % ny + ny) % ny;
        int64_t iz = d.dim == 3 ? ((z - D) % nz + nz) % nz : 0;
        std::memcpy(out, src + es * (ix + nx * (iy + ny * iz)), es);
      }
}

void padded_to_interior(const sps_fd_desc& d, const char* src, char* dst,
                        size_t es) {
  const int D = SPS_BOUNDARY_DEPTH;
  const int64_t X = d.n[0] + 2 * D, Y = d.n[1] + 2 * D;
  const int64_t nx = d.n[0], ny = d.n[1], nz = d.dim == 3 ? d.n[2] : 1;
  const int64_t zo = d.dim == 3 ? D : 0;
  for (int64_t z = 0; z < nz; ++z)
    for (int64_t y = 0; y < ny; ++y)
      std::memcpy(dst + es * (nx * (y + ny * z)),
                  src + es * (D + X * ((y + D) + Y * (z + zo))), es * nx);
}

}  // namespace sps

int sps_fd_plan::launch_provisionals() {
  const void* in = buf[cur];
  for (int k = 0; k < d.nprov; ++k) {
    int rc = d.dtype == SPS_F32
                 ? launch_ref<float>((const float*)in, (float*)vars[k],
                                     prov32[k], d.dim, stream)
                 : launch_ref<double>((const double*)in, (double*)vars[k],
                                      prov64[k], d.dim, stream);
    if (rc != SPS_OK) return rc;
  }
  return SPS_OK;
}

int sps_fd_plan::launch_one(int k) {
  const void* in = buf[cur];
  void* out = buf[1 - cur];
  int rc;
  if (path == SPS_PATH_REFERENCE) {
    // provisional variables first, from the current field (Model::update,
    // sol/inc/modelspecialized.h:430-434), then the step reading them
    if ((rc = launch_provisionals()) != SPS_OK) return rc;
    rc = d.dtype == SPS_F32
             ? launch_ref<float>((const float*)in, (float*)out, rp32, d.dim,
                                 stream)
             : launch_ref<double>((const double*)in, (double*)out, rp64,
                                  d.dim, stream);
  } else if (d.dim == 2) {
    rc = d.dtype == SPS_F32
             ? launch_fused2d<float>((const float*)in, (float*)out, d.n[0],
                                     d.n[1], k, TY, VEC, fc32, stream)
             : launch_fused2d<double>((const double*)in, (double*)out, d.n[0],
                                      d.n[1], k, TY, VEC, fc64, stream);
  } else {
    rc = d.dtype == SPS_F32
             ? launch_fused3d<float>((const float*)in, (float*)out, d.n[0],
                                     d.n[1], d.n[2], k, TY, TZ, fc32, stream)
             : launch_fused3d<double>((const double*)in, (double*)out, d.n[0],
                                      d.n[1], d.n[2], k, TY, TZ, fc64, stream);
  }
  if (rc == SPS_OK) cur = 1 - cur;
  return rc;
}

int sps_fd_plan::steps_per_launch() const {
  return path == SPS_PATH_FUSED ? K : 1;
}

void sps_fd_plan::drop_graphs() {
  for (auto& g : graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g = GraphSlot{};
  }
}

int sps_fd_plan::step(int nsteps) {
  const int spl = steps_per_launch();
  int launches = nsteps / spl;
  const int rem = nsteps % spl;
  // the null stream cannot be captured: launch one by one there
  if (launches >= 2 * kGraphLaunches && stream != nullptr) {
    GraphSlot& g = graphs[cur];
    if (!g.exec) {
      const int start = cur;
      SPS_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
      int rc = SPS_OK;
      for (int i = 0; i < kGraphLaunches && rc == SPS_OK; ++i)
        rc = launch_one(spl);
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamEndCapture(stream, &graph);
      cur = start;  // capture records work; nothing ran
      if (rc != SPS_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      if (e != hipSuccess) {
        set_error("graph capture failed: %s", hipGetErrorString(e));
        return SPS_E_HIP;
      }
      hipGraphExec_t exec = nullptr;
      e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        set_error("graph instantiate failed: %s", hipGetErrorString(e));
        return SPS_E_HIP;
      }
      g.graph = graph;
      g.exec = exec;
    }
    while (launches >= kGraphLaunches) {
      SPS_HIP_TRY(hipGraphLaunch(g.exec, stream));
      launches -= kGraphLaunches;  // even count: parity unchanged
    }
  }
  for (int i = 0; i < launches; ++i) SPS_TRY(launch_one(spl));
  if (rem) SPS_TRY(launch_one(rem));
  return SPS_OK;
}

// ------------------------------------------------------------------ C ABI

extern "C" {

const char* sps_fd_last_error(void) { return sps::last_error(); }
int sps_fd_abi_version(void) { return SPS_FD_ABI_VERSION; }
size_t sps_fd_desc_size(void) { return sizeof(sps_fd_desc); }

int sps_fd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int sps_fd_create(const sps_fd_desc* desc, sps_fd_plan** out) {
  if (!out) {
    set_error("out is NULL");
    return SPS_E_INVALID;
  }
  *out = nullptr;
  SPS_TRY(validate_desc(desc));
  if (sps_fd_device_count() < 1) {
    set_error("no HIP device: this library runs the step on an MI355X only");
    return SPS_E_HIP;
  }
  sps_fd_plan* p = new (std::nothrow) sps_fd_plan();
  if (!p) {
    set_error("host allocation failed");
    return SPS_E_NOMEM;
  }
  p->d = *desc;
  p->low = lower(*desc);
  const int vec = desc->dtype == SPS_F32 ? 2 : 1;  // narrowest lane width
  const bool fusable =
      p->low.fusable && (desc->dim == 3 || desc->n[0] % vec == 0);
  if (desc->path == SPS_PATH_FUSED && !fusable) {
    set_error("fused path unavailable: %s",
              p->low.fusable ? "x extent not a multiple of the vector width"
                             : p->low.why_not);
    delete p;
    return SPS_E_UNSUPPORTED;
  }
  p->path = (desc->path == SPS_PATH_REFERENCE || !fusable)
                ? SPS_PATH_REFERENCE
                : SPS_PATH_FUSED;
  make_ref_params(*desc, p->rp32);
  make_ref_params(*desc, p->rp64);
  for (int k = 0; k < desc->nprov; ++k) {
    make_ref_params(*desc, p->prov32[k], k + 1);
    make_ref_params(*desc, p->prov64[k], k + 1);
    p->prov32[k].mode = 1;  // out = sum of the terms
    p->prov64[k].mode = 1;
  }
  make_fused_coef(*desc, p->low, p->fc32);
  make_fused_coef(*desc, p->low, p->fc64);
  p->nsites = sites_of(*desc);
  p->esize = desc->dtype == SPS_F32 ? 4 : 8;
  p->choose_default_shape();
  (void)make_energy_params(nullptr, p->d.h, p->energy);  // the default is valid
  int rc = p->allocate();
  if (rc != SPS_OK) {
    sps_fd_destroy(p);
    return rc;
  }
  *out = p;
  return SPS_OK;
}

int sps_fd_destroy(sps_fd_plan* p) {
  if (!p) return SPS_OK;
  p->drop_graphs();
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  if (p->own_stream && p->own_stream != p->stream)
    (void)hipStreamSynchronize(p->own_stream);
  for (void* b : p->buf)
    if (b) (void)hipFree(b);
  if (p->work) (void)hipFree(p->work);
  for (void* v : p->vars)
    if (v) (void)hipFree(v);
  if (p->partials) (void)hipFree(p->partials);
  if (p->sums) (void)hipFree(p->sums);
  if (p->own_stream) (void)hipStreamDestroy(p->own_stream);
  delete p;
  return SPS_OK;
}

int sps_fd_upload(sps_fd_plan* p, const void* host, int layout) {
  if (!p || !host) {
    set_error("plan or host pointer is NULL");
    return SPS_E_INVALID;
  }
  const size_t bytes = p->nsites * p->esize;
  if (layout == SPS_LAYOUT_INTERIOR) {
    SPS_HIP_TRY(hipMemcpyAsync(p->buf[p->cur], host, bytes,
                               hipMemcpyHostToDevice, p->stream));
  } else if (layout == SPS_LAYOUT_PADDED) {
    std::vector<char> tmp(bytes);
    padded_to_interior(p->d, (const char*)host, tmp.data(), p->esize);
    SPS_HIP_TRY(hipMemcpyAsync(p->buf[p->cur], tmp.data(), bytes,
                               hipMemcpyHostToDevice, p->stream));
    SPS_HIP_TRY(hipStreamSynchronize(p->stream));
  } else {
    set_error("unknown layout %d", layout);
    return SPS_E_INVALID;
  }
  SPS_HIP_TRY(hipStreamSynchronize(p->stream));
  p->has_data = true;
  return SPS_OK;
}

static int download_buffer(sps_fd_plan* p, const void* dev, void* host,
                           int layout, bool fill_ghosts) {
  const size_t bytes = p->nsites * p->esize;
  if (layout == SPS_LAYOUT_INTERIOR) {
    SPS_HIP_TRY(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost,
                               p->stream));
    SPS_HIP_TRY(hipStreamSynchronize(p->stream));
    return SPS_OK;
  }
  if (layout != SPS_LAYOUT_PADDED) {
    set_error("unknown layout %d", layout);
    return SPS_E_INVALID;
  }
  std::vector<char> tmp(bytes);
  SPS_HIP_TRY(hipMemcpyAsync(tmp.data(), dev, bytes, hipMemcpyDeviceToHost,
                             p->stream));
  SPS_HIP_TRY(hipStreamSynchronize(p->stream));
  interior_to_padded(p->d, tmp.data(), (char*)host, p->esize, fill_ghosts);
  return SPS_OK;
}

int sps_fd_download(sps_fd_plan* p, void* host, int layout) {
  if (!p || !host) {
    set_error("plan or host pointer is NULL");
    return SPS_E_INVALID;
  }
  if (!p->has_data) {
    set_error("no field uploaded");
    return SPS_E_STATE;
  }
  return download_buffer(p, p->buf[p->cur], host, layout, true);
}

int sps_fd_step(sps_fd_plan* p, int nsteps) {
  if (!p || nsteps < 0) {
    set_error("plan is NULL or nsteps < 0");
    return SPS_E_INVALID;
  }
  if (!p->has_data) {
    set_error("no field uploaded");
    return SPS_E_STATE;
  }
  return p->step(nsteps);
}

static int run_ref_mode(sps_fd_plan* p, int mode, int op, int axis,
                        int component) {
  if (!p->work) SPS_HIP_TRY(hipMalloc(&p->work, p->nsites * p->esize));
  if (mode == 1) SPS_TRY(p->launch_provisionals());
  if (p->d.dtype == SPS_F32) {
    RefParams<float> rp = p->rp32;
    rp.mode = mode;
    rp.apply_op = op;
    rp.apply_axis = axis;
    rp.apply_component = component;
    return launch_ref<float>((const float*)p->buf[p->cur], (float*)p->work,
                             rp, p->d.dim, p->stream);
  }
  RefParams<double> rp = p->rp64;
  rp.mode = mode;
  rp.apply_op = op;
  rp.apply_axis = axis;
  rp.apply_component = component;
  return launch_ref<double>((const double*)p->buf[p->cur], (double*)p->work,
                            rp, p->d.dim, p->stream);
}

int sps_fd_equation(sps_fd_plan* p, void* host, int layout) {
  if (!p || !host) {
    set_error("plan or host pointer is NULL");
    return SPS_E_INVALID;
  }
  if (!p->has_data) {
    set_error("no field uploaded");
    return SPS_E_STATE;
  }
  SPS_TRY(run_ref_mode(p, 1, 0, 0, 0));
  return download_buffer(p, p->work, host, layout, false);
}

int sps_fd_apply(sps_fd_plan* p, int op, int axis, int component, void* host,
                 int layout) {
  if (!p || !host) {
    set_error("plan or host pointer is NULL");
    return SPS_E_INVALID;
  }
  if (!p->has_data) {
    set_error("no field uploaded");
    return SPS_E_STATE;
  }
  if (op != SPS_OP_LAP && op != SPS_OP_BILAP && op != SPS_OP_GRADLAP) {
    set_error("apply: op must be LAP, BILAP or GRADLAP");
    return SPS_E_INVALID;
  }
  if (op == SPS_OP_GRADLAP) {
    const int ncomp = (p->d.dim == 2 && p->d.gradlap_points == 6) ? 2 : 1;
    if (axis < 0 || axis >= p->d.dim || component < 0 || component >= ncomp) {
      set_error("apply: gradlaplacian axis %d / component %d invalid", axis,
                component);
      return SPS_E_INVALID;
    }
  }
  SPS_TRY(run_ref_mode(p, 2, op, axis, component));
  return download_buffer(p, p->work, host, layout, false);
}

int sps_fd_reduce(sps_fd_plan* p, double* mass, double* energy) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  if (!p->has_data) {
    set_error("no field uploaded");
    return SPS_E_STATE;
  }
  int rc = p->d.dtype == SPS_F32
               ? launch_reduce<float>((const float*)p->buf[p->cur], p->d.dim,
                                      p->d.n, p->d.h, p->energy, p->partials,
                                      p->sums, p->stream)
               : launch_reduce<double>((const double*)p->buf[p->cur],
                                       p->d.dim, p->d.n, p->d.h, p->energy,
                                       p->partials, p->sums, p->stream);
  if (rc != SPS_OK) return rc;
  double h[2];
  SPS_HIP_TRY(hipMemcpyAsync(h, p->sums, sizeof(h), hipMemcpyDeviceToHost,
                             p->stream));
  SPS_HIP_TRY(hipStreamSynchronize(p->stream));
  if (mass) *mass = h[0];
  if (energy) *energy = h[1];
  return SPS_OK;
}

int sps_fd_set_energy(sps_fd_plan* p, const sps_fd_energy* energy) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  EnergyParams ep;
  SPS_TRY(make_energy_params(energy, p->d.h, ep));
  p->energy = ep;
  return SPS_OK;
}

int sps_fd_sync(sps_fd_plan* p) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  SPS_HIP_TRY(hipStreamSynchronize(p->stream));
  return SPS_OK;
}

int sps_fd_set_stream(sps_fd_plan* p, void* hip_stream) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  // order the switch after the work already on the old stream
  SPS_HIP_TRY(hipStreamSynchronize(p->stream));
  p->stream = (hipStream_t)hip_stream;
  return SPS_OK;
}

int sps_fd_copy_device(sps_fd_plan* p, void* dev, int direction) {
  if (!p || !dev || (direction != 0 && direction != 1)) {
    set_error("copy_device: plan, device pointer and direction 0/1 required");
    return SPS_E_INVALID;
  }
  const size_t bytes = p->nsites * p->esize;
  if (direction == 0) {
    if (!p->has_data) {
      set_error("no field uploaded");
      return SPS_E_STATE;
    }
    SPS_HIP_TRY(hipMemcpyAsync(dev, p->buf[p->cur], bytes,
                               hipMemcpyDeviceToDevice, p->stream));
  } else {
    SPS_HIP_TRY(hipMemcpyAsync(p->buf[p->cur], dev, bytes,
                               hipMemcpyDeviceToDevice, p->stream));
    p->has_data = true;
  }
  return SPS_OK;
}

int sps_fd_device_field(sps_fd_plan* p, void** dptr) {
  if (!p || !dptr) {
    set_error("plan or out pointer is NULL");
    return SPS_E_INVALID;
  }
  *dptr = p->buf[p->cur];
  return SPS_OK;
}

int sps_fd_query(const sps_fd_plan* p, int* path, int* spl) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  if (path) *path = p->path;
  if (spl) *spl = p->steps_per_launch();
  return SPS_OK;
}

int sps_fd_tune(sps_fd_plan* p, int steps_per_launch, int rows_per_tile,
                int planes_per_chunk, int values_per_lane) {
  if (!p) {
    set_error("plan is NULL");
    return SPS_E_INVALID;
  }
  if (p->d.dim == 2) {
    const int wide = 16 / (int)p->esize, narrow = 8 / (int)p->esize;
    const int vec = values_per_lane ? values_per_lane : p->VEC;
    if (steps_per_launch < 0 || steps_per_launch > 4 || planes_per_chunk != 0 ||
        (rows_per_tile != 0 && (rows_per_tile < 4 || rows_per_tile > 512)) ||
        (vec != wide && vec != narrow) || p->d.n[0] % vec != 0) {
      set_error("tune (2D): steps_per_launch in 0..4, rows_per_tile 0 or "
                "4..512, planes_per_chunk 0, values_per_lane 0, %d or %d "
                "dividing the x extent", wide, narrow);
      return SPS_E_INVALID;
    }
    if (steps_per_launch) p->K = steps_per_launch;
    if (rows_per_tile) p->TY = rows_per_tile;
    p->VEC = vec;
  } else {
    const int k = steps_per_launch ? steps_per_launch : p->K;
    const int rows = rows_per_tile ? rows_per_tile : p->TY;
    if (steps_per_launch < 0 || rows_per_tile < 0 || planes_per_chunk < 0 ||
        planes_per_chunk > 1024 || (values_per_lane != 0 && values_per_lane != 1) ||
        !sps::fused3d_shape_ok((int)p->esize, k, rows)) {
      set_error("tune (3D): (steps_per_launch, rows_per_tile) in "
                "(1, 4|6|8|12|16), (2, 4) or, FP32 only, (2, 6|8|12); "
                "planes_per_chunk in 0..1024; values_per_lane 0 or 1");
      return SPS_E_INVALID;
    }
    p->K = k;
    p->TY = rows;
    if (planes_per_chunk) p->TZ = planes_per_chunk;
  }
  p->drop_graphs();
  return SPS_OK;
}

}  // extern "C"
