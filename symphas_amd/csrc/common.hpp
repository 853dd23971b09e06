// common.hpp - shared definitions for the gfx950 finite-difference library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/sps_fd.h"

namespace sps {

// Last error text, per host thread (sps_fd_last_error).
void set_error(const char* fmt, ...);

#define SPS_HIP_TRY(expr)                                                    \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ::sps::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                  \
      return SPS_E_HIP;                                                      \
    }                                                                        \
  } while (0)

#define SPS_TRY(expr)          \
  do {                         \
    int rc_ = (expr);          \
    if (rc_ != SPS_OK) return rc_; \
  } while (0)

constexpr int kMaxDeg = SPS_MAX_DEGREE;
constexpr int kMaxTerms = SPS_MAX_TERMS;

// One term in device-friendly form (the reference-order kernels read these).
template <typename T>
struct DevTerm {
  int op, axis, component, identity;  // identity: P is exactly the operand
  int operand;                        // 0: the kernel's input, k: var(k)
  T coeff;
  T poly[kMaxDeg + 1];
};

// Everything the reference-order kernels need; passed by value.
template <typename T>
struct RefParams {
  int n[3];
  int L, G, B;
  int nterms;
  int mode;  // 0: out = in + dt*rhs; 1: out = rhs; 2: out = D[in] (apply)
  int apply_op, apply_axis, apply_component;
  T divh2, divh3, divh4, dt;
  const T* vars[SPS_MAX_PROV + 1];  // vars[k]: grid of var(k) (k >= 1)
  DevTerm<T> terms[kMaxTerms];
};

// Canonical fused form: dpsi/dt = A(psi) + lap(B(psi)) + kc*bilap(psi), with
// B free of a constant term (its laplacian is zero).  With S the sum of the
// 2*dim axial neighbours, w = 2*dim and D = S - w*psi the unscaled 5/7-point
// difference, a radius-two step is
//   mu   = B(psi) + kmu*D(psi),          kmu = kc*divh2
//   psi' = psi + Adt(psi) + kd*D(mu),    kd = dt*divh2, Adt = dt*A
// which the kernels evaluate with the centre weight and kmu folded into the
// polynomial of the first stage:
//   mu~  = ks*S + psi*p(psi)    mu~ = mu/kmu (ks = 1), or mu if kmu = 0 (ks = 0)
//   psi' = psi + q(psi) + k2*D(mu~)   q = Adt, k2 = kd*kmu (kd if kmu = 0)
// In "radius one" form (B = b1*psi, kc = 0) a step is
//   psi' = psi + q(psi) + k1*D(psi)   k1 = kd*b1.
// Two things stay unfolded on purpose: psi's own coefficient 1, and D as
// fma(-w, psi, S), which is exactly zero on a uniform field.  Folded into the
// rounded polynomial coefficients, either would bias the uniform mode's rate
// every step, an error that grows linearly with the step count.  The first
// stage's fold only perturbs mu~ by a rounding, which the exact D of the
// second stage does not turn into such a bias.
template <typename T>
struct FusedCoef {
  T q[kMaxDeg + 1];
  T p[kMaxDeg];
  T ks, k1, k2;
  int dega, degb;  // degrees of A and B, -1 when identically zero
  int radius;      // 1 or 2
  int poly;        // degree bound of the kernel instance: 3, or 6 (also when
                   // kmu = 0, which needs ks)
};

// Host-side lowering result.
struct Lowered {
  bool fusable = false;
  const char* why_not = "";
  double a[kMaxDeg + 1] = {0};
  double b[kMaxDeg + 1] = {0};
  double kc = 0;  // coefficient of bilap(psi)
};

}  // namespace sps
