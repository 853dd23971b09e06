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
// the coefficients pre-scaled for the kernels:
//   mu   = B(psi) + kmu * D(psi)          D = unscaled 5/7-point difference
//   psi' = psi + Adt(psi) + kd * D(mu)
// kmu = kc*divh2, kd = dt*divh2, Adt = dt*A.  In "radius one" form (B linear,
// kc = 0) the update is psi' = psi + Adt(psi) + k1 * D(psi), k1 = kd*b1.
template <typename T>
struct FusedCoef {
  T adt[kMaxDeg + 1];
  T b[kMaxDeg + 1];
  T kmu, kd, k1;
  int dega, degb;  // -1 when the polynomial is identically zero
  int radius;      // 1 or 2
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
