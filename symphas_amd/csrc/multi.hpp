// multi.hpp - coupled-field (several order parameters) definitions shared by
// the reference-order and fused kernels of sps_fdm_*.
#pragma once

#include "common.hpp"

namespace sps {

// One term coeff * D[prod_f psi_f^pow[f]] in device form.
template <typename T>
struct MDevTerm {
  int op, axis, component;
  int single;  // the field read directly when the monomial is one field, else -1
  int pow[SPS_MAX_FIELDS];
  T coeff;
};

// Reference-order parameters: all fields' term lists, passed by value.
template <typename T>
struct MRefParams {
  int n[3];
  int L, G, B;
  int nfields;
  int nterms[SPS_MAX_FIELDS];
  T divh2, divh3, divh4, dt;
  const T* in[SPS_MAX_FIELDS];
  T* out[SPS_MAX_FIELDS];
  MDevTerm<T> terms[SPS_MAX_FIELDS][SPS_MAX_MTERMS];
};

template <typename T>
int launch_ref_multi(const MRefParams<T>& p, int dim, hipStream_t stream);

// Fused two-field form, per field f (fields 0 and 1, 2D):
//   dpsi_f/dt = A_f(psi0, psi1) + lap(B_f(psi0, psi1)) + kc_f * bilap(psi_f)
// with A_f, B_f dense bivariate polynomials, a[f][i][j] multiplying
// psi0^i psi1^j (degree <= 3 in each field), evaluated as
//   radius 2:  mu_f = B_f + kmu_f * D(psi_f),  psi_f' = psi_f + Adt_f + kd * D(mu_f)
//   radius 1:  psi_f' = psi_f + Adt_f + k1_f * D(psi_f)   (B_f = b_f * psi_f)
// Each of the four polynomials carries a shape class chosen on the host from
// its nonzero coefficients, so the kernel evaluates only what is there:
//   kPolyZero  identically 0                                   (0 FMAs)
//   kPolyLin   c00 + c10 x + c01 y                             (2 FMAs)
//   kPolyXLin  total degree <= 3, degree <= 1 in x             (6 FMAs)
//   kPolyYLin  total degree <= 3, degree <= 1 in y             (6 FMAs)
//   kPolySep   f(x) + g(y), each of degree <= 3                (6 FMAs)
//   kPolyTri   total degree <= 3, triangular two-level Horner  (9 FMAs)
//   kPolyDense every c[i][j], i, j <= 3                        (15 FMAs)
// The class is a kernel argument, so the branch on it is uniform.
constexpr int kMDeg = 3;
enum {
  kPolyZero = 0,
  kPolyLin = 1,
  kPolyTri = 2,
  kPolyDense = 3,
  kPolyXLin = 4,
  kPolyYLin = 5,
  kPolySep = 6
};
template <typename T>
struct MFusedCoef {
  T adt[2][kMDeg + 1][kMDeg + 1];  // dt * A_f
  T b[2][kMDeg + 1][kMDeg + 1];
  T kmu[2], k1[2], kd;
  int radius;
  int adt_class[2], b_class[2];
  int lin;  // radius 2: 1 (field 1) or 2 (field 0) takes the radius-one form
};

struct MLowered {
  bool fusable = false;
  const char* why_not = "";
  double a[2][kMDeg + 1][kMDeg + 1] = {};
  double b[2][kMDeg + 1][kMDeg + 1] = {};
  double kc[2] = {0, 0};
  int radius = 2;
};

template <typename T>
int launch_fused_multi(const T* const* in, T* const* out, int X, int Y, int K,
                       int TY, int vec, const MFusedCoef<T>& c, hipStream_t s);

// 3D: radius one K in {1, 2}, radius two K = 1; TY rows per wave in {4, 8},
// TZ planes per streamed chunk
bool fused3d_multi_shape_ok(int radius, int K, int rows);
// Scratch (register spill) bytes per lane of the 2D coupled kernel instance a
// launch of K steps at vec values per lane would run; 0 when it does not spill.
template <typename T>
size_t fused_multi_scratch(int K, int vec, const MFusedCoef<T>& c);
template <typename T>
int launch_fused3d_multi(const T* const* in, T* const* out, int X, int Y,
                         int Z, int K, int TY, int TZ, const MFusedCoef<T>& c,
                         hipStream_t s);

}  // namespace sps
