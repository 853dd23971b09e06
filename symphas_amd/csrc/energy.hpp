// energy.hpp - the Landau free-energy density the FP64 reductions sum.
//
//   f(psi) = -c1/2 psi^2 + c2/4 psi^4 + kappa/2 |grad psi|^2
//
// (landau_fe, sol/inc/modelspecialized.h:31-45, with the reference's fixed
// gradient coefficient 1/2 generalised to kappa/2).  The gradient is the
// forward difference (psi[i+1] - psi[i]) / h, or the central one
// (psi[i+1] - psi[i-1]) / (2h) that the reference generates for a
// first-order derivative at its default accuracy (sol/inc/stencil.h:45,
// 610-660).  The coefficients are pre-scaled on the host so a site costs
// three multiply-adds after the differences.
#pragma once

#include <cmath>

#include "common.hpp"

namespace sps {

struct EnergyParams {
  double a2;     // -c1/2
  double a4;     // c2/4
  double ag;     // kappa/2 * (1/h)^2 forward, kappa/2 * (1/(2h))^2 central
  int central;   // 0 forward, 1 central
  int mean;      // 0: scale the sum by h^D; 1: divide it by the site count
};

inline int make_energy_params(const sps_fd_energy* e, double h,
                              EnergyParams& out) {
  sps_fd_energy v{1.0, 1.0, 1.0, SPS_GRAD_FORWARD, 0};
  if (e) v = *e;
  if (!std::isfinite(v.c1) || !std::isfinite(v.c2) || !std::isfinite(v.kappa)) {
    set_error("energy coefficients must be finite");
    return SPS_E_INVALID;
  }
  if (v.gradient != SPS_GRAD_FORWARD && v.gradient != SPS_GRAD_CENTRAL) {
    set_error("energy gradient must be SPS_GRAD_FORWARD or SPS_GRAD_CENTRAL "
              "(got %d)", v.gradient);
    return SPS_E_INVALID;
  }
  if (v.mean != 0 && v.mean != 1) {
    set_error("energy mean must be 0 or 1 (got %d)", v.mean);
    return SPS_E_INVALID;
  }
  const double dh = v.gradient == SPS_GRAD_CENTRAL ? 0.5 / h : 1.0 / h;
  out.a2 = -0.5 * v.c1;
  out.a4 = 0.25 * v.c2;
  out.ag = 0.5 * v.kappa * dh * dh;
  out.central = v.gradient == SPS_GRAD_CENTRAL;
  out.mean = v.mean;
  return SPS_OK;
}

// Density at one site from psi and its unscaled gradient differences' squared
// sum g2 (sum over axes of (difference)^2).
__host__ __device__ __forceinline__ double landau_density(
    const EnergyParams& ep, double p, double g2) {
  const double p2 = p * p;
  return ep.a2 * p2 + ep.a4 * (p2 * p2) + ep.ag * g2;
}

// Final scale of the summed density: h^D for the integral, 1/sites for the
// mean.
inline double energy_scale(const EnergyParams& ep, int dim, double h,
                           double sites) {
  if (ep.mean) return 1.0 / sites;
  double hd = h * h;
  if (dim == 3) hd *= h;
  return hd;
}

}  // namespace sps
