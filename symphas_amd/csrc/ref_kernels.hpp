// ref_kernels.hpp - launchers of the reference-order kernels.
#pragma once

#include "common.hpp"

namespace sps {

// mode 0: out = in + dt*rhs; 1: out = rhs; 2: out = D[in] for one derivative.
template <typename T>
int launch_ref(const T* in, T* out, const RefParams<T>& p, int dim,
               hipStream_t stream);

// Mode 0 over a box of dims p.n whose boundary layer (SPS_BOUNDARY_DEPTH
// cells per side) supplies the neighbours: interior sites get psi + dt*rhs,
// layer sites are copied.
template <typename T>
int launch_ref_box(const T* in, T* out, const RefParams<T>& p, int dim,
                   hipStream_t stream);

}  // namespace sps
