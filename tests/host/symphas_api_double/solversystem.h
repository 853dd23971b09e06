// Test double of SolverSystemFD (sol/inc/solversystem.h:50-60): a padded
// host field with BOUNDARY_DEPTH ghost cells per side, x fastest, whose
// update() refreshes the ghosts periodically (boundarysystem.h:210-214).
// Written for this test, not taken from SymPhas.
#pragma once

#include <cstddef>
#include <vector>

#include "solver.h"

namespace symphas {
struct init_data_type {
  std::vector<double> interior;  // x fastest
};
struct interval_data_type {
  len_type n[3] = {1, 1, 1};
};
struct b_data_type {};
}  // namespace symphas

template <typename T, size_t D>
struct SolverSystemFD {
  len_type dims[D] = {};
  std::vector<T> storage;
  T* values = nullptr;
  std::vector<T> dframe;

  SolverSystemFD() = default;
  SolverSystemFD(symphas::init_data_type const& tdata,
                 symphas::interval_data_type const& vdata,
                 symphas::b_data_type const&, size_t = 0) {
    size_t len = 1;
    for (size_t a = 0; a < D; ++a) {
      dims[a] = vdata.n[a] + 2 * BOUNDARY_DEPTH;
      len *= (size_t)dims[a];
    }
    storage.assign(len, T(0));
    dframe.assign(len, T(0));
    values = storage.data();
    for (size_t i = 0; i < tdata.interior.size(); ++i) values[padded(i)] = tdata.interior[i];
    refresh();
  }
  SolverSystemFD(const SolverSystemFD& o)
      : storage(o.storage), dframe(o.dframe) {
    for (size_t a = 0; a < D; ++a) dims[a] = o.dims[a];
    values = storage.data();
  }

  size_t padded(size_t i) const {
    size_t idx = 0, stride = 1;
    for (size_t a = 0; a < D; ++a) {
      const size_t n = (size_t)dims[a] - 2 * BOUNDARY_DEPTH;
      idx += (i % n + BOUNDARY_DEPTH) * stride;
      i /= n;
      stride *= (size_t)dims[a];
    }
    return idx;
  }

  void refresh() {
    const size_t len = storage.size();
    for (size_t k = 0; k < len; ++k) {
      size_t r = k, src = 0, stride = 1;
      for (size_t a = 0; a < D; ++a) {
        const long n = dims[a] - 2 * BOUNDARY_DEPTH;
        long c = (long)(r % (size_t)dims[a]) - BOUNDARY_DEPTH;
        r /= (size_t)dims[a];
        c = ((c % n) + n) % n;
        src += (size_t)(c + BOUNDARY_DEPTH) * stride;
        stride *= (size_t)dims[a];
      }
      values[k] = values[src];
    }
  }

  void update(iter_type, double) { refresh(); }
};
