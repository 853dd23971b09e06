// Test double of the SymPhas solver-plugin surface the SolverFT_MI355X
// header uses (sol/inc/solver.h:385-451, :289-304, :368-373; stencils of
// sol/inc/stencilincludes.h:95-135, 315-409; problem parameters).  Written
// for this test, not taken from SymPhas.
#pragma once

#include <cstddef>
#include <vector>

#include "expressions.h"

#define BOUNDARY_DEPTH 3

template <typename Sp, size_t N>
struct Solver {
  double dt = 0;
  template <Axis ax, size_t O>
  struct derivative {
    static const size_t order = O;
    static const Axis axis = ax;
  };
};

template <size_t D>
struct StencilDims {
  len_type dims[D] = {};
  double divh = 1;
  StencilDims(const len_type* d, double h) : divh(1.0 / h) {
    for (size_t i = 0; i < D; ++i) dims[i] = d ? d[i] : 0;
  }
  StencilDims() = default;
};

template <size_t... Ns>
struct Stencil2d2h;
template <size_t L, size_t G, size_t B>
struct Stencil2d2h<L, G, B> : StencilDims<2> {
  using StencilDims<2>::StencilDims;
};
template <size_t... Ns>
struct Stencil3d2h;
template <size_t L, size_t G, size_t B>
struct Stencil3d2h<L, G, B> : StencilDims<3> {
  using StencilDims<3>::StencilDims;
};

namespace grid {
struct dim_list {
  const len_type* d;
  dim_list(const len_type* p, size_t) : d(p) {}
  operator const len_type*() const { return d; }
};
struct h_list {
  double h = 1.0;
  h_list(const double*, size_t) {}
  operator double() const { return h; }
};
}  // namespace grid

namespace symphas {
struct interval {
  double w;
  len_type count;
  double width() const { return w; }
  len_type get_count() const { return count; }
};
struct interval_data {
  interval ax[3];
  interval const& at(Axis a) const { return ax[(int)a]; }
};
struct problem_parameters_type {
  size_t dim = 2;
  std::vector<interval_data> systems;
  size_t length() const { return systems.size(); }
  size_t get_dimension() const { return dim; }
  const interval_data* get_interval_data() const { return systems.data(); }
};
inline Axis index_to_axis(iter_type i) { return (Axis)i; }
}  // namespace symphas

#define START_NEW_SOLVER_WITH_STENCIL(NAME)                  \
  template <typename stencil_t, size_t N = 0>                \
  struct NAME : Solver<NAME<stencil_t, N>, N>, stencil_t {   \
    using this_type = NAME<stencil_t, N>;                    \
    using parent_type = Solver<this_type, N>;                \
    using stencil_t::stencil_t;                              \
    using parent_type::dt;

#define END_SOLVER \
  }                \
  ;

// The double records the association so a test can name the system type.
template <template <typename, size_t> class Solver_>
struct associated_system;
#define ASSOCIATE_SELECTABLE_SOLVER_SYSTEM_TYPE(SOLVER, SYSTEM)         \
  template <>                                                        \
  struct associated_system<SOLVER> {                                 \
    template <typename T, size_t D>                                  \
    using type = SYSTEM<T, D>;                                       \
  };
#define SYMPHAS_SOLVER_ALL_SUPPORTED(SOLVER)
