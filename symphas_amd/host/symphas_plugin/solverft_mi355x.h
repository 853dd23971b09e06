// solverft_mi355x.h - the SymPhas solver plugin SolverFT_MI355X.
//
// Drop this header next to the reference's solvers (examples/solvers/, listed
// in solverinclude.h) and a model selects the MI355X step with
//   model_MB_t<2, SolverFT_MI355X<Stencil2d2h<5, 6, 13>>>
// while its MODEL/EVOLUTION definition stays as written.  Link the executable
// with -L<repo>/symphas_amd/lib -lsps_fd.
//
// What the plugin provides, against the plugin interface of
// sol/inc/solver.h:162-274:
//   START_NEW_SOLVER_WITH_STENCIL(SolverFT_MI355X) ... END_SOLVER
//                        the solver struct (solver.h:385-451), deriving the
//                        stencil's dims and divh and Solver's dt as SolverFT does
//   form_expr_one        canonicalises the right-hand side as SolverFT does
//                        (examples/solvers/solverft.h:112-119) and lowers its
//                        type, OpAdd / OpTerms / OpDerivative<derivative<ax,O>>
//                        (sym/inc/expressionapply.h:1368-1488), into the term
//                        list of sps_fd_desc; a form outside that vocabulary
//                        fails to compile with a static_assert
//   step                 one sps_fd_step of the system's device plan, created
//                        on the first step once Model::step has set dt
//                        (sol/inc/model.h:932-936)
//   equation             nothing: the right-hand side is evaluated in the step
//   make_solver          dims = count + 2*BOUNDARY_DEPTH and h from system 0,
//                        as SolverFT::make_solver (solverft.h:121-161)
//   SolverSystemFD_MI355X
//                        SolverFT's SolverSystemFD (sol/inc/solversystem.h:
//                        50-60) with a device plan; update() refreshes the host
//                        values (ghosts included) from the device so every
//                        reader of values, persist or save sees the field the
//                        reference would hold
//   ASSOCIATE_SELECTABLE_SOLVER_SYSTEM_TYPE / SYMPHAS_SOLVER_ALL_SUPPORTED
//                        the registrations of solver.h:289-304, :368-373
//
// Scope: one real (FP64) field per equation; the right-hand side may hold
// pointwise monomials of that field and derivatives of order 2 (laplacian),
// 3 (gradlaplacian) and 4 (bilaplacian) of polynomials in it, with constant
// coefficients.  A monomial of another field is a coupled model, which runs
// through sps_fdm_* (a runtime error here names it).  Provisional variables
// are not lowered by this plugin.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "expressions.h"
#include "solver.h"
#include "solversystem.h"
#include "sps_fd.h"  // <repo>/include

namespace sps_bind {

// Errors throw, carrying the C ABI's message (the reference prints and
// exits, lib/inc/definitions.h:42-50).
struct plugin_error : std::runtime_error {
  int status;
  plugin_error(int st, const std::string& what)
      : std::runtime_error(what), status(st) {}
};

inline void check(int rc, const char* where) {
  if (rc != SPS_OK)
    throw plugin_error(rc, std::string(where) + ": " + sps_fd_last_error());
}

// Host values are refreshed from the device at every update() whose index is
// a multiple of this interval.  1 keeps the host field current after every
// step, as the reference's is; a run whose host readers (saves, data
// processes) act only every k steps may set k to skip the other copies.
inline int host_sync_interval = 1;

// ------------------------------------------------------ stencil point counts

template <typename S>
struct stencil_points {
  static_assert(sizeof(S) == 0,
                "SolverFT_MI355X takes Stencil2d2h<L, G, B> or Stencil3d2h<L, G, B>");
};
template <size_t L, size_t G, size_t B>
struct stencil_points<Stencil2d2h<L, G, B>> {
  static constexpr int dim = 2, lap = (int)L, gradlap = (int)G, bilap = (int)B;
};
template <size_t L, size_t G, size_t B>
struct stencil_points<Stencil3d2h<L, G, B>> {
  static constexpr int dim = 3, lap = (int)L, gradlap = (int)G, bilap = (int)B;
};

// ------------------------------------------------------------ field indices

// Variable<Z, G> is the Z-th field of the model (op(Z + 1) in EVOLUTION).
template <typename G>
struct field_index {
  static constexpr bool is_field = false;
};
template <size_t Z, typename G>
struct field_index<Variable<Z, G>> {
  static constexpr bool is_field = true;
  static constexpr int value = (int)Z;
};

// Exponent key of a Term: non-negative integer powers only
// (expr::_Xk_t, sym/inc/expressionprototypes.h:60-117).
template <expr::exp_key_t X>
constexpr int integer_power() {
  using K = expr::_Xk_t<X>;
  static_assert(K::D == 1 && !K::sign,
                "SolverFT_MI355X lowers non-negative integer powers of a field only");
  return (int)K::N;
}

// ----------------------------------------------------------------- the IR

constexpr int kMaxFields = 4;

struct Monomial {
  double coeff = 0;
  int pow[kMaxFields] = {0, 0, 0, 0};
  int degree() const { return pow[0] + pow[1] + pow[2] + pow[3]; }
};

struct Term {
  int op = SPS_OP_ID;
  int axis = 0;
  double coeff = 1;
  std::vector<Monomial> operand;  // coeff * D[sum of monomials]
};

template <typename V>
double value_of(V const& v) {
  return static_cast<double>(expr::eval(v));
}

// Constants that can stand alone in a sum or as a derivative's operand.
template <typename E>
struct is_constant : std::false_type {};
template <>
struct is_constant<OpIdentity> : std::true_type {};
template <>
struct is_constant<OpNegIdentity> : std::true_type {};
template <>
struct is_constant<OpVoid> : std::true_type {};
template <typename T>
struct is_constant<OpLiteral<T>> : std::true_type {};
template <size_t N, size_t D>
struct is_constant<OpFractionLiteral<N, D>> : std::true_type {};
template <size_t N, size_t D>
struct is_constant<OpNegFractionLiteral<N, D>> : std::true_type {};

template <typename E>
struct unsupported : std::false_type {};

// Operand monomials: sum of products of fields with a constant coefficient.
template <typename E>
void monomials(E const& e, double scale, std::vector<Monomial>& out) {
  if constexpr (is_constant<E>::value) {
    Monomial m;
    m.coeff = scale * value_of(e);
    if (m.coeff != 0.0) out.push_back(m);
  } else {
    static_assert(unsupported<E>::value,
                  "SolverFT_MI355X: a derivative's operand or a pointwise term must be a "
                  "sum of constant-coefficient products of fields (OpAdd of OpTerms)");
  }
}

template <typename V, typename... Gs, expr::exp_key_t... Xs>
void monomials(OpTerms<V, ::Term<Gs, Xs>...> const& e, double scale,
               std::vector<Monomial>& out) {
  static_assert((field_index<Gs>::is_field && ...),
                "SolverFT_MI355X: every factor of a product must be a field (op(n))");
  Monomial m;
  m.coeff = scale * value_of(expr::coeff(e));
  ((m.pow[field_index<Gs>::value] += integer_power<Xs>()), ...);
  out.push_back(m);
}

template <typename... Es, size_t... Is>
void monomials_add(OpAdd<Es...> const& e, double scale,
                   std::vector<Monomial>& out, std::index_sequence<Is...>) {
  (monomials(expr::get<Is>(e), scale, out), ...);
}

template <typename... Es>
void monomials(OpAdd<Es...> const& e, double scale, std::vector<Monomial>& out) {
  monomials_add(e, scale, out, std::make_index_sequence<sizeof...(Es)>{});
}

// Right-hand side terms.
template <size_t O>
constexpr int op_of_order() {
  static_assert(O == 2 || O == 3 || O == 4,
                "SolverFT_MI355X lowers laplacians (order 2), gradlaplacians (3) and "
                "bilaplacians (4)");
  return O == 2 ? SPS_OP_LAP : (O == 3 ? SPS_OP_GRADLAP : SPS_OP_BILAP);
}

template <typename E>
void lower(E const& e, double scale, std::vector<Term>& out) {
  // a pointwise constant or an unsupported form
  Term t;
  t.op = SPS_OP_ID;
  monomials(e, scale, t.operand);
  if (!t.operand.empty()) out.push_back(t);
}

template <typename V, typename... Gs, expr::exp_key_t... Xs>
void lower(OpTerms<V, ::Term<Gs, Xs>...> const& e, double scale,
           std::vector<Term>& out) {
  Term t;
  t.op = SPS_OP_ID;
  monomials(e, scale, t.operand);
  out.push_back(t);
}

template <typename... Es, size_t... Is>
void lower_add(OpAdd<Es...> const& e, double scale, std::vector<Term>& out,
               std::index_sequence<Is...>) {
  (lower(expr::get<Is>(e), scale, out), ...);
}

template <typename... Es>
void lower(OpAdd<Es...> const& e, double scale, std::vector<Term>& out) {
  lower_add(e, scale, out, std::make_index_sequence<sizeof...(Es)>{});
}

// Derivative of a non-trivial operand: the temporary grid of
// sym/inc/expressionderivatives.h:1545-1593.
template <typename Dd, typename V, typename E, typename Sp>
void lower(OpDerivative<Dd, V, E, Sp> const& d, double scale,
           std::vector<Term>& out) {
  using D = OpDerivative<Dd, V, E, Sp>;
  Term t;
  t.op = op_of_order<D::order>();
  t.axis = (int)D::axis;
  t.coeff = scale * value_of(d.value);
  monomials(d.e, 1.0, t.operand);
  if (!t.operand.empty()) out.push_back(t);
}

// Derivative applied to a field directly (expressionderivatives.h:1658-1682).
template <typename Dd, typename V, typename G, typename Sp>
void lower(OpDerivative<Dd, V, OpTerm<OpIdentity, G>, Sp> const& d,
           double scale, std::vector<Term>& out) {
  using D = OpDerivative<Dd, V, OpTerm<OpIdentity, G>, Sp>;
  static_assert(field_index<G>::is_field,
                "SolverFT_MI355X: a derivative must act on a field (op(n))");
  Term t;
  t.op = op_of_order<D::order>();
  t.axis = (int)D::axis;
  t.coeff = scale * value_of(d.value);
  Monomial m;
  m.coeff = 1.0;
  m.pow[field_index<G>::value] = 1;
  t.operand.push_back(m);
  out.push_back(t);
}

// The single-field descriptor of field `own` from the lowered terms.
inline void fill_terms(sps_fd_desc& d, const std::vector<Term>& terms,
                       int own) {
  if ((int)terms.size() > SPS_MAX_TERMS)
    throw plugin_error(SPS_E_UNSUPPORTED,
                       "SolverFT_MI355X: more than SPS_MAX_TERMS terms in the "
                       "right-hand side");
  d.nterms = 0;
  for (const Term& t : terms) {
    sps_fd_term& o = d.terms[d.nterms++];
    o = sps_fd_term{};
    o.op = t.op;
    o.axis = t.op == SPS_OP_GRADLAP ? t.axis : 0;
    o.component = 0;
    o.operand = 0;
    o.coeff = t.coeff;
    for (const Monomial& m : t.operand) {
      for (int f = 0; f < kMaxFields; ++f)
        if (f != own && m.pow[f] != 0)
          throw plugin_error(
              SPS_E_UNSUPPORTED,
              "SolverFT_MI355X: the right-hand side of field " +
                  std::to_string(own) + " reads field " + std::to_string(f) +
                  "; coupled models run through sps_fdm_*");
      const int k = m.pow[own];
      if (k > SPS_MAX_DEGREE)
        throw plugin_error(SPS_E_UNSUPPORTED,
                           "SolverFT_MI355X: a power above SPS_MAX_DEGREE");
      o.poly[k] += m.coeff;
    }
  }
}

// Device state of one system, shared by copies of the system object.
struct device_state {
  sps_fd_desc desc{};
  bool lowered = false;
  sps_fd_plan* plan = nullptr;
  double dt = 0;
  bool device_newer = false;  // the device holds steps the host has not seen
  bool host_changed = false;  // the host values were written since upload
  ~device_state() {
    if (plan) (void)sps_fd_destroy(plan);
  }
};

}  // namespace sps_bind

START_NEW_SOLVER_WITH_STENCIL(SolverFT_MI355X)

// One forward-Euler step of the system's field on the device (replaces
// SolverFT::step, examples/solvers/solverft.h:46-50, and the equation pass
// before it).
template <typename S>
void step(S& sys) const {
  auto& st = *sys.dev;
  if (!st.lowered)
    throw sps_bind::plugin_error(SPS_E_STATE,
                                 "SolverFT_MI355X: step before form_expr_one");
  if (!st.plan || st.dt != dt) {
    if (st.plan) {
      if (st.device_newer) sys.download_from_device();
      (void)sps_fd_destroy(st.plan);
      st.plan = nullptr;
    }
    st.desc.dt = dt;
    sps_bind::check(sps_fd_create(&st.desc, &st.plan), "sps_fd_create");
    st.dt = dt;
    st.host_changed = true;
  }
  if (st.host_changed) {
    sps_bind::check(sps_fd_upload(st.plan, sys.values, SPS_LAYOUT_PADDED),
                    "sps_fd_upload");
    st.host_changed = false;
  }
  sps_bind::check(sps_fd_step(st.plan, 1), "sps_fd_step");
  st.device_newer = true;
}

// The right-hand side is evaluated inside step.
template <typename S, typename E>
void equation(std::pair<S, E>&) const {}

template <typename S, typename E>
void equation(std::pair<S, E>*, len_type) const {}

template <size_t En, typename SS, typename S, typename E>
auto form_expr_one(SS&&, std::pair<S, E> const& e) const {
  auto [sys, equation] = e;
  auto eq = expr::transform::optimize(expr::apply_operators(equation));
  expr::prune::update(eq);
  static_assert(sps_bind::field_index<S>::is_field,
                "SolverFT_MI355X: an equation's left side must be a field");
  constexpr int own = sps_bind::field_index<S>::value;
  using P = sps_bind::stencil_points<stencil_t>;
  std::vector<sps_bind::Term> terms;
  sps_bind::lower(eq, 1.0, terms);
  auto& system = sys.get();
  sps_fd_desc& d = system.dev->desc;
  d = sps_fd_desc{};
  d.dim = P::dim;
  d.dtype = SPS_F64;
  for (int a = 0; a < P::dim; ++a)
    d.n[a] = (int32_t)(stencil_t::dims[a] - 2 * BOUNDARY_DEPTH);
  d.lap_points = P::lap;
  d.gradlap_points = P::gradlap;
  d.bilap_points = P::bilap;
  d.path = SPS_PATH_AUTO;
  d.h = 1.0 / stencil_t::divh;
  d.dt = dt;
  sps_bind::fill_terms(d, terms, own);
  system.dev->lowered = true;
  system.dev->host_changed = true;
  return std::make_pair(sys, eq);
}

// As SolverFT::make_solver (examples/solvers/solverft.h:121-161): the grid
// spacing of system 0's X axis, dims = count + 2 * BOUNDARY_DEPTH.
static auto make_solver(symphas::problem_parameters_type const& parameters) {
  if (parameters.length()) {
    double h = parameters.get_interval_data()[0].at(Axis::X).width();
    size_t dim = parameters.get_dimension();
    std::vector<len_type> dims(dim);
    for (size_t i = 0; i < dim; ++i)
      dims[i] = parameters.get_interval_data()[0]
                    .at(symphas::index_to_axis((iter_type)i))
                    .get_count() +
                2 * BOUNDARY_DEPTH;
    return this_type{dims.data(), h};
  }
  return this_type{grid::dim_list(nullptr, 3), grid::h_list(nullptr, 3)};
}
END_SOLVER

// SolverSystemFD with a device plan.  The device holds the field between
// steps; update() copies it back into values (ghosts filled periodically by
// the library), the work the reference's update does on the host
// (sol/inc/boundarysystem.h:210-214).
template <typename T, size_t D>
struct SolverSystemFD_MI355X : SolverSystemFD<T, D> {
  static_assert(std::is_same<T, double>::value,
                "SolverFT_MI355X steps real FP64 fields");
  using parent_type = SolverSystemFD<T, D>;
  using parent_type::values;

  std::shared_ptr<sps_bind::device_state> dev =
      std::make_shared<sps_bind::device_state>();

  SolverSystemFD_MI355X(symphas::init_data_type const& tdata,
                        symphas::interval_data_type const& vdata,
                        symphas::b_data_type const& bdata, size_t id = 0)
      : parent_type(tdata, vdata, bdata, id) {}
  SolverSystemFD_MI355X() : parent_type() {}

  void download_from_device() {
    sps_bind::check(sps_fd_download(dev->plan, values, SPS_LAYOUT_PADDED),
                    "sps_fd_download");
    dev->device_newer = false;
  }

  // Call after writing values on the host (a loaded checkpoint, a changed
  // initial condition): the next step uploads them.
  void mark_host_changed() { dev->host_changed = true; }

  void update(iter_type index, double time) {
    if (dev->plan && dev->device_newer) {
      if (index % sps_bind::host_sync_interval == 0) download_from_device();
      return;
    }
    parent_type::update(index, time);
  }
};

ASSOCIATE_SELECTABLE_SOLVER_SYSTEM_TYPE(SolverFT_MI355X, SolverSystemFD_MI355X)
SYMPHAS_SOLVER_ALL_SUPPORTED(SolverFT_MI355X)
