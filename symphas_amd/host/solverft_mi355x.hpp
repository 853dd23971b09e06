// solverft_mi355x.hpp - C++ drop-in host side of the MI355X SolverFT step.
//
// A header-only mirror of the SymPhas solver-plugin interface for this path
// (sol/inc/solver.h:162-274, examples/solvers/solverft.h:41-184), driving the
// HIP kernels through the C ABI of include/sps_fd.h:
//
//   Stencil2d2h<L,G,B>, Stencil3d2h<L,G,B>      template order and defaults of
//                                               sol/inc/stencilincludes.h:315-409,
//                                               :894-917 (<9,6,13>, <15,10,25>)
//   op(1), lap, bilap, gradlap<ax>, power, +,-,*  the EVOLUTION vocabulary
//                                               (sol/inc/modelmacros.h:1153-1180,
//                                               :1263), lowered as apply_operators
//                                               does (sym/inc/expressionapply.h:
//                                               1368-1394,1466-1488)
//   SolverFT_MI355X<stencil_t>                  make_solver / form_expr_one /
//                                               equation / step
//   SolverSystemFD_MI355X<D>                    the system: values, dframe, update
//   Model<D, Solver>, run_model, find_solution  modules-models.h:436-492,
//                                               modules-none.h:43-53
//
// Errors throw sps::symphas::solver_error carrying the C ABI's status and
// message (the reference exits; see sol/inc/stencil.h:671-679).
#pragma once

#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/sps_fd.h"

namespace sps::symphas {

struct solver_error : std::runtime_error {
  int status;
  solver_error(int st, const std::string& what)
      : std::runtime_error(what), status(st) {}
};

inline void check(int rc, const char* where) {
  if (rc != SPS_OK)
    throw solver_error(rc, std::string(where) + ": " + sps_fd_last_error());
}

enum class Axis { X = 0, Y = 1, Z = 2 };

// ------------------------------------------------------------- stencils

template <size_t L = 9, size_t G = 6, size_t B = 13>
struct Stencil2d2h {
  static constexpr int dimension = 2;
  static constexpr int lap_points = (int)L, gradlap_points = (int)G,
                       bilap_points = (int)B;
};

template <size_t L = 15, size_t G = 10, size_t B = 25>
struct Stencil3d2h {
  static constexpr int dimension = 3;
  static constexpr int lap_points = (int)L, gradlap_points = (int)G,
                       bilap_points = (int)B;
};

// ---------------------------------------------------------- expressions
//
// A small runtime expression tree: the order-parameter field psi, constants,
// sums, products, powers and derivatives.  lower() brings it to the
// canonical sum of  coeff * D[ P(psi) ]  terms of sps_fd_term.

struct Node;
using NodeP = std::shared_ptr<const Node>;

struct Node {
  enum Kind { Const, Coeff, Field, Add, Mul, Neg, Deriv } kind;
  double value = 0;  // Const
  int index = 0;     // Coeff: c(index)
  int order = 0;     // Deriv: 2 lap, 3 gradlap, 4 bilap
  int axis = 0;      // Deriv order 3
  NodeP a, b;
};

struct Expr {
  NodeP n;
};

inline Expr make_const(double v) {
  auto p = std::make_shared<Node>();
  p->kind = Node::Const;
  p->value = v;
  return {p};
}

// op(1): the single order parameter of the model
inline Expr op(int index) {
  if (index != 1)
    throw solver_error(SPS_E_UNSUPPORTED,
                       "op(" + std::to_string(index) +
                           "): the MI355X step advances single-field models");
  auto p = std::make_shared<Node>();
  p->kind = Node::Field;
  return {p};
}

inline Expr bin(Node::Kind k, Expr x, Expr y) {
  auto p = std::make_shared<Node>();
  p->kind = k;
  p->a = x.n;
  p->b = y.n;
  return {p};
}
inline Expr operator+(Expr x, Expr y) { return bin(Node::Add, x, y); }
inline Expr operator*(Expr x, Expr y) { return bin(Node::Mul, x, y); }
inline Expr operator-(Expr x) {
  auto p = std::make_shared<Node>();
  p->kind = Node::Neg;
  p->a = x.n;
  return {p};
}
inline Expr operator-(Expr x, Expr y) { return x + (-y); }
inline Expr operator+(Expr x, double c) { return x + make_const(c); }
inline Expr operator+(double c, Expr x) { return make_const(c) + x; }
inline Expr operator-(Expr x, double c) { return x - make_const(c); }
inline Expr operator-(double c, Expr x) { return make_const(c) - x; }
inline Expr operator*(Expr x, double c) { return x * make_const(c); }
inline Expr operator*(double c, Expr x) { return make_const(c) * x; }

inline Expr deriv(int order, Expr x, int axis = 0) {
  auto p = std::make_shared<Node>();
  p->kind = Node::Deriv;
  p->order = order;
  p->axis = axis;
  p->a = x.n;
  return {p};
}
inline Expr lap(Expr x) { return deriv(2, x); }
inline Expr bilap(Expr x) { return deriv(4, x); }
template <Axis ax>
inline Expr gradlap(Expr x) { return deriv(3, x, (int)ax); }
inline Expr power(Expr x, int n) {
  if (n < 0) throw solver_error(SPS_E_UNSUPPORTED, "negative power");
  Expr r = make_const(1.0);
  for (int i = 0; i < n; ++i) r = (i == 0) ? x : r * x;
  return r;
}
// c(i): the i-th model coefficient (1-based), resolved against the model's
// coefficient list when the expression is lowered; a coefficient the list
// does not provide is 1, as in the reference (lib/inc/definitions.h:628-630).
inline Expr c(int index) {
  auto p = std::make_shared<Node>();
  p->kind = Node::Coeff;
  p->index = index;
  return {p};
}

// A polynomial in psi, coefficients of psi^0 .. psi^6
struct Poly {
  std::array<double, SPS_MAX_DEGREE + 1> k{};
};

namespace detail {

inline Poly poly_add(const Poly& a, const Poly& b, double s = 1.0) {
  Poly r = a;
  for (int i = 0; i <= SPS_MAX_DEGREE; ++i) r.k[i] += s * b.k[i];
  return r;
}

inline Poly poly_mul(const Poly& a, const Poly& b) {
  Poly r;
  for (int i = 0; i <= SPS_MAX_DEGREE; ++i)
    for (int j = 0; j <= SPS_MAX_DEGREE; ++j)
      if (a.k[i] != 0.0 && b.k[j] != 0.0) {
        if (i + j > SPS_MAX_DEGREE)
          throw solver_error(SPS_E_UNSUPPORTED,
                             "power of psi above 6 (pow<N> semantics, "
                             "lib/inc/spslibmath.h:255-269)");
        r.k[i + j] += a.k[i] * b.k[j];
      }
  return r;
}

inline bool is_zero(const Poly& p) {
  for (double v : p.k)
    if (v != 0.0) return false;
  return true;
}

inline bool is_const(const Poly& p) {
  for (int i = 1; i <= SPS_MAX_DEGREE; ++i)
    if (p.k[i] != 0.0) return false;
  return true;
}

// Lowered value of a sub-expression: a polynomial part plus derivative terms.
struct Value {
  Poly poly;
  std::map<std::pair<int, int>, Poly> derivs;
};

inline Value scale(Value v, double s) {
  for (double& x : v.poly.k) x *= s;
  for (auto& kv : v.derivs)
    for (double& x : kv.second.k) x *= s;
  return v;
}

inline Value lower_node(const NodeP& n, const std::vector<double>& coeffs) {
  Value v;
  switch (n->kind) {
    case Node::Const:
      v.poly.k[0] = n->value;
      return v;
    case Node::Coeff:
      v.poly.k[0] = (n->index >= 1 && (size_t)n->index <= coeffs.size())
                        ? coeffs[n->index - 1]
                        : 1.0;
      return v;
    case Node::Field:
      v.poly.k[1] = 1.0;
      return v;
    case Node::Neg:
      return scale(lower_node(n->a, coeffs), -1.0);
    case Node::Add: {
      Value a = lower_node(n->a, coeffs), b = lower_node(n->b, coeffs);
      a.poly = poly_add(a.poly, b.poly);
      for (auto& kv : b.derivs)
        a.derivs[kv.first] = poly_add(a.derivs[kv.first], kv.second);
      return a;
    }
    case Node::Mul: {
      Value a = lower_node(n->a, coeffs), b = lower_node(n->b, coeffs);
      // a derivative may only be multiplied by a constant
      if (!a.derivs.empty() && !b.derivs.empty())
        throw solver_error(SPS_E_UNSUPPORTED, "product of two derivatives");
      if (!b.derivs.empty()) std::swap(a, b);
      if (!a.derivs.empty()) {
        if (!is_const(b.poly))
          throw solver_error(SPS_E_UNSUPPORTED,
                             "derivative multiplied by a non-constant");
        Value r = scale(a, b.poly.k[0]);
        return r;
      }
      v.poly = poly_mul(a.poly, b.poly);
      return v;
    }
    case Node::Deriv: {
      Value in = lower_node(n->a, coeffs);
      // derivatives distribute over sums (expressionapply.h:1476-1488)
      Poly operand = in.poly;
      operand.k[0] = 0.0;  // the derivative of a constant is zero
      if (!is_zero(operand)) v.derivs[{n->order, n->axis}] = operand;
      for (auto& kv : in.derivs) {
        // derivative of a derivative adds orders (expressionapply.h:
        // 1368-1394): lap(lap) = bilap; other combinations are not on the
        // Stencil{2,3}d2h path
        const int order = n->order + kv.first.first;
        if (!(n->order == 2 && kv.first.first == 2))
          throw solver_error(SPS_E_UNSUPPORTED,
                             "derivative of order " + std::to_string(order) +
                                 " other than lap(lap)");
        v.derivs[{4, 0}] = poly_add(v.derivs[{4, 0}], kv.second);
      }
      return v;
    }
  }
  return v;
}

}  // namespace detail

// Lower an EVOLUTION right-hand side into the descriptor's term list, in the
// order: bilaplacian, laplacian, gradlaplacian (by axis), pointwise.
inline int lower_terms(const Expr& rhs, sps_fd_term* terms, int max_terms,
                       const std::vector<double>& coeffs = {}) {
  detail::Value v = detail::lower_node(rhs.n, coeffs);
  int nt = 0;
  auto push = [&](int op, int axis, const Poly& p) {
    if (detail::is_zero(p)) return;
    if (nt >= max_terms)
      throw solver_error(SPS_E_UNSUPPORTED, "more than 8 terms");
    sps_fd_term& t = terms[nt++];
    std::memset(&t, 0, sizeof(t));
    t.op = op;
    t.axis = axis;
    t.coeff = 1.0;
    // single-monomial operands keep their coefficient outside, as
    // OpDerivative's value does (expressionderivatives.h:1545-1593)
    int nz = 0, last = 0;
    for (int k = 0; k <= SPS_MAX_DEGREE; ++k)
      if (p.k[k] != 0.0) ++nz, last = k;
    if (op != SPS_OP_ID && nz == 1) {
      t.coeff = p.k[last];
      t.poly[last] = 1.0;
    } else {
      for (int k = 0; k <= SPS_MAX_DEGREE; ++k) t.poly[k] = p.k[k];
    }
  };
  for (auto& kv : v.derivs) {
    if (kv.first.first != 4) continue;
    push(SPS_OP_BILAP, 0, kv.second);
  }
  for (auto& kv : v.derivs) {
    if (kv.first.first != 2) continue;
    // a laplacian of a sum is a sum of laplacians: one term per monomial
    for (int k = 1; k <= SPS_MAX_DEGREE; ++k)
      if (kv.second.k[k] != 0.0) {
        Poly m;
        m.k[k] = kv.second.k[k];
        push(SPS_OP_LAP, 0, m);
      }
  }
  for (auto& kv : v.derivs) {
    if (kv.first.first != 3) continue;
    push(SPS_OP_GRADLAP, kv.first.second, kv.second);
  }
  push(SPS_OP_ID, 0, v.poly);
  return nt;
}

// ------------------------------------------------------------ parameters

// Periodic problem on an interior grid of `dims` cells with spacing h (the
// configuration the reference builds with BoundaryType::PERIODIC ||
// n_h / h_dh intervals, README.md:91-97).
template <size_t D>
struct problem_parameters {
  std::array<int, D> dims{};
  double h = 1.0;
  std::vector<double> initial;  // interior values, x fastest; empty = zeros
};

// ------------------------------------------------------------ system

// The system of one field: a host mirror of the padded values (refreshed on
// demand) and the device plan that owns the field.
template <size_t D>
struct SolverSystemFD_MI355X {
  std::array<int, D> dims{};      // interior counts
  std::array<int, D> padded{};    // dims + 2*BOUNDARY_DEPTH
  double h = 1.0;
  std::vector<double> values;     // padded host mirror (FP64)
  std::vector<double> dframe;     // padded host mirror of the RHS
  sps_fd_plan* plan = nullptr;
  bool host_stale = false;

  explicit SolverSystemFD_MI355X(const problem_parameters<D>& p)
      : dims(p.dims), h(p.h) {
    size_t len = 1, inner = 1;
    for (size_t a = 0; a < D; ++a) {
      padded[a] = dims[a] + 2 * SPS_BOUNDARY_DEPTH;
      len *= (size_t)padded[a];
      inner *= (size_t)dims[a];
    }
    values.assign(len, 0.0);  // Block zero-fills (datatypes/inc/grid.h:1476)
    dframe.assign(len, 0.0);
    if (!p.initial.empty()) {
      if (p.initial.size() != inner)
        throw solver_error(SPS_E_INVALID, "initial field size mismatch");
      // the interior only (sol/inc/boundarysystem.h:505-507)
      std::vector<double> pad(len, 0.0);
      scatter_interior(p.initial.data(), pad.data());
      values.swap(pad);
    }
  }
  SolverSystemFD_MI355X(const SolverSystemFD_MI355X&) = delete;
  SolverSystemFD_MI355X& operator=(const SolverSystemFD_MI355X&) = delete;
  ~SolverSystemFD_MI355X() {
    if (plan) sps_fd_destroy(plan);
  }

  void scatter_interior(const double* in, double* pad) const {
    const int B = SPS_BOUNDARY_DEPTH;
    const int nz = D == 3 ? dims[2] : 1;
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < dims[1]; ++y)
        for (int x = 0; x < dims[0]; ++x) {
          size_t src = x + (size_t)dims[0] * (y + (size_t)dims[1] * z);
          size_t dst = (x + B) + (size_t)padded[0] *
                                     ((y + B) + (size_t)padded[1] *
                                                    (D == 3 ? z + B : 0));
          pad[dst] = in[src];
        }
  }

  // periodic ghost refresh of the host mirror: on the device the kernels
  // wrap indices instead (datatypes/inc/boundaryupdate.h:154-356,650-708)
  void update(long /*index*/, double /*time*/) { host_stale = true; }

  // padded values with fresh ghosts, as as_grid() after update()
  const std::vector<double>& as_grid() {
    if (host_stale && plan) {
      if (plan_dtype == SPS_F64) {
        check(sps_fd_download(plan, values.data(), SPS_LAYOUT_PADDED),
              "download");
      } else {
        std::vector<float> tmp(values.size());
        check(sps_fd_download(plan, tmp.data(), SPS_LAYOUT_PADDED),
              "download");
        for (size_t i = 0; i < tmp.size(); ++i) values[i] = tmp[i];
      }
      host_stale = false;
    }
    return values;
  }

  int plan_dtype = SPS_F64;
};

// ------------------------------------------------------------ solver

// SolverFT_MI355X<stencil_t, N>: the SolverFT step on the MI355X.  `dtype`
// selects FP64 (reference arithmetic) or FP32 storage and arithmetic;
// `path` the kernel family (include/sps_fd.h: sps_path).
template <typename stencil_t, size_t N = 0>
struct SolverFT_MI355X : stencil_t {
  static constexpr size_t D = stencil_t::dimension;
  double dt = 0;
  std::array<int, D> dims{};  // padded counts, as the reference's solver
  double h = 1.0;
  int dtype = SPS_F64;
  int path = SPS_PATH_AUTO;
  std::vector<double> coeffs;  // c(1), c(2), ... of the model

  // SolverFT::make_solver (solverft.h:121-161): dims = count + 2*depth,
  // h from the first system.
  static SolverFT_MI355X make_solver(const problem_parameters<D>& p) {
    SolverFT_MI355X s;
    for (size_t a = 0; a < D; ++a) s.dims[a] = p.dims[a] + 2 * SPS_BOUNDARY_DEPTH;
    s.h = p.h;
    return s;
  }

  // form_expr_one (solverft.h:112-119): canonicalise the RHS, then build the
  // device plan of the system and upload its values.
  template <size_t En = 0, typename S>
  std::pair<S*, Expr> form_expr_one(S& sys, const Expr& rhs) const {
    sps_fd_desc d;
    std::memset(&d, 0, sizeof(d));
    d.dim = (int)D;
    d.dtype = dtype;
    for (size_t a = 0; a < 3; ++a) d.n[a] = a < D ? sys.dims[a] : 1;
    d.lap_points = stencil_t::lap_points;
    d.gradlap_points = stencil_t::gradlap_points;
    d.bilap_points = stencil_t::bilap_points;
    d.path = path;
    d.h = h;
    d.dt = dt;
    d.nterms = lower_terms(rhs, d.terms, SPS_MAX_TERMS, coeffs);
    if (sys.plan) sps_fd_destroy(sys.plan);
    sys.plan = nullptr;
    check(sps_fd_create(&d, &sys.plan), "sps_fd_create");
    sys.plan_dtype = dtype;
    if (dtype == SPS_F64) {
      check(sps_fd_upload(sys.plan, sys.values.data(), SPS_LAYOUT_PADDED),
            "upload");
    } else {
      std::vector<float> tmp(sys.values.begin(), sys.values.end());
      check(sps_fd_upload(sys.plan, tmp.data(), SPS_LAYOUT_PADDED), "upload");
    }
    sys.host_stale = true;
    return {&sys, rhs};
  }

  // SolverFT::equation (solverft.h:90-103): the RHS is evaluated inside the
  // fused step; the host dframe mirror is filled only on request.
  template <typename S>
  void equation(std::pair<S*, Expr>& r) const {
    (void)r;
  }
  template <typename S>
  void evaluate_dframe(S& sys) const {
    if (sys.plan_dtype == SPS_F64) {
      check(sps_fd_equation(sys.plan, sys.dframe.data(), SPS_LAYOUT_PADDED),
            "equation");
    } else {
      std::vector<float> tmp(sys.dframe.size());
      check(sps_fd_equation(sys.plan, tmp.data(), SPS_LAYOUT_PADDED),
            "equation");
      for (size_t i = 0; i < tmp.size(); ++i) sys.dframe[i] = tmp[i];
    }
  }

  // SolverFT::step (solverft.h:46-50): psi += dt * dframe over the interior,
  // with the periodic refresh folded into the next launch.
  template <typename S>
  void step(S& sys) const {
    check(sps_fd_step(sys.plan, 1), "step");
    sys.host_stale = true;
  }

  // several steps in one call (graph-replayed on the device)
  template <typename S>
  void step_n(S& sys, int n) const {
    check(sps_fd_step(sys.plan, n), "step");
    sys.host_stale = true;
  }
};

// ------------------------------------------------------------ model

// Model<D, Sp> with one field: construction builds the solver and the
// equation (model.h:100-117, modelspecialized.h:370-375); step/update/
// equation follow model_iteration (modules-models.h:436-467).
template <size_t D, typename Sp>
struct Model {
  Sp solver;
  SolverSystemFD_MI355X<D> system0;
  std::pair<SolverSystemFD_MI355X<D>*, Expr> eq;
  long index = 0;
  double time = 0;
  double eq_dt = -1;
  Expr rhs;

  Model(const problem_parameters<D>& p, Expr evolution, int dtype = SPS_F64,
        int path = SPS_PATH_AUTO, std::vector<double> coeffs = {})
      : solver(Sp::make_solver(p)), system0(p), rhs(evolution) {
    solver.dtype = dtype;
    solver.path = path;
    solver.coeffs = std::move(coeffs);
  }

  SolverSystemFD_MI355X<D>& system() { return system0; }
  long get_index() const { return index; }

  void equation() {
    if (eq.first) solver.equation(eq);
  }
  void step(double dt) {
    ensure_plan(dt);
    solver.step(system0);
    ++index;
  }
  void step_n(double dt, int n) {
    ensure_plan(dt);
    solver.step_n(system0, n);
    index += n;
  }
  void update(double t) {
    time = t;
    system0.update(index, t);
  }

 private:
  // the descriptor carries dt, so the plan is (re)built when dt changes
  void ensure_plan(double dt) {
    if (dt != eq_dt || !system0.plan) {
      if (system0.plan && system0.host_stale) system0.as_grid();
      solver.dt = dt;
      eq = solver.form_expr_one(system0, rhs);
      eq_dt = dt;
    }
  }
};

// run_model (modules-models.h:481-492) with a fixed time step; ghosts are
// fresh before the first evaluation, as in the IO driver (modules-io.h:
// 450-452).  Consecutive steps are handed to the device in one call.
template <typename M>
bool run_model(M& model, long n, double dt, double starttime = 0) {
  model.update(starttime);
  model.step_n(dt, (int)n);
  model.update(starttime + dt * n);
  return true;
}

template <typename M>
void find_solution(M& model, double dt, long n) {
  run_model(model, n, dt, 0.0);
}

}  // namespace sps::symphas
