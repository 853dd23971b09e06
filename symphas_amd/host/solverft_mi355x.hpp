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
//   var(k), var(k) <= expr                      provisional variables
//                                               (PROVISIONAL_DEF, MBB at
//                                               examples/models/modeldefinitions.h:
//                                               150-155)
//   SolverFT_MI355X<stencil_t>                  make_solver / form_expr_one /
//                                               equation / step
//   SolverSystemFD_MI355X<D>                    the system: values, dframe, update
//   Model<D, Solver>, run_model, find_solution  modules-models.h:436-492,
//                                               modules-none.h:43-53
//   op(1) .. op(4), MultiModel<D, Solver, F>    coupled fields (MC, GRAYSCOTT,
//                                               Turing: examples/models/
//                                               modeldefinitions.h:63-81,
//                                               123-137), through sps_fdm_*
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
#include <tuple>
#include <utility>
#include <vector>
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
// A small runtime expression tree: the order-parameter field psi, the
// provisional variables var(k), constants, sums, products, powers and
// derivatives.  lower() brings it to the canonical sum of
// coeff * D[ P(u) ] terms of sps_fd_term, u = psi or var(k).

struct Node;
using NodeP = std::shared_ptr<const Node>;

struct Node {
  enum Kind { Const, Coeff, Field, Var, Add, Mul, Neg, Deriv } kind;
  double value = 0;  // Const
  int index = 0;     // Coeff: c(index); Var: var(index)
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

// op(k): the k-th order parameter of the model (1-based).  A single-field
// Model takes op(1) only; a MultiModel takes op(1) .. op(SPS_MAX_FIELDS).
inline Expr op(int index) {
  if (index < 1 || index > SPS_MAX_FIELDS)
    throw solver_error(SPS_E_UNSUPPORTED,
                       "op(" + std::to_string(index) + "): at most " +
                           std::to_string(SPS_MAX_FIELDS) + " fields");
  auto p = std::make_shared<Node>();
  p->kind = Node::Field;
  p->index = index;
  return {p};
}

// var(k): the k-th provisional variable (1-based, k <= SPS_MAX_PROV)
inline Expr var(int index) {
  if (index < 1 || index > SPS_MAX_PROV)
    throw solver_error(SPS_E_UNSUPPORTED,
                       "var(" + std::to_string(index) + "): at most " +
                           std::to_string(SPS_MAX_PROV) +
                           " provisional variables");
  auto p = std::make_shared<Node>();
  p->kind = Node::Var;
  p->index = index;
  return {p};
}

// The definition of a provisional variable, written var(k) <= expr as in
// PROVISIONAL_DEF (examples/models/modeldefinitions.h:153-154).
struct ProvisionalDef {
  int index;
  Expr def;
};
inline ProvisionalDef operator<=(Expr v, Expr def) {
  if (v.n->kind != Node::Var)
    throw solver_error(SPS_E_INVALID,
                       "the left side of <= must be a provisional variable");
  return {v.n->index, def};
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

// A polynomial in one operand, coefficients of u^0 .. u^6
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
                             "power above 6 (pow<N> semantics, "
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

// Lowered value of a sub-expression: a constant, a polynomial per operand
// (0 = psi, k = var(k); no constant coefficient) and derivative terms keyed
// by (order, axis, operand).
struct Value {
  double c0 = 0;
  std::map<int, Poly> poly;
  std::map<std::tuple<int, int, int>, Poly> derivs;

  bool is_const() const {
    for (auto& kv : poly)
      if (!is_zero(kv.second)) return false;
    return derivs.empty();
  }
};

inline Value scale(Value v, double s) {
  v.c0 *= s;
  for (auto& kv : v.poly)
    for (double& x : kv.second.k) x *= s;
  for (auto& kv : v.derivs)
    for (double& x : kv.second.k) x *= s;
  return v;
}

inline Value lower_node(const NodeP& n, const std::vector<double>& coeffs) {
  Value v;
  switch (n->kind) {
    case Node::Const:
      v.c0 = n->value;
      return v;
    case Node::Coeff:
      v.c0 = (n->index >= 1 && (size_t)n->index <= coeffs.size())
                 ? coeffs[n->index - 1]
                 : 1.0;
      return v;
    case Node::Field:
      if (n->index != 1)
        throw solver_error(SPS_E_UNSUPPORTED,
                           "op(" + std::to_string(n->index) +
                               ") in a single-field model (MultiModel steps "
                               "several fields)");
      v.poly[0].k[1] = 1.0;
      return v;
    case Node::Var:
      v.poly[n->index].k[1] = 1.0;
      return v;
    case Node::Neg:
      return scale(lower_node(n->a, coeffs), -1.0);
    case Node::Add: {
      Value a = lower_node(n->a, coeffs), b = lower_node(n->b, coeffs);
      a.c0 += b.c0;
      for (auto& kv : b.poly) a.poly[kv.first] = poly_add(a.poly[kv.first], kv.second);
      for (auto& kv : b.derivs)
        a.derivs[kv.first] = poly_add(a.derivs[kv.first], kv.second);
      return a;
    }
    case Node::Mul: {
      Value a = lower_node(n->a, coeffs), b = lower_node(n->b, coeffs);
      if (b.is_const()) return scale(a, b.c0);
      if (a.is_const()) return scale(b, a.c0);
      // a derivative may only be multiplied by a constant
      if (!a.derivs.empty() || !b.derivs.empty())
        throw solver_error(SPS_E_UNSUPPORTED,
                           "derivative multiplied by a non-constant");
      // a product of polynomials in one operand; psi * var(k) has no term form
      int operand = -1;
      for (const Value* x : {&a, &b})
        for (auto& kv : x->poly) {
          if (is_zero(kv.second)) continue;
          if (operand >= 0 && operand != kv.first)
            throw solver_error(SPS_E_UNSUPPORTED,
                               "product of two different operands");
          operand = kv.first;
        }
      Poly pa = a.poly[operand], pb = b.poly[operand];
      pa.k[0] = a.c0;
      pb.k[0] = b.c0;
      Poly r = poly_mul(pa, pb);
      v.c0 = r.k[0];
      r.k[0] = 0.0;
      v.poly[operand] = r;
      return v;
    }
    case Node::Deriv: {
      Value in = lower_node(n->a, coeffs);
      // derivatives distribute over sums (expressionapply.h:1476-1488); the
      // derivative of a constant is zero
      for (auto& kv : in.poly)
        if (!is_zero(kv.second))
          v.derivs[{n->order, n->axis, kv.first}] =
              poly_add(v.derivs[{n->order, n->axis, kv.first}], kv.second);
      for (auto& kv : in.derivs) {
        // derivative of a derivative adds orders (expressionapply.h:
        // 1368-1394): lap(lap) = bilap; other combinations are not on the
        // Stencil{2,3}d2h path
        const int inner = std::get<0>(kv.first);
        if (!(n->order == 2 && inner == 2))
          throw solver_error(SPS_E_UNSUPPORTED,
                             "derivative of order " +
                                 std::to_string(n->order + inner) +
                                 " other than lap(lap)");
        const auto key = std::make_tuple(4, 0, std::get<2>(kv.first));
        v.derivs[key] = poly_add(v.derivs[key], kv.second);
      }
      return v;
    }
  }
  return v;
}

}  // namespace detail

// Lower an EVOLUTION right-hand side (or a provisional definition) into the
// descriptor's term list, in the order: bilaplacian, laplacian,
// gradlaplacian (by axis), pointwise; within each, psi before var(1), var(2).
inline int lower_terms(const Expr& rhs, sps_fd_term* terms, int max_terms,
                       const std::vector<double>& coeffs = {}) {
  detail::Value v = detail::lower_node(rhs.n, coeffs);
  int nt = 0;
  auto push = [&](int op, int axis, int operand, const Poly& p) {
    if (detail::is_zero(p)) return;
    if (nt >= max_terms)
      throw solver_error(SPS_E_UNSUPPORTED, "more than 8 terms");
    sps_fd_term& t = terms[nt++];
    std::memset(&t, 0, sizeof(t));
    t.op = op;
    t.axis = axis;
    t.operand = operand;
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
  auto each = [&](int order, auto fn) {
    for (int operand = 0; operand <= SPS_MAX_PROV; ++operand)
      for (auto& kv : v.derivs)
        if (std::get<0>(kv.first) == order && std::get<2>(kv.first) == operand)
          fn(std::get<1>(kv.first), operand, kv.second);
  };
  each(4, [&](int, int operand, const Poly& p) { push(SPS_OP_BILAP, 0, operand, p); });
  each(2, [&](int, int operand, const Poly& p) {
    // a laplacian of a sum is a sum of laplacians: one term per monomial
    for (int k = 1; k <= SPS_MAX_DEGREE; ++k)
      if (p.k[k] != 0.0) {
        Poly m;
        m.k[k] = p.k[k];
        push(SPS_OP_LAP, 0, operand, m);
      }
  });
  each(3, [&](int axis, int operand, const Poly& p) {
    push(SPS_OP_GRADLAP, axis, operand, p);
  });
  // pointwise part: psi's polynomial carries the constant
  Poly id0 = v.poly.count(0) ? v.poly.at(0) : Poly{};
  id0.k[0] += v.c0;
  push(SPS_OP_ID, 0, 0, id0);
  for (int operand = 1; operand <= SPS_MAX_PROV; ++operand)
    if (v.poly.count(operand)) push(SPS_OP_ID, 0, operand, v.poly.at(operand));
  return nt;
}

// ------------------------------------------------ several coupled fields

// Lowering of the right-hand side of one field of a multi-field model into
// sps_fdm_term's coeff * D[ prod_f psi_f^pow[f] ]: products are expanded into
// monomials of the fields, derivatives distribute over the monomials (one term
// per derivative of a monomial, its coefficient outside), lap(lap) is the
// bilaplacian.  Terms come out by category (bilaplacian, laplacian,
// gradlaplacian by axis, pointwise) and, within one, in the order in which
// their monomials first appear in the expression; equal monomials merge.
namespace detail {

using Mono = std::array<int, SPS_MAX_FIELDS>;

struct MValue {
  std::vector<std::pair<Mono, double>> poly;  // pointwise, constant included
  std::vector<std::pair<std::tuple<int, int, Mono>, double>> derivs;

  bool is_const() const {
    for (auto& kv : poly)
      if (kv.second != 0.0 && kv.first != Mono{}) return false;
    return derivs.empty();
  }
  double constant() const {
    for (auto& kv : poly)
      if (kv.first == Mono{}) return kv.second;
    return 0.0;
  }
};

template <typename K>
inline void madd(std::vector<std::pair<K, double>>& v, const K& k, double c) {
  for (auto& kv : v)
    if (kv.first == k) {
      kv.second += c;
      return;
    }
  v.emplace_back(k, c);
}

inline MValue mscale(MValue v, double s) {
  for (auto& kv : v.poly) kv.second *= s;
  for (auto& kv : v.derivs) kv.second *= s;
  return v;
}

inline MValue lower_mnode(const NodeP& n, const std::vector<double>& coeffs) {
  MValue v;
  switch (n->kind) {
    case Node::Const:
      madd(v.poly, Mono{}, n->value);
      return v;
    case Node::Coeff:
      madd(v.poly, Mono{},
           (n->index >= 1 && (size_t)n->index <= coeffs.size())
               ? coeffs[n->index - 1]
               : 1.0);
      return v;
    case Node::Field: {
      Mono m{};
      m[n->index - 1] = 1;
      madd(v.poly, m, 1.0);
      return v;
    }
    case Node::Var:
      throw solver_error(SPS_E_UNSUPPORTED,
                         "provisional variables in a multi-field model");
    case Node::Neg:
      return mscale(lower_mnode(n->a, coeffs), -1.0);
    case Node::Add: {
      MValue a = lower_mnode(n->a, coeffs), b = lower_mnode(n->b, coeffs);
      for (auto& kv : b.poly) madd(a.poly, kv.first, kv.second);
      for (auto& kv : b.derivs) madd(a.derivs, kv.first, kv.second);
      return a;
    }
    case Node::Mul: {
      MValue a = lower_mnode(n->a, coeffs), b = lower_mnode(n->b, coeffs);
      if (b.is_const()) return mscale(a, b.constant());
      if (a.is_const()) return mscale(b, a.constant());
      if (!a.derivs.empty() || !b.derivs.empty())
        throw solver_error(SPS_E_UNSUPPORTED,
                           "derivative multiplied by a non-constant");
      for (auto& x : a.poly)
        for (auto& y : b.poly) {
          if (x.second == 0.0 || y.second == 0.0) continue;
          Mono m{};
          int deg = 0;
          for (int f = 0; f < SPS_MAX_FIELDS; ++f) {
            m[f] = x.first[f] + y.first[f];
            deg += m[f];
          }
          if (deg > SPS_MAX_DEGREE)
            throw solver_error(SPS_E_UNSUPPORTED, "monomial of degree above 6");
          madd(v.poly, m, x.second * y.second);
        }
      return v;
    }
    case Node::Deriv: {
      MValue in = lower_mnode(n->a, coeffs);
      for (auto& kv : in.poly)
        if (kv.second != 0.0 && kv.first != Mono{})
          madd(v.derivs, std::make_tuple(n->order, n->axis, kv.first),
               kv.second);
      for (auto& kv : in.derivs) {
        const int inner = std::get<0>(kv.first);
        if (!(n->order == 2 && inner == 2))
          throw solver_error(SPS_E_UNSUPPORTED,
                             "derivative of order " +
                                 std::to_string(n->order + inner) +
                                 " other than lap(lap)");
        madd(v.derivs, std::make_tuple(4, 0, std::get<2>(kv.first)),
             kv.second);
      }
      return v;
    }
  }
  return v;
}

}  // namespace detail

inline int lower_mterms(const Expr& rhs, sps_fdm_term* terms, int max_terms,
                        const std::vector<double>& coeffs = {}) {
  detail::MValue v = detail::lower_mnode(rhs.n, coeffs);
  int nt = 0;
  auto push = [&](int op, int axis, const detail::Mono& m, double c) {
    if (c == 0.0) return;
    if (nt >= max_terms)
      throw solver_error(SPS_E_UNSUPPORTED,
                         "more than " + std::to_string(max_terms) + " terms");
    sps_fdm_term& t = terms[nt++];
    std::memset(&t, 0, sizeof(t));
    t.op = op;
    t.axis = axis;
    t.coeff = c;
    for (int f = 0; f < SPS_MAX_FIELDS; ++f) t.pow[f] = m[f];
  };
  for (int order : {4, 2})
    for (auto& kv : v.derivs)
      if (std::get<0>(kv.first) == order)
        push(order == 4 ? SPS_OP_BILAP : SPS_OP_LAP, 0, std::get<2>(kv.first),
             kv.second);
  for (int axis = 0; axis < 3; ++axis)
    for (auto& kv : v.derivs)
      if (std::get<0>(kv.first) == 3 && std::get<1>(kv.first) == axis)
        push(SPS_OP_GRADLAP, axis, std::get<2>(kv.first), kv.second);
  for (auto& kv : v.poly) push(SPS_OP_ID, 0, kv.first, kv.second);
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
  std::vector<ProvisionalDef> provisionals;  // var(1), var(2), ... in order

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
    // provisional variables, each lowered like a right-hand side
    // (PROVISIONAL_DEF; their grids live on the device with the plan)
    if (provisionals.size() > SPS_MAX_PROV)
      throw solver_error(SPS_E_UNSUPPORTED, "more than 2 provisional variables");
    d.nprov = (int)provisionals.size();
    for (size_t k = 0; k < provisionals.size(); ++k) {
      if (provisionals[k].index != (int)k + 1)
        throw solver_error(SPS_E_INVALID,
                           "provisional definitions must be var(1), var(2), "
                           "... in order");
      d.prov_nterms[k] =
          lower_terms(provisionals[k].def, d.prov[k], SPS_MAX_TERMS, coeffs);
    }
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
// equation follow model_iteration (modules-models.h:436-467).  Provisional
// variables are evaluated before every right-hand side, after the field's
// refresh, as Specialized::update does (modelspecialized.h:430-434).
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
        int path = SPS_PATH_AUTO, std::vector<double> coeffs = {},
        std::vector<ProvisionalDef> provisionals = {})
      : solver(Sp::make_solver(p)), system0(p), rhs(evolution) {
    solver.dtype = dtype;
    solver.path = path;
    solver.coeffs = std::move(coeffs);
    solver.provisionals = std::move(provisionals);
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

// MultiModel<D, Sp, F>: a model of F coupled fields (MODEL(MC, (SCALARS(2)),
// ...), GRAYSCOTT, Turing: examples/models/modeldefinitions.h:63-81,123-137),
// one EVOLUTION right-hand side per field, every field stepped from the same
// current fields (modelspecialized.h:383-385 then model.h:931-936).  The
// fields live in one device plan (sps_fdm_*); system(i) is field i's padded
// host mirror, refreshed on demand.
template <size_t D, typename Sp, size_t F>
struct MultiModel {
  static_assert(F >= 1 && F <= SPS_MAX_FIELDS, "1 .. SPS_MAX_FIELDS fields");
  Sp solver;
  std::array<int, D> dims{};
  std::array<int, D> padded{};
  std::array<std::vector<double>, F> values;  // padded host mirrors
  std::array<Expr, F> rhs;
  sps_fdm_plan* plan = nullptr;
  bool host_stale = false;
  long index = 0;
  double time = 0;
  double eq_dt = -1;

  // initial[f]: interior values of field f, x fastest (empty = zeros)
  MultiModel(const problem_parameters<D>& p, std::array<Expr, F> evolution,
             std::array<std::vector<double>, F> initial = {},
             int dtype = SPS_F64, int path = SPS_PATH_AUTO,
             std::vector<double> coeffs = {})
      : solver(Sp::make_solver(p)), dims(p.dims), rhs(std::move(evolution)) {
    solver.dtype = dtype;
    solver.path = path;
    solver.coeffs = std::move(coeffs);
    size_t len = 1, inner = 1;
    for (size_t a = 0; a < D; ++a) {
      padded[a] = dims[a] + 2 * SPS_BOUNDARY_DEPTH;
      len *= (size_t)padded[a];
      inner *= (size_t)dims[a];
    }
    for (size_t f = 0; f < F; ++f) {
      values[f].assign(len, 0.0);
      if (initial[f].empty()) continue;
      if (initial[f].size() != inner)
        throw solver_error(SPS_E_INVALID, "initial field size mismatch");
      const int B = SPS_BOUNDARY_DEPTH;
      const int nz = D == 3 ? dims[2] : 1;
      for (int z = 0; z < nz; ++z)
        for (int y = 0; y < dims[1]; ++y)
          for (int x = 0; x < dims[0]; ++x)
            values[f][(x + B) + (size_t)padded[0] *
                                    ((y + B) + (size_t)padded[1] *
                                                   (D == 3 ? z + B : 0))] =
                initial[f][x + (size_t)dims[0] * (y + (size_t)dims[1] * z)];
    }
  }
  MultiModel(const MultiModel&) = delete;
  MultiModel& operator=(const MultiModel&) = delete;
  ~MultiModel() {
    if (plan) sps_fdm_destroy(plan);
  }

  // field f's padded values with fresh ghosts
  const std::vector<double>& system(size_t f) {
    if (host_stale && plan) {
      for (size_t g = 0; g < F; ++g) {
        if (solver.dtype == SPS_F64) {
          check(sps_fdm_download(plan, (int)g, values[g].data(),
                                 SPS_LAYOUT_PADDED),
                "download");
        } else {
          std::vector<float> tmp(values[g].size());
          check(sps_fdm_download(plan, (int)g, tmp.data(), SPS_LAYOUT_PADDED),
                "download");
          for (size_t i = 0; i < tmp.size(); ++i) values[g][i] = tmp[i];
        }
      }
      host_stale = false;
    }
    return values[f];
  }
  long get_index() const { return index; }

  void equation() {}
  void step(double dt) { step_n(dt, 1); }
  void step_n(double dt, int n) {
    ensure_plan(dt);
    check(sps_fdm_step(plan, n), "step");
    host_stale = true;
    index += n;
  }
  void update(double t) { time = t; }

 private:
  // form_expr_one for every field into one descriptor; the plan is
  // (re)built when dt changes, carrying the current fields over
  void ensure_plan(double dt) {
    if (plan && dt == eq_dt) return;
    if (plan && host_stale) system(0);
    sps_fdm_desc d;
    std::memset(&d, 0, sizeof(d));
    d.dim = (int)D;
    d.dtype = solver.dtype;
    for (size_t a = 0; a < 3; ++a) d.n[a] = a < D ? dims[a] : 1;
    d.lap_points = Sp::lap_points;
    d.gradlap_points = Sp::gradlap_points;
    d.bilap_points = Sp::bilap_points;
    d.nfields = (int)F;
    d.path = solver.path;
    d.h = solver.h;
    d.dt = dt;
    for (size_t f = 0; f < F; ++f)
      d.nterms[f] = lower_mterms(rhs[f], d.terms[f], SPS_MAX_MTERMS,
                                 solver.coeffs);
    if (plan) sps_fdm_destroy(plan);
    plan = nullptr;
    check(sps_fdm_create(&d, &plan), "sps_fdm_create");
    for (size_t f = 0; f < F; ++f) {
      if (solver.dtype == SPS_F64) {
        check(sps_fdm_upload(plan, (int)f, values[f].data(), SPS_LAYOUT_PADDED),
              "upload");
      } else {
        std::vector<float> tmp(values[f].begin(), values[f].end());
        check(sps_fdm_upload(plan, (int)f, tmp.data(), SPS_LAYOUT_PADDED),
              "upload");
      }
    }
    host_stale = true;
    eq_dt = dt;
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
