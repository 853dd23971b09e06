// Test double of the SymPhas expression types that the SolverFT_MI355X
// plugin header names (symphas_amd/host/symphas_plugin/solverft_mi355x.h).
// Written for this test, not taken from SymPhas: each type has the template
// parameters and the members the plugin reads, as the reference declares
// them (sym/inc/expressionprototypes.h:231-330), and nothing else.  The
// canonicalising transforms are identities here, so a test builds the
// expression in its canonical form directly.
#pragma once

#include <cstddef>
#include <tuple>
#include <type_traits>
#include <utility>

using iter_type = int;
using len_type = int;

enum class Axis { X = 0, Y = 1, Z = 2, NONE = 3 };

namespace expr {
using exp_key_t = unsigned int;
// power key: numerator in bits 0-15, denominator - 1 in bits 16-30, sign in
// bit 31 (the double's own encoding; the plugin reads N, D and sign only)
template <exp_key_t X>
struct _Xk_t {
  static const exp_key_t N = X & 0xffffu;
  static const exp_key_t D = ((X >> 16) & 0x7fffu) + 1;
  static const bool sign = (X >> 31) != 0;
};
}  // namespace expr

struct OpIdentity {
  constexpr double eval(iter_type = 0) const { return 1.0; }
};
struct OpNegIdentity {
  constexpr double eval(iter_type = 0) const { return -1.0; }
};
struct OpVoid {
  constexpr double eval(iter_type = 0) const { return 0.0; }
};
template <typename T>
struct OpLiteral {
  T value;
  constexpr T eval(iter_type = 0) const { return value; }
};
template <size_t N, size_t D>
struct OpFractionLiteral {
  constexpr double eval(iter_type = 0) const { return double(N) / D; }
};
template <size_t N, size_t D>
struct OpNegFractionLiteral {
  constexpr double eval(iter_type = 0) const { return -double(N) / D; }
};

namespace symphas {
template <typename T>
struct ref {
  T* p = nullptr;
  T& get() const { return *p; }
};
}  // namespace symphas

template <size_t Z, typename G = void>
struct Variable : G {
  Variable() = default;
  explicit Variable(G g) : G(g) {}
};
template <size_t Z>
struct Variable<Z, void> {};

template <typename G, expr::exp_key_t X = 1>
struct Term {
  G data;
};

template <typename... Ts>
struct OpTerms;
template <typename V, typename... Gs, expr::exp_key_t... Xs>
struct OpTerms<V, Term<Gs, Xs>...> {
  V term;
  std::tuple<Term<Gs, Xs>...> factors;
};
template <typename V, typename G>
using OpTerm = OpTerms<V, Term<G, 1>>;

template <typename... Es>
struct OpAdd {
  std::tuple<Es...> terms;
};

template <typename Dd, typename V, typename E, typename Sp>
struct OpDerivative {
  static const size_t order = Dd::order;
  static const Axis axis = Dd::axis;
  V value;
  E e;
};
template <typename Dd, typename V, typename G, typename Sp>
struct OpDerivative<Dd, V, OpTerm<OpIdentity, G>, Sp> {
  static const size_t order = Dd::order;
  static const Axis axis = Dd::axis;
  V value;
  G data;
};

namespace expr {
template <typename T>
double eval(T const& v) {
  if constexpr (std::is_arithmetic<T>::value)
    return double(v);
  else
    return double(v.eval());
}
template <size_t I, typename... Es>
const auto& get(OpAdd<Es...> const& e) {
  return std::get<I>(e.terms);
}
template <typename... Ts>
decltype(auto) coeff(OpTerms<Ts...> const& e) {
  return (e.term);
}
template <typename E>
E apply_operators(E const& e) {
  return e;
}
namespace transform {
template <typename E>
E optimize(E const& e) {
  return e;
}
}  // namespace transform
namespace prune {
template <typename E>
void update(E&) {}
}  // namespace prune
}  // namespace expr
