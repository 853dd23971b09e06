// Must not compile: a first-order derivative is outside the plugin's
// vocabulary, and the static_assert says so.
#include "symphas_plugin/solverft_mi355x.h"

using Sys = SolverSystemFD_MI355X<double, 2>;
using Sv = SolverFT_MI355X<Stencil2d2h<5, 6, 13>>;
using Psi = Variable<0, symphas::ref<Sys>>;
using Grad = Solver<Sv, 0>::derivative<Axis::X, 1>;

int main() {
  OpDerivative<Grad, OpIdentity, OpTerm<OpIdentity, Psi>, Sv> e{};
  std::vector<sps_bind::Term> terms;
  sps_bind::lower(e, 1.0, terms);
  return 0;
}
