// Driver of the SolverFT_MI355X SymPhas plugin header, compiled against the
// test double of the SymPhas API (tests/host/symphas_api_double).  Each model
// is written as the canonical expression type SymPhas hands form_expr_one
// (OpAdd of OpTerms and OpDerivative<Solver::derivative<ax, O>, ...>):
//   lower <model>                   print the lowered term list
//   run <model> <n0> <n1> <steps> <in> <out>
//                                   step a seeded field through the plugin
//                                   (Model::step / Model::update order) and
//                                   write the interior
//   coupled                         lower Gray-Scott's first equation, which
//                                   reads the second field: must be refused
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "symphas_plugin/solverft_mi355x.h"

using Sys = SolverSystemFD_MI355X<double, 2>;
using Sv = SolverFT_MI355X<Stencil2d2h<5, 6, 13>>;
using Sv9 = SolverFT_MI355X<Stencil2d2h<9, 6, 17>>;
template <typename S, Axis ax, size_t O>
using Der = typename Solver<S, 0>::template derivative<ax, O>;
using Psi = Variable<0, symphas::ref<Sys>>;
using Rho = Variable<1, symphas::ref<Sys>>;

// Model B: -bilap(psi) - c1*lap(psi) + c2*lap(psi^3)
// (examples/models/modeldefinitions.h:53-55 after apply_operators)
template <typename S>
auto model_b(Psi f, double c1, double c2) {
  using E = OpAdd<OpDerivative<Der<S, Axis::X, 4>, OpNegIdentity, OpTerm<OpIdentity, Psi>, S>,
                  OpDerivative<Der<S, Axis::X, 2>, OpLiteral<double>, OpTerm<OpIdentity, Psi>, S>,
                  OpDerivative<Der<S, Axis::X, 2>, OpLiteral<double>,
                               OpTerms<OpIdentity, Term<Psi, 3>>, S>>;
  E e;
  std::get<0>(e.terms).data = f;
  std::get<1>(e.terms).value = OpLiteral<double>{-c1};
  std::get<1>(e.terms).data = f;
  std::get<2>(e.terms).value = OpLiteral<double>{c2};
  return e;
}

// Allen-Cahn: lap(psi) + psi - psi^3 (README.md:83-85)
template <typename S>
auto allen_cahn(Psi f) {
  using E = OpAdd<OpDerivative<Der<S, Axis::X, 2>, OpIdentity, OpTerm<OpIdentity, Psi>, S>,
                  OpTerm<OpIdentity, Psi>, OpTerms<OpNegIdentity, Term<Psi, 3>>>;
  E e;
  std::get<0>(e.terms).data = f;
  return e;
}

// A gradlaplacian along y of psi^2 plus a constant source: exercises order 3,
// the axis and a standalone literal.
template <typename S>
auto gradlap_source(Psi) {
  using E = OpAdd<OpDerivative<Der<S, Axis::Y, 3>, OpFractionLiteral<1, 2>,
                               OpTerms<OpIdentity, Term<Psi, 2>>, S>,
                  OpLiteral<double>>;
  E e;
  std::get<1>(e.terms).value = 0.25;
  return e;
}

symphas::problem_parameters_type params(int n0, int n1, double h) {
  symphas::problem_parameters_type p;
  p.dim = 2;
  symphas::interval_data d;
  d.ax[0] = {h, n0};
  d.ax[1] = {h, n1};
  d.ax[2] = {h, 1};
  p.systems.push_back(d);
  return p;
}

template <typename S, typename E>
void print_lowered(Sys& sys, S& solver, E eq) {
  Psi f;
  static_cast<symphas::ref<Sys>&>(f).p = &sys;
  solver.template form_expr_one<0>(0, std::make_pair(f, eq));
  const sps_fd_desc& d = sys.dev->desc;
  std::printf("dim %d n %d %d L %d G %d B %d h %.17g dt %.17g nterms %d\n", d.dim, d.n[0],
              d.n[1], d.lap_points, d.gradlap_points, d.bilap_points, d.h, d.dt, d.nterms);
  for (int t = 0; t < d.nterms; ++t) {
    std::printf("term %d %d %.17g", d.terms[t].op, d.terms[t].axis, d.terms[t].coeff);
    for (int k = 0; k <= SPS_MAX_DEGREE; ++k) std::printf(" %.17g", d.terms[t].poly[k]);
    std::printf("\n");
  }
}

Sys make_system(int n0, int n1, const std::vector<double>& interior) {
  symphas::init_data_type tdata{interior};
  symphas::interval_data_type vdata;
  vdata.n[0] = n0;
  vdata.n[1] = n1;
  return Sys(tdata, vdata, symphas::b_data_type{}, 0);
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string mode = argv[1];
  try {
    if (mode == "lower") {
      const std::string model = argc > 2 ? argv[2] : "mb";
      Sys sys = make_system(16, 12, std::vector<double>(16 * 12, 0.0));
      if (model == "mb9") {
        auto solver = Sv9::make_solver(params(16, 12, 1.0));
        solver.dt = 0.005;
        Psi f;
        print_lowered(sys, solver, model_b<Sv9>(f, 1.0, 1.0));
      } else {
        auto solver = Sv::make_solver(params(16, 12, model == "ac" ? 0.5 : 1.0));
        solver.dt = 0.01;
        Psi f;
        if (model == "mb") print_lowered(sys, solver, model_b<Sv>(f, 0.7, 1.3));
        if (model == "ac") print_lowered(sys, solver, allen_cahn<Sv>(f));
        if (model == "gl") print_lowered(sys, solver, gradlap_source<Sv>(f));
      }
      return 0;
    }
    if (mode == "coupled") {
      // Gray-Scott's first equation: c1*lap(psi) - psi*rho^2 + c3 - c3*psi
      using E = OpAdd<OpDerivative<Der<Sv, Axis::X, 2>, OpLiteral<double>,
                                   OpTerm<OpIdentity, Psi>, Sv>,
                      OpTerms<OpNegIdentity, Term<Psi, 1>, Term<Rho, 2>>, OpLiteral<double>,
                      OpTerms<OpLiteral<double>, Term<Psi, 1>>>;
      Sys sys = make_system(16, 12, std::vector<double>(16 * 12, 0.0));
      auto solver = Sv::make_solver(params(16, 12, 1.0));
      solver.dt = 0.01;
      E e;
      std::get<0>(e.terms).value = OpLiteral<double>{0.2};
      std::get<2>(e.terms).value = 0.04;
      std::get<3>(e.terms).term = OpLiteral<double>{-0.04};
      Psi f;
      static_cast<symphas::ref<Sys>&>(f).p = &sys;
      solver.template form_expr_one<0>(0, std::make_pair(f, e));
      std::printf("lowered\n");
      return 0;
    }
    if (mode == "run" && argc == 8) {
      const std::string model = argv[2];
      const int n0 = std::atoi(argv[3]), n1 = std::atoi(argv[4]), steps = std::atoi(argv[5]);
      std::vector<double> in((size_t)n0 * n1);
      std::ifstream fi(argv[6], std::ios::binary);
      fi.read(reinterpret_cast<char*>(in.data()), (std::streamsize)(in.size() * sizeof(double)));
      if (!fi) return 3;
      Sys sys = make_system(n0, n1, in);
      const double h = model == "ac" ? 0.5 : 1.0, dt = 0.01;
      auto solver = Sv::make_solver(params(n0, n1, h));
      Psi f;
      static_cast<symphas::ref<Sys>&>(f).p = &sys;
      if (model == "ac")
        solver.template form_expr_one<0>(0, std::make_pair(f, allen_cahn<Sv>(f)));
      else
        solver.template form_expr_one<0>(0, std::make_pair(f, model_b<Sv>(f, 1.0, 1.0)));
      // run_model: update(time), equation, step (modules-models.h:460-466)
      double time = 0;
      for (int i = 0; i < steps; ++i) {
        sys.update(i, time);
        solver.dt = dt;  // Model::step sets dt before stepping (model.h:932-936)
        solver.step(sys);
        time += dt;
      }
      sys.update(steps, time);
      std::vector<double> out((size_t)n0 * n1);
      for (size_t i = 0; i < out.size(); ++i) out[i] = sys.values[sys.padded(i)];
      std::ofstream fo(argv[7], std::ios::binary);
      fo.write(reinterpret_cast<const char*>(out.data()), (std::streamsize)(out.size() * sizeof(double)));
      return fo ? 0 : 4;
    }
  } catch (const sps_bind::plugin_error& e) {
    std::printf("plugin_error %d %s\n", e.status, e.what());
    return 5;
  }
  return 2;
}
