// Test driver of the C++ drop-in adapter (symphas_amd/host/solverft_mi355x.hpp).
//   adapter_driver lower            print the lowered term lists (no GPU)
//   adapter_driver run <case> <out> run a case on the GPU, write the interior
//                                   field (FP64, x fastest) to <out>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../symphas_amd/host/solverft_mi355x.hpp"

using namespace sps::symphas;

static void print_terms(const char* name, const Expr& e,
                        const std::vector<double>& coeffs = {}) {
  sps_fd_term t[SPS_MAX_TERMS];
  int n = lower_terms(e, t, SPS_MAX_TERMS, coeffs);
  std::printf("%s", name);
  for (int i = 0; i < n; ++i) {
    std::printf(" | op=%d axis=%d operand=%d coeff=%.17g poly=", t[i].op, t[i].axis,
                t[i].operand, t[i].coeff);
    for (int k = 0; k <= SPS_MAX_DEGREE; ++k) std::printf("%s%.17g", k ? "," : "", t[i].poly[k]);
  }
  std::printf("\n");
}

static void print_mterms(const char* name, const Expr& e,
                         const std::vector<double>& coeffs = {}) {
  sps_fdm_term t[SPS_MAX_MTERMS];
  int n = lower_mterms(e, t, SPS_MAX_MTERMS, coeffs);
  std::printf("%s", name);
  for (int i = 0; i < n; ++i) {
    std::printf(" | op=%d axis=%d coeff=%.17g pow=", t[i].op, t[i].axis, t[i].coeff);
    for (int f = 0; f < SPS_MAX_FIELDS; ++f) std::printf("%s%d", f ? "," : "", t[i].pow[f]);
  }
  std::printf("\n");
}

// The right-hand sides of MC, GRAYSCOTT and Turing as written in
// examples/models/modeldefinitions.h, with c(i) the model coefficients.
static std::vector<std::pair<std::string, Expr>> coupled_rhs() {
  auto psi = op(1), rho = op(2);
  return {
      {"model_c_psi", -bilap(psi) - lap((c(1) - c(2) * psi * psi) * psi + c(5) * rho * rho)},
      {"model_c_rho", lap(rho) + (c(3) - c(4) * rho * rho) * rho + 2.0 * c(5) * psi * rho},
      {"gray_scott_psi", c(1) * lap(psi) - psi * rho * rho + c(3) * (1.0 - psi)},
      {"gray_scott_rho", c(2) * lap(rho) + psi * rho * rho - (c(3) + c(4)) * rho},
      {"turing_psi", c(1) * lap(psi) + c(2) * (psi + c(3) * rho - psi * rho * rho -
                                                 c(4) * psi * rho)},
      {"turing_rho", lap(rho) + c(2) * (c(5) * rho + c(6) * psi + psi * rho * rho +
                                          c(4) * psi * rho)},
  };
}

static Expr rhs_of(const std::string& name) {
  for (const auto& [n, e] : coupled_rhs())
    if (n == name) return e;
  throw solver_error(SPS_E_INVALID, "no right-hand side " + name);
}

template <size_t D, typename M>
static void write_fields(M& model, const std::array<int, D>& dims, size_t nf,
                         const char* path) {
  std::vector<double> out;
  const int B = SPS_BOUNDARY_DEPTH;
  const int nz = D == 3 ? dims[2] : 1;
  const int px = dims[0] + 2 * B, py = dims[1] + 2 * B;
  for (size_t f = 0; f < nf; ++f) {
    const auto& g = model.system(f);
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < dims[1]; ++y)
        for (int x = 0; x < dims[0]; ++x)
          out.push_back(g[(x + B) + (size_t)px * ((y + B) + (size_t)py * (D == 3 ? z + B : 0))]);
  }
  FILE* f = std::fopen(path, "wb");
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
}

static void print_error(const char* name, Expr (*make)()) {
  try {
    Expr e = make();
    sps_fd_term t[SPS_MAX_TERMS];
    lower_terms(e, t, SPS_MAX_TERMS);
    std::printf("%s | accepted\n", name);
  } catch (const solver_error& err) {
    std::printf("%s | error %d: %s\n", name, err.status, err.what());
  }
}

// SplitMix64 uniform field (same generator as symphas_amd/inputs.py)
static std::vector<double> uniform(size_t n, unsigned long long seed, double lo, double hi) {
  std::vector<double> v(n);
  unsigned long long s = seed;
  for (size_t i = 0; i < n; ++i) {
    s += 0x9E3779B97F4A7C15ull;
    unsigned long long z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    v[i] = lo + (hi - lo) * ((double)(z >> 11) * (1.0 / 9007199254740992.0));
  }
  return v;
}

template <size_t D>
static void write_interior(SolverSystemFD_MI355X<D>& sys, const char* path) {
  const auto& g = sys.as_grid();
  std::vector<double> out;
  const int B = SPS_BOUNDARY_DEPTH;
  const int nz = D == 3 ? sys.dims[2] : 1;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < sys.dims[1]; ++y)
      for (int x = 0; x < sys.dims[0]; ++x)
        out.push_back(g[(x + B) + (size_t)sys.padded[0] *
                                      ((y + B) + (size_t)sys.padded[1] * (D == 3 ? z + B : 0))]);
  FILE* f = std::fopen(path, "wb");
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc >= 2 && std::string(argv[1]) == "lower") {
    auto psi = op(1);
    // README Allen-Cahn: dop(1) = lap(op(1)) - (power(op(1), 2) - 1_n) * op(1)
    print_terms("allen_cahn", lap(psi) - (power(psi, 2) - 1.0) * psi);
    // MB: dpsi = -bilap(psi) - lap((c(1) - c(2) * psi * psi) * psi)
    print_terms("model_b", -bilap(psi) - lap((c(1) - c(2) * psi * psi) * psi));
    // MA: dpsi = lap(psi) + (c(1) - c(2) * psi * psi) * psi
    print_terms("model_a", lap(psi) + (c(1) - c(2) * psi * psi) * psi);
    // lap applied twice becomes a bilaplacian
    print_terms("model_b_coeffs", -bilap(psi) - lap((c(1) - c(2) * psi * psi) * psi),
                {0.7, 1.3});
    print_terms("lap_lap", lap(lap(psi)) + 0.5 * lap(psi * psi));
    print_terms("gradlap_y", gradlap<Axis::Y>(psi * psi) * 2.0);
    print_error("derivative_product", [] { auto p = op(1); return lap(p) * lap(p); });
    print_error("derivative_times_field", [] { auto p = op(1); return lap(p) * p; });
    print_error("power7", [] { auto p = op(1); return power(p, 7); });
    print_error("second_field", [] { return op(2); });
    // provisional variables: MBB (examples/models/modeldefinitions.h:150-155)
    // and BB (tests/testdefs.h:29-37) with c(1) = 1.1, c(2) = 0.4
    const std::vector<double> cf = {1.1, 0.4};
    print_terms("mbb_evolution", -bilap(psi) + lap(var(1)), cf);
    print_terms("mbb_var1", -(c(1) - 4.0 * c(2) * psi * psi) * psi, cf);
    print_terms("bb_evolution", -lap(var(1)), cf);
    print_terms("bb_var1", lap(psi) + (c(1) - 4.0 * c(2) * psi * psi) * psi, cf);
    print_terms("mixed_operands", 0.5 * var(2) + psi - 2.0 + lap(var(1) * var(1)));
    print_error("field_times_var", [] { auto p = op(1); return p * var(1); });
    print_error("var3", [] { return var(3); });
    // coupled fields (examples/models/modeldefinitions.h:74-80,123-136)
    const std::vector<double> mc = {1.1, 0.8, 0.6, 1.3, 0.4};
    for (const auto& [name, e] : coupled_rhs())
      print_mterms(name.c_str(), e, mc);
    print_error("second_field_single", [] { return op(2); });
    print_error("field5", [] { return op(5); });
    return 0;
  }
  if (argc < 4 || std::string(argv[1]) != "run") {
    std::fprintf(stderr, "usage: adapter_driver lower | run <case> <out>\n");
    return 2;
  }
  const std::string which = argv[2];
  try {
    auto psi = op(1);
    if (which == "allen_cahn") {
      // README.md:83-97: 64x64 periodic, h = 0.5, dt = 0.01, 1000 steps
      problem_parameters<2> p;
      p.dims = {64, 64};
      p.h = 0.5;
      p.initial = uniform(64 * 64, 1, -0.1, 0.1);
      Model<2, SolverFT_MI355X<Stencil2d2h<5, 6, 13>>> model(
          p, lap(psi) - (power(psi, 2) - 1.0) * psi, SPS_F64, SPS_PATH_REFERENCE);
      find_solution(model, 0.01, 1000);
      write_interior(model.system(), argv[3]);
    } else if (which == "model_b_steps") {
      // stepping one iteration at a time, as model_iteration does
      problem_parameters<2> p;
      p.dims = {48, 40};
      p.h = 1.0;
      p.initial = uniform(48 * 40, 2, -0.1, 0.1);
      Model<2, SolverFT_MI355X<Stencil2d2h<9, 6, 13>>> model(
          p, -bilap(psi) - lap((c(1) - c(2) * psi * psi) * psi), SPS_F64,
          SPS_PATH_REFERENCE);
      model.update(0.0);
      for (int i = 0; i < 50; ++i) {
        model.equation();
        model.step(0.01);
        model.update(0.01 * (i + 1));
      }
      write_interior(model.system(), argv[3]);
    } else if (which == "model_b_3d_fp32") {
      problem_parameters<3> p;
      p.dims = {32, 24, 16};
      p.h = 1.0;
      p.initial = uniform(32 * 24 * 16, 3, -0.05, 0.05);
      Model<3, SolverFT_MI355X<Stencil3d2h<7, 10, 25>>> model(
          p, -bilap(psi) - lap((c(1) - c(2) * psi * psi) * psi), SPS_F32,
          SPS_PATH_AUTO);
      find_solution(model, 0.005, 40);
      write_interior(model.system(), argv[3]);
    } else if (which == "mbb" || which == "bb") {
      // provisional variables, FP64; MBB on the reference-order path (var
      // grid pass, then the step), BB on whichever path AUTO picks (fused)
      problem_parameters<2> p;
      p.dims = {40, 28};
      p.h = 0.9;
      p.initial = uniform(40 * 28, 4, -0.5, 0.5);
      const std::vector<double> cf = {1.1, 0.4};
      if (which == "mbb") {
        Model<2, SolverFT_MI355X<Stencil2d2h<5, 6, 13>>> model(
            p, -bilap(psi) + lap(var(1)), SPS_F64, SPS_PATH_REFERENCE, cf,
            {var(1) <= -(c(1) - 4.0 * c(2) * psi * psi) * psi});
        find_solution(model, 0.002, 30);
        write_interior(model.system(), argv[3]);
      } else {
        Model<2, SolverFT_MI355X<Stencil2d2h<5, 6, 13>>> model(
            p, -lap(var(1)), SPS_F64, SPS_PATH_AUTO, cf,
            {var(1) <= lap(psi) + (c(1) - 4.0 * c(2) * psi * psi) * psi});
        find_solution(model, 0.002, 30);
        write_interior(model.system(), argv[3]);
      }
    } else if (which == "model_c" || which == "gray_scott" || which == "turing" ||
               which == "gray_scott_fused") {
      // coupled fields, FP64: the reference-order path (bit for bit against
      // the oracle), and Gray-Scott on the fused two-field path
      const bool fused = which == "gray_scott_fused";
      const std::string base = fused ? "gray_scott" : which;
      problem_parameters<2> p;
      p.dims = {40, 28};
      p.h = 1.0;
      std::vector<double> cf;
      double dt = 0.002;
      if (base == "model_c") cf = {1.1, 0.8, 0.6, 1.3, 0.4};
      if (base == "gray_scott") cf = {0.16, 0.08, 0.035, 0.065}, dt = 0.5;
      if (base == "turing") cf = {1.0, 0.7, 1.2, 0.5, -0.4, 0.9}, dt = 0.01;
      std::array<std::vector<double>, 2> init = {uniform(40 * 28, 5, -0.5, 0.5),
                                                 uniform(40 * 28, 6, -0.5, 0.5)};
      MultiModel<2, SolverFT_MI355X<Stencil2d2h<5, 6, 13>>, 2> model(
          p, {rhs_of(base + "_psi"), rhs_of(base + "_rho")}, init, SPS_F64,
          fused ? SPS_PATH_FUSED : SPS_PATH_REFERENCE, cf);
      find_solution(model, dt, 30);
      write_fields<2>(model, p.dims, 2, argv[3]);
    } else {
      std::fprintf(stderr, "unknown case %s\n", which.c_str());
      return 2;
    }
  } catch (const solver_error& e) {
    std::fprintf(stderr, "solver_error %d: %s\n", e.status, e.what());
    return 1;
  }
  return 0;
}
