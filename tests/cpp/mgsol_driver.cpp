// mgsol_driver.cpp -- drives the C++ host layer (include/femus_hip.hpp) the way a
// FEMuS application drives MGSolT / MGSolNS (template_appl/nse/SRC/main.C:
// MGDofBcOp -> GenIc -> MGTimeStep / MGSolve), and prints the result as JSON
// for the GPU tests.
//
//   mgsol_driver heat   <dim> <nx> <nlev>         steady heat, T = 0 on the boundary, f = 1
//   mgsol_driver cavity <nx> <nlev> <steps>       2D lid-driven cavity, Picard steps (Re 100)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "femus_hip.hpp"

using namespace femus_hip;

namespace {

// User_T.C-style application: Dirichlet T = 0 on every face, T_old = 0
class HeatApp : public MGSolT {
 public:
  using MGSolT::MGSolT;
  void bc_read(const double *, int bc_Neum[], int bc_flag[]) override { bc_Neum[0] = 0; bc_flag[0] = 0; }
  void ic_read(const double *, double u[]) override { u[0] = 0.0; }
};

// User_NS.C-style application: lid u = (1, 0) on y = 1, no-slip elsewhere, pressure free
class CavityApp : public MGSolNS {
 public:
  using MGSolNS::MGSolNS;
  void bc_read(const double xp[], int bc_Neum[], int bc_flag[]) override {
    const double tol = 1e-9;
    bc_Neum[0] = bc_Neum[1] = 0;
    bc_flag[0] = bc_flag[1] = 0;
    if (xp[1] > 1.0 - tol) bc_flag[0] = 4;
    if (xp[0] < tol || xp[0] > 1.0 - tol || xp[1] < tol) bc_flag[0] = bc_flag[1] = 0;
    bc_Neum[2] = 1;
    bc_flag[2] = 0;
  }
  void ic_read(const double xp[], double u[]) override {
    const double tol = 1e-9;
    u[0] = (xp[1] > 1.0 - tol && xp[0] > tol && xp[0] < 1.0 - tol) ? 1.0 : 0.0;
    u[1] = 0.0;
    u[2] = 0.0;
  }
};

void print_vec(const char *name, const std::vector<double> &v) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < v.size(); i++) std::printf(i ? ",%.17g" : "%.17g", v[i]);
  std::printf("]");
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: mgsol_driver heat|cavity ...\n"); return 2; }
  const std::string mode = argv[1];
  Context ctx(0);
  if (mode == "heat" && argc == 5) {
    const int dim = std::atoi(argv[2]), nx = std::atoi(argv[3]), nlev = std::atoi(argv[4]);
    const int nx0[3] = {nx, nx, dim == 3 ? nx : 1};
    const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, dim == 3 ? 1.0 : 0.0};
    Mesh mesh(dim, nx0, lo, hi, nlev);
    fm_heat_params p{};
    p.alpha0 = 1.0; p.rhocp = 1.0; p.dt = 0.1; p.time_coeff = 0.0; p.source = 1.0; p.upwind = 1.0;
    HeatApp T(ctx, mesh, p);
    T.GenBc();
    T.init();
    T.GenIc();
    T.opts.precision = FM_PREC_FP64_FGMRES;
    T.MGTimeStep(0.0, 1);
    std::printf("{\"mode\": \"heat\", \"ndof\": %lld, \"iters\": %d, \"resid\": %.17g, ", (long long)T._Dim.back(),
                T.report.iters, T.report.rel_resid);
    print_vec("x", T.x_old);
    std::printf("}\n");
    return T.report.converged ? 0 : 1;
  }
  if (mode == "cavity" && argc == 5) {
    const int nx = std::atoi(argv[2]), nlev = std::atoi(argv[3]), steps = std::atoi(argv[4]);
    const int nx0[3] = {nx, nx, 1};
    const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 0};
    Mesh mesh(2, nx0, lo, hi, nlev);
    fm_ns_params p{};
    p.IRe = 0.01; p.dt = 0.1; p.time_coeff = 1.0; p.upwind = 0.1;
    CavityApp ns(ctx, mesh, p);
    ns.GenBc();
    ns.init();
    ns.GenIc();
    ns.opts.precision = FM_PREC_FP64_FGMRES;
    ns.opts.nu_pre = ns.opts.nu_post = 4;
    ns.opts.tol = 1e-10;
    ns.opts.max_iter = 200;
    double du = 0.0;
    for (int k = 0; k < steps; k++) {
      std::vector<double> prev = ns.x_old;
      ns.MGTimeStep(0.1 * (k + 1), 1);
      double a = 0.0, b = 0.0;
      for (size_t i = 0; i < prev.size(); i++) { a += (ns.x_old[i] - prev[i]) * (ns.x_old[i] - prev[i]); b += ns.x_old[i] * ns.x_old[i]; }
      du = std::sqrt(a / b);
      if (!ns.report.converged) { std::fprintf(stderr, "step %d: not converged (%g)\n", k, ns.report.rel_resid); return 1; }
    }
    std::printf("{\"mode\": \"cavity\", \"ndof\": %lld, \"iters\": %d, \"resid\": %.17g, \"last_update\": %.17g, ",
                (long long)ns._Dim.back(), ns.report.iters, ns.report.rel_resid, du);
    print_vec("x", ns.x_old);
    std::printf("}\n");
    return 0;
  }
  std::fprintf(stderr, "bad arguments\n");
  return 2;
}
