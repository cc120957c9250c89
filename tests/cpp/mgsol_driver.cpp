// mgsol_driver.cpp -- drives the C++ host layer (include/femus_hip.hpp) the way a
// FEMuS application drives MGSolT / MGSolNS (template_appl/nse/SRC/main.C:
// MGDofBcOp -> GenIc -> MGTimeStep / MGSolve), and prints the result as JSON
// for the GPU tests.
//
//   mgsol_driver heat   <dim> <nx> <nlev>         steady heat, T = 0 on the boundary, f = 1
//   mgsol_driver cavity <nx> <nlev> <steps>       2D lid-driven cavity, Picard steps (Re 100)
//   mgsol_driver dhc <dim> <nx> <nlev> <Ra> <steps>  Boussinesq differentially heated box: NS and heat stepped in
//                                                 turn (MGEquationsSystem::eqnmap_timestep_loop), then the Nusselt
//                                                 number and its conductivity gradient (steady heat adjoint at the
//                                                 frozen velocity)
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

// differentially heated box: no-slip walls, pressure free; T = 1 at x = 0, T = 0 at x = 1 (values from x_old),
// adiabatic elsewhere
class DhcNS : public MGSolNS {
 public:
  using MGSolNS::MGSolNS;
  void bc_read(const double *, int bc_Neum[], int bc_flag[]) override {
    const int dim = mesh_.dim();
    for (int d = 0; d < dim; d++) { bc_Neum[d] = 0; bc_flag[d] = 0; }
    bc_Neum[dim] = 1;
    bc_flag[dim] = 0;
  }
  void ic_read(const double *, double u[]) override {
    for (int d = 0; d <= mesh_.dim(); d++) u[d] = 0.0;
  }
};
class DhcT : public MGSolT {
 public:
  using MGSolT::MGSolT;
  void bc_read(const double xp[], int bc_Neum[], int bc_flag[]) override {
    const double tol = 1e-9;
    const bool wall = xp[0] < tol || xp[0] > 1.0 - tol;
    bc_Neum[0] = wall ? 0 : 1;
    bc_flag[0] = wall ? 1 : 0;
  }
  void ic_read(const double xp[], double u[]) override { u[0] = 1.0 - xp[0]; }
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
  if (mode == "dhc" && argc == 7) {
    const int dim = std::atoi(argv[2]), nx = std::atoi(argv[3]), nlev = std::atoi(argv[4]);
    const double Ra = std::atof(argv[5]), Pr = 0.71;
    const int steps = std::atoi(argv[6]);
    const int nx0[3] = {nx, nx, dim == 3 ? nx : 1};
    const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, dim == 3 ? 1.0 : 0.0};
    Mesh mesh(dim, nx0, lo, hi, nlev);
    fm_ns_params pn{};
    pn.IRe = std::sqrt(Pr / Ra); pn.dt = 1.0; pn.time_coeff = 1.0; pn.upwind = 0.1;
    fm_heat_params pt{};
    pt.alpha0 = 1.0 / std::sqrt(Ra * Pr); pt.rhocp = 1.0; pt.dt = 1.0; pt.time_coeff = 1.0; pt.source = 0.0; pt.upwind = 1.0;
    DhcNS ns(ctx, mesh, pn);
    ns.GenBc(); ns.init(); ns.GenIc();
    DhcT T(ctx, mesh, pt);
    T.GenBc(); T.init(); T.GenIc();
    ns.opts.precision = FM_PREC_FP32_FGMRES; ns.opts.nu_pre = ns.opts.nu_post = 3; ns.opts.coarse_mult = 1;
    ns.opts.tol = 1e-9; ns.opts.max_iter = 200;
    T.opts.precision = FM_PREC_FP32_FGMRES; T.opts.nu_pre = T.opts.nu_post = 3; T.opts.cheb_lo = 0.1;
    T.opts.tol = 1e-10; T.opts.max_iter = 200;
    double buoy[3] = {0, 0, 0};
    buoy[dim - 1] = 1.0;
    for (int k = 0; k < steps; k++) {
      ns.set_buoyancy(T.node_values(T.x_old), buoy);
      ns.MGTimeStep(k + 1.0, 1);
      if (!ns.report.converged) { std::fprintf(stderr, "NS step %d: %g\n", k, ns.report.rel_resid); return 1; }
      T.set_velocity(ns.velocity_nodes(ns.x_old));
      T.MGTimeStep(k + 1.0, 1);
      if (!T.report.converged) { std::fprintf(stderr, "T step %d: %g\n", k, T.report.rel_resid); return 1; }
    }
    const double Nu = T.nusselt(T.x_old, 0);
    // steady heat at the frozen velocity, its Nusselt number and dNu/dk_e by one transposed solve
    const int top = nlev - 1;
    T.par.time_coeff = 0.0;
    T.set_xold(T.x_old);
    for (int L = 0; L <= top; L++) {
      T.GenMatRhs(0.0, L, L == top ? 1 : 0);
      T.GenMatTranspose(L);
    }
    T.opts.precision = FM_PREC_FP64_FGMRES; T.opts.tol = 1e-12;
    T.set_x(T.x_old);
    T.MGSolve(1e-12, 200, 1, 3, 40, 3);
    const std::vector<double> Ts = T.get_x();
    std::vector<double> dJdT, dJdk;
    const double Nus = T.nusselt(Ts, 0, &dJdT, &dJdk);
    const std::vector<double> lam = T.AdjointSolve(dJdT);
    std::vector<double> g = T.grad_coef(lam, Ts);
    for (size_t e = 0; e < g.size(); e++) g[e] += dJdk[e];
    std::printf("{\"mode\": \"dhc\", \"Nu\": %.17g, \"Nu_steady\": %.17g, ", Nu, Nus);
    print_vec("grad", g);
    std::printf("}\n");
    return 0;
  }
  std::fprintf(stderr, "bad arguments\n");
  return 2;
}
