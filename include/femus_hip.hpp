// femus_hip.hpp -- C++ host layer over the C ABI of femus_hip.h that keeps the
// reference's operator interface for this path: MGSolDA / MGSolT / MGSolNS with
// the application callbacks bc_read / ic_read and the entry points GenBc,
// GenIc, init, GenMatRhs(time, Level, mode), MGSolve(Eps, MaxIter, Gamma,
// Nc_pre, Nc_coarse, Nc_post) and MGTimeStep(time, mode).
//
//   MGSolBase        include/MGSolverBase.h:24-389 (MGSolve :333-340, MGStep :351-359,
//                    MGTimeStep :308-311), src/MGSolverBase.C:163-289, 426-454
//   MGSolDA          include/MGSolverDA.h:55-376 (GenBc / GenIc / init / init_dof),
//                    src/MGSolverDA.C:552-948, 1690-2444
//   MGSolT / MGSolNS template_appl/nse/SRC/MGSolverT.C:93-530, MGSolverNS.C:108-756
//
// Errors follow the reference: it reports on std::cerr and aborts
// (laspack_linear_solverM.C:118-122); MGSolve itself never fails, it returns
// the last residual.  The handler is replaceable (set_error_handler) so a
// test can turn failures into exceptions.  Header only; link libfemus_hip.so.
#ifndef FEMUS_HIP_HPP
#define FEMUS_HIP_HPP
#include <algorithm>
#include <cstdlib>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "femus_hip.h"

namespace femus_hip {

using ErrorHandler = std::function<void(int code, const std::string &where, const std::string &msg)>;

inline ErrorHandler &error_handler() {
  static ErrorHandler h = [](int code, const std::string &where, const std::string &msg) {
    std::cerr << "femus_hip: " << where << " failed (" << code << "): " << msg << std::endl;
    std::abort();
  };
  return h;
}
inline void set_error_handler(ErrorHandler h) { error_handler() = std::move(h); }
inline void check(int rc, const char *where) {
  if (rc != FM_OK) error_handler()(rc, where, fm_last_error() ? fm_last_error() : "");
}
// true when a host vector holds the number of values the library reads through its pointer; otherwise reports
// FM_EINVAL through the error handler and returns false (the call is then skipped)
inline bool sized(const std::vector<double> &v, size_t n, const char *where) {
  if (v.size() == n) return true;
  error_handler()(FM_EINVAL, where,
                  "expected " + std::to_string(n) + " values, got " + std::to_string(v.size()));
  return false;
}

// One HIP device + stream (the reference's one MPI rank).
class Context {
 public:
  explicit Context(int device = 0) { check(fm_ctx_create(device, &h_), "fm_ctx_create"); }
  ~Context() { if (h_) fm_ctx_destroy(h_); }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  fm_ctx *get() const { return h_; }

 private:
  fm_ctx *h_ = nullptr;
};

// gencase box mesh hierarchy (MGGenCase::GenCase + printMesh, MGMesh::read_c).
class Mesh {
 public:
  Mesh(int dim, const int nx0[3], const double lo[3], const double hi[3], int nlev) {
    check(fm_mesh_create_box(dim, nx0, lo, hi, nlev, &h_), "fm_mesh_create_box");
    int info[6];
    check(fm_mesh_info(h_, info), "fm_mesh_info");
    dim_ = info[0]; nlev_ = info[1]; n_nodes_ = info[5];
    xyz_.resize((size_t)dim_ * n_nodes_);
    check(fm_mesh_coords(h_, xyz_.data()), "fm_mesh_coords");
  }
  ~Mesh() { if (h_) fm_mesh_destroy(h_); }
  Mesh(const Mesh &) = delete;
  Mesh &operator=(const Mesh &) = delete;
  fm_mesh *get() const { return h_; }
  int dim() const { return dim_; }
  int nlev() const { return nlev_; }
  int n_nodes() const { return n_nodes_; }
  const std::vector<double> &xyz() const { return xyz_; }   // xyz[node + d * n_nodes]

 private:
  fm_mesh *h_ = nullptr;
  int dim_ = 0, nlev_ = 0, n_nodes_ = 0;
  std::vector<double> xyz_;
};

// MGSolDA: one equation system on the whole hierarchy.
class MGSolDA {
 public:
  // MGSolDA::MGSolDA + init_dof: nq2 Q2 variables followed by nq1 Q1 variables
  MGSolDA(Context &ctx, Mesh &mesh, int nq2, int nq1) : mesh_(mesh), nq2_(nq2), nq1_(nq1) {
    check(fm_system_create(ctx.get(), mesh.get(), nq2, nq1, &h_), "fm_system_create");
    _NoLevels = mesh.nlev();
    _Dim.resize(_NoLevels);
    for (int L = 0; L < _NoLevels; L++) _Dim[L] = fm_system_ndof(h_, L);
    _node_dof.resize(_NoLevels);
    for (int L = 0; L < _NoLevels; L++) {
      _node_dof[L].resize((size_t)(nq2 + nq1) * mesh.n_nodes());
      check(fm_system_node_dof(h_, L, _node_dof[L].data()), "fm_system_node_dof");
    }
    opts = fm_mg_opts{FM_PREC_FP32_FGMRES, FM_SMOOTH_CHEBYSHEV, 2, 2, 1, 40, 30, 1e-8, 0.1, 1.05, 0.6, 0, 1, 1, 0};
  }
  virtual ~MGSolDA() { if (h_) fm_system_destroy(h_); }
  MGSolDA(const MGSolDA &) = delete;
  MGSolDA &operator=(const MGSolDA &) = delete;

  // application callbacks (User_NS.C / User_T.C): boundary condition flags and initial values at a node
  virtual void bc_read(const double xp[], int bc_Neum[], int bc_flag[]) = 0;
  virtual void ic_read(const double xp[], double u_value[]) = 0;

  // MGSolDA::GenBc (src/MGSolverDA.C:696-862)
  void GenBc() {
    check(fm_system_genbc(h_, &MGSolDA::bc_trampoline, this), "fm_system_genbc");
    bc[0].assign(_Dim.back(), 0);
    bc[1].assign(_Dim.back(), 0);
    check(fm_system_get_bc(h_, bc[0].data(), bc[1].data()), "fm_system_get_bc");
  }
  // MGSolDA::init (src/MGSolverDA.C:1690-1764): A, Rst, Prl on every level
  void init() { check(fm_system_init(h_), "fm_system_init"); }
  // MGSolDA::GenIc (src/MGSolverDA.C:552-692): x_old of the top level from ic_read at every node
  void GenIc() {
    const int top = _NoLevels - 1, N = mesh_.n_nodes(), dim = mesh_.dim(), nv = nq2_ + nq1_;
    std::vector<double> x(_Dim[top], 0.0), u(nv);
    for (int k = 0; k < N; k++) {
      double xp[3] = {0, 0, 0};
      for (int d = 0; d < dim; d++) xp[d] = mesh_.xyz()[k + (size_t)d * N];
      ic_read(xp, u.data());
      for (int iv = 0; iv < nv; iv++) {
        const int d = _node_dof[top][k + (size_t)iv * N];
        if (d >= 0) x[d] = u[iv];
      }
    }
    x_old = x;
    check(fm_system_set_xold(h_, x_old.data()), "fm_system_set_xold");
  }

  virtual void GenMatRhs(double time, int Level, int mode) = 0;

  // MGSolBase::MGSolve (src/MGSolverBase.C:163-208): cycles on the top level until the relative
  // residual ||b - A x|| / ||b|| <= Eps or MaxIter; returns that residual (never aborts on it)
  virtual double MGSolve(double Eps, int MaxIter, int Gamma = 1, int Nc_pre = 8, int Nc_coarse = 40,
                         int Nc_post = 8) {
    (void)Nc_coarse;  // the coarse level is solved directly
    fm_mg_opts o = opts;
    o.tol = Eps;
    o.max_iter = MaxIter;
    o.gamma = Gamma;
    o.nu_pre = Nc_pre;
    o.nu_post = Nc_post;
    const int rc = fm_mg_solve(h_, &o, &report);
    if (rc != FM_OK && rc != FM_ENOCONV) check(rc, "fm_mg_solve");
    return report.rel_resid;
  }

  // MGSolBase::MGTimeStep (src/MGSolverBase.C:426-454): assemble the top level with its rhs and the
  // coarse levels, solve, x_old <- x
  virtual void MGTimeStep(double time, int mode = 1) {
    (void)mode;
    const int top = _NoLevels - 1;
    check(fm_system_set_xold(h_, x_old.data()), "fm_system_set_xold");
    GenMatRhs(time, top, 1);
    for (int L = 0; L < top; L++) GenMatRhs(time, L, 0);
    MGSolve(opts.tol, opts.max_iter, opts.gamma, opts.nu_pre, 40, opts.nu_post);
    x_old = get_x();
  }

  // x_old of the top level (advection field and Dirichlet data of the next assembly)
  void set_xold(const std::vector<double> &v) {
    if (!sized(v, (size_t)_Dim.back(), "set_xold")) return;
    x_old = v;
    check(fm_system_set_xold(h_, x_old.data()), "fm_system_set_xold");
  }
  std::vector<double> get_x(int L = -1) const {
    if (L < 0) L = _NoLevels - 1;
    std::vector<double> v(_Dim[L]);
    check(fm_system_get_x(h_, L, v.data()), "fm_system_get_x");
    return v;
  }
  void set_x(const std::vector<double> &v, int L = -1) {
    if (L < 0) L = _NoLevels - 1;
    if (L >= _NoLevels || !sized(v, (size_t)_Dim[L], "set_x")) return;
    check(fm_system_set_x(h_, L, v.data()), "fm_system_set_x");
  }
  std::vector<double> get_rhs(int L = -1) const {
    if (L < 0) L = _NoLevels - 1;
    std::vector<double> v(_Dim[L]);
    check(fm_system_get_rhs(h_, L, v.data()), "fm_system_get_rhs");
    return v;
  }
  void set_rhs(const std::vector<double> &v, int L = -1) {
    if (L < 0) L = _NoLevels - 1;
    if (L >= _NoLevels || !sized(v, (size_t)_Dim[L], "set_rhs")) return;
    check(fm_system_set_rhs(h_, L, v.data()), "fm_system_set_rhs");
  }
  // a caller-assembled A[L]: values in the CSR order of the level's pattern (fm_system_set_matrix)
  void set_matrix(int L, const std::vector<double> &val) {
    if (L < 0 || L >= _NoLevels || !sized(val, (size_t)fm_system_csr_nnz(h_, 0, L), "set_matrix")) return;
    check(fm_system_set_matrix(h_, L, val.data()), "fm_system_set_matrix");
  }
  void set_elem_coef(int L, const std::vector<double> &c) {
    if (!sized(c, (size_t)n_elem(L), "set_elem_coef")) return;
    check(fm_system_set_elem_coef(h_, L, c.data()), "fm_system_set_elem_coef");
  }
  // nodal values of variable ivar on the top-level nodes from a top-level DoF vector (0 where the variable is absent)
  std::vector<double> node_values(const std::vector<double> &x, int ivar = 0) const {
    const int N = mesh_.n_nodes();
    std::vector<double> v(N, 0.0);
    if (!sized(x, (size_t)_Dim.back(), "node_values")) return v;
    const std::vector<int> &nd = _node_dof[_NoLevels - 1];
    for (int k = 0; k < N; k++) {
      const int d = nd[k + (size_t)ivar * N];
      if (d >= 0) v[k] = x[d];
    }
    return v;
  }
  // Dirichlet data of the penalty RHS separate from x_old (top level; empty: the reference's x_old)
  void set_bcdata(const std::vector<double> &g) {
    if (!g.empty() && !sized(g, (size_t)_Dim.back(), "set_bcdata")) return;
    check(fm_system_set_bcdata(h_, g.empty() ? nullptr : g.data()), "fm_system_set_bcdata");
  }
  // transposed multigrid solve A^T lambda = g on the transposed hierarchy (GenMatTranspose on every level first)
  std::vector<double> AdjointSolve(const std::vector<double> &g) {
    std::vector<double> lam(_Dim.back(), 0.0);
    if (!sized(g, lam.size(), "AdjointSolve")) return lam;
    const int rc = fm_adjoint_solve(h_, &opts, g.data(), lam.data(), &report);
    if (rc != FM_OK && rc != FM_ENOCONV) check(rc, "fm_adjoint_solve");
    return lam;
  }
  // coupling field from src's top-level solution gathered on the device (fm_system_couple: which 0 = src's
  // temperature as this system's buoyancy field, 1 = src's velocity as its advection field)
  void couple(const MGSolDA &src, int which, const double coef[3] = nullptr) {
    const double zero[3] = {0.0, 0.0, 0.0};
    check(fm_system_couple(h_, src.h_, which, coef ? coef : zero), "fm_system_couple");
  }
  // x_old <- x on the device (fm_system_advance); returns |x - x_old| / |x|.  The host copy x_old of this
  // class is left as it is: call get_x() when it is needed.
  double advance() {
    double r = 0.0;
    check(fm_system_advance(h_, &r), "fm_system_advance");
    return r;
  }
  // memory of the largest meshes: precisions in which the Vanka inverses are kept (bit 0 FP64, bit 1 FP32;
  // returns the previous mask), and release of one hierarchy's patch and coarse inverses
  int set_vanka_store(int mask) {
    const int rc = fm_system_set_vanka_store(h_, mask);
    if (rc < 0) check(rc, "fm_system_set_vanka_store");
    return rc;
  }
  void release_mg(bool adjoint = false) { check(fm_mg_release(h_, adjoint ? 1 : 0), "fm_mg_release"); }
  fm_system *handle() const { return h_; }

  // public state of the reference (MGSolverBase.h:44-98)
  int _NoLevels = 0;
  std::vector<int64_t> _Dim;
  std::vector<std::vector<int>> _node_dof;
  std::vector<int> bc[2];
  std::vector<double> x_old;
  fm_mg_opts opts;
  fm_mg_report report{};

 protected:
  Mesh &mesh_;
  int nq2_, nq1_;
  fm_system *h_ = nullptr;
  int n_elem(int L) const {
    int ne = 0, n2 = 0, n1 = 0;
    check(fm_mesh_level(mesh_.get(), L, &ne, &n2, &n1), "fm_mesh_level");
    return ne;
  }
  int n_elem_top() const { return n_elem(_NoLevels - 1); }

 private:
  static void bc_trampoline(const double *xp, int *neu, int *flag, void *self) {
    static_cast<MGSolDA *>(self)->bc_read(xp, neu, flag);
  }
};

// MGSolT: Q2 heat / temperature equation (template_appl/nse/SRC/MGSolverT.C)
class MGSolT : public MGSolDA {
 public:
  MGSolT(Context &ctx, Mesh &mesh, const fm_heat_params &p) : MGSolDA(ctx, mesh, 1, 0), par(p) {}
  void GenMatRhs(double time, int Level, int mode) override {
    (void)time;
    check(fm_heat_genmatrhs(h_, Level, mode, &par), "fm_heat_genmatrhs");
  }
  // transpose-Jacobian assembly of level L (the discrete adjoint's operator)
  void GenMatTranspose(int Level) { check(fm_heat_genmat_transpose(h_, Level, &par), "fm_heat_genmat_transpose"); }
  // advecting velocity [dim * Ntop] (component-major nodal values); empty removes it
  void set_velocity(const std::vector<double> &vel) {
    if (!vel.empty() && !sized(vel, (size_t)mesh_.dim() * mesh_.n_nodes(), "set_velocity")) return;
    check(fm_system_set_velocity(h_, vel.empty() ? nullptr : vel.data()), "fm_system_set_velocity");
  }
  // Nusselt number along axis of the top-level temperature T, with dNu/dT and the partials dNu/dk_e
  double nusselt(const std::vector<double> &T, int axis, std::vector<double> *dJdT = nullptr,
                 std::vector<double> *dJdk = nullptr) {
    double J = 0.0;
    if (!sized(T, (size_t)_Dim.back(), "nusselt")) return J;
    if (dJdT) dJdT->assign(_Dim.back(), 0.0);
    if (dJdk) dJdk->assign(n_elem_top(), 0.0);
    check(fm_heat_nusselt(h_, &par, axis, T.data(), &J, dJdT ? dJdT->data() : nullptr, dJdk ? dJdk->data() : nullptr),
          "fm_heat_nusselt");
    return J;
  }
  // -lambda^T (dA/dk_e) T for every top-level element
  std::vector<double> grad_coef(const std::vector<double> &lam, const std::vector<double> &T) {
    std::vector<double> g(n_elem_top(), 0.0);
    if (!sized(lam, (size_t)_Dim.back(), "grad_coef") || !sized(T, (size_t)_Dim.back(), "grad_coef")) return g;
    check(fm_heat_grad_coef(h_, &par, lam.data(), T.data(), g.data()), "fm_heat_grad_coef");
    return g;
  }
  fm_heat_params par;
};

// MGSolNS: coupled Q2-Q1 Navier-Stokes (template_appl/nse/SRC/MGSolverNS.C, NS_EQUATIONS == 1)
class MGSolNS : public MGSolDA {
 public:
  MGSolNS(Context &ctx, Mesh &mesh, const fm_ns_params &p) : MGSolDA(ctx, mesh, mesh.dim(), 1), par(p) {
    opts.smoother = FM_SMOOTH_VANKA;
    opts.vanka_colored = 1;
    opts.omega = 0.5;
  }
  void GenMatRhs(double time, int Level, int mode) override {
    (void)time;
    check(fm_ns_genmatrhs(h_, Level, mode, &par), "fm_ns_genmatrhs");
  }
  // transpose of the Oseen (Picard) matrix of level L
  void GenMatTranspose(int Level) { check(fm_ns_genmat_transpose(h_, Level, &par), "fm_ns_genmat_transpose"); }
  // Newton Jacobian of the steady residual at x_old (trans = 1: its transpose, for the adjoint)
  void GenJac(int Level, int trans) { check(fm_ns_genjac(h_, Level, trans, &par), "fm_ns_genjac"); }
  // Boussinesq body force coef * T from a nodal temperature field [Ntop]; empty removes it
  void set_buoyancy(const std::vector<double> &T_node, const double coef[3]) {
    if (!T_node.empty() && !sized(T_node, (size_t)mesh_.n_nodes(), "set_buoyancy")) return;
    check(fm_ns_set_buoyancy(h_, T_node.empty() ? nullptr : T_node.data(), coef), "fm_ns_set_buoyancy");
  }
  // nodal velocity [dim * Ntop] (component-major) of a top-level DoF vector
  std::vector<double> velocity_nodes(const std::vector<double> &U) const {
    const int N = mesh_.n_nodes(), dim = mesh_.dim();
    std::vector<double> v((size_t)dim * N);
    for (int d = 0; d < dim; d++) {
      const std::vector<double> c = node_values(U, d);
      std::copy(c.begin(), c.end(), v.begin() + (size_t)d * N);
    }
    return v;
  }
  fm_ns_params par;
};

}  // namespace femus_hip
#endif
