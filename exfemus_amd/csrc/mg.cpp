// mg.cpp -- multigrid for the MI355X FEMuS hot path.
//
// MGSolBase::MGStep / MGSolve (src/MGSolverBase.C:163-289) on the device:
// recursive V-cycle (Gamma = 1) with nu_pre / nu_post sweeps doubled below
// the top level, restriction b[L-1] = Rst res[L], x[L-1] = 0, coarse direct
// solve, prolongation x[L] += Prl x[L-1] (reference rule: 0.95 and only when
// the coarse residual is below the fine one), post-smoothing.
// Precision modes: all-FP64 cycles; FP32 V-cycle (FP32 values, FP32
// accumulation) inside FP64 iterative refinement or FP64 flexible GMRES,
// where the outer residual b - A x is always formed in FP64.
// The coarse-level inverse and the Vanka local inverses are built in FP64.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include "fm_internal.h"
#include "fm_kernels.h"

int fm_vanka_setup(fm_system *s, int L, std::vector<struct LOps> *ops);
int fm_vanka_sweep64(fm_system *s, int L, const struct LOps &o, double *x, const double *b, int its, double omega,
                     int colored, bool xz, bool post);
int fm_vanka_sweep32(fm_system *s, int L, const struct LOps &o, float *x, const float *b, int its, double omega,
                     int colored, bool xz, bool post);

// Per-level operator set used by the V-cycle (forward or transposed).
struct LOps {
  int64_t n = 0;
  const DevCSR *A = nullptr, *R = nullptr, *P = nullptr;
  double *dinv = nullptr;
  float *dinvf = nullptr;
  double lam = 0.;
  double *Ainv = nullptr;
  float *Ainvf = nullptr;
  Level *v = nullptr;  // work vectors
  double *vinv = nullptr;  // Vanka local inverses (transposed), FP64 and FP32 copies
  float *vinvf = nullptr;
};

template <typename T> static inline CsrView<T> cview(const DevCSR &c) { return fm_view<T>(c); }

// typed access to a level's work vectors
template <typename T> struct Vec;
template <> struct Vec<double> {
  static double *x(Level *l) { return l->x; }  static double *b(Level *l) { return l->b; }
  static double *r(Level *l) { return l->r; }  static double *t(Level *l) { return l->t; }
  static double *d(Level *l) { return l->d; }  static double *d2(Level *l) { return l->d2; }
  static double *dinv(const LOps &o) { return o.dinv; }
  static double *Ainv(const LOps &o) { return o.Ainv; }
};
template <> struct Vec<float> {
  static float *x(Level *l) { return l->xf; }  static float *b(Level *l) { return l->bf; }
  static float *r(Level *l) { return l->rf; }  static float *t(Level *l) { return l->tf; }
  static float *d(Level *l) { return l->df; }  static float *d2(Level *l) { return l->d2f; }
  static float *dinv(const LOps &o) { return o.dinvf; }
  static float *Ainv(const LOps &o) { return o.Ainvf; }
};

// ---- device reductions read back on the host -----------------------------
template <typename T> static double dev_dot(fm_ctx *c, int64_t n, const T *x, const T *y, bool do_sqrt) {
  if (fmk::dot_dev<T>(c->stream, n, x, y, c->scratch, c->scratch + 2048, do_sqrt ? 1 : 0) != 0) return NAN;
  if (hipMemcpyAsync(c->h_scratch, c->scratch + 2048, sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess) return NAN;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return NAN;
  return c->h_scratch[0];
}

// ---- lambda_max(D^-1 A) by power iteration (same start vector as the oracle,
// 1 + (dof mod 7) / 7 in the reference numbering: ord[i] is the reference DoF of internal row i)
__global__ void k_start_vec(int64_t n, const int *__restrict__ ord, double *x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = 1.0 + (double)(ord[i] % 7) / 7.0;
}
// y = dinv .* x
__global__ void k_scale_mul(int64_t n, double *y, const double *__restrict__ dinv, const double *__restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = dinv[i] * x[i];
}
static int lambda_max(fm_ctx *c, const LOps &o, int iters, double *lam) {
  Level *l = o.v;
  hipStream_t st = c->stream;
  int64_t n = o.n;
  unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_start_vec, dim3(g), dim3(256), 0, st, n, l->ord, l->t);
  double lm = 0.;
  for (int it = 0; it < iters; it++) {
    double xn = dev_dot<double>(c, n, l->t, l->t, true);
    if (!(xn > 0.)) return FM_ESINGULAR;
    // t = t / ||t||  (scale folded into the product below)
    FM_CHECK(fmk::axpby<double>(st, n, 1.0 / xn, l->t, 0.0, l->t));
    FM_CHECK(fmk::spmv(st, cview<double>(*o.A), l->t, l->d));
    hipLaunchKernelGGL(k_scale_mul, dim3(g), dim3(256), 0, st, n, l->t, o.dinv, l->d);
    lm = dev_dot<double>(c, n, l->t, l->t, true);
  }
  *lam = lm;
  return 0;
}

// ---- coarse inverse (coarse.hip): blocked exchange inversion on the device --------
int fm_dense_inverse(hipStream_t st, const DevCSR &A, const int *d_rank, int pin, double thr, double **dAinv,
                     float **dAinvf);

// The dense inverse is laid out in the level's internal DoF order (rank[dof]).  For a system with a Q1
// (pressure) variable the last reference DoF may be pinned: in an enclosed flow the pressure is defined up
// to a constant and the coarse solve then sets that DoF to 0.
static int coarse_inverse(fm_system *s, const DevCSR &A, const Level &lv, double **dAinv, float **dAinvf) {
  const int n = (int)A.nr;
  hipStream_t st = s->ctx->stream;
  if ((int)lv.h_rank.size() != n) { fm_set_error("coarse level without internal order"); return FM_ESTATE; }
  FM_CHECK(fm_csr_from_sell(st, const_cast<DevCSR &>(A)));
  std::vector<double> vals(A.nnz);
  FM_HIP(hipMemcpyAsync(vals.data(), A.v, A.nnz * sizeof(double), hipMemcpyDeviceToHost, st));
  FM_HIP(hipStreamSynchronize(st));
  double amax = 0.;
  for (double v : vals) amax = std::max(amax, std::fabs(v));
  const int pin = s->nq1 > 0 ? lv.h_rank[n - 1] : -1;
  return fm_dense_inverse(st, A, lv.rank, pin, 1e-13 * amax, dAinv, dAinvf);
}

// ---- forward operator set
static std::vector<LOps> forward_ops(fm_system *s) {
  std::vector<LOps> ops(s->lev.size());
  for (size_t L = 0; L < s->lev.size(); L++) {
    Level &l = s->lev[L];
    LOps &o = ops[L];
    o.n = l.n; o.A = &l.A; o.R = &l.R; o.P = &l.P;
    o.dinv = l.dinv; o.dinvf = l.dinvf; o.lam = l.lam; o.Ainv = l.Ainv; o.Ainvf = l.Ainvf; o.v = &l; o.vinv = l.vinv;
    o.vinvf = l.vinvf;
  }
  return ops;
}

// ---- smoothers ----------------------------------------------------------
// Chebyshev sweeps: the first direction d_0 = a D^-1 r_0 and x += d_0 come either from one elementwise
// kernel (xz: x is zero on entry, r_0 = b, x = d_0; x need not be cleared beforehand) or from the residual
// product's epilogue, with x += d_0 folded into the next step (x += d_0 + d_1).
template <typename T>
static int cheb_sweeps(hipStream_t st, const CsrView<T> &A, const LOps &o, T *x, const T *b, T *r, T *d, T *d2,
                       const T *dinv, int its, bool xz, double a0, const std::vector<std::pair<double, double>> &c) {
  bool pending = false;  // d_0 not yet added to x
  if (xz) {
    FM_CHECK(fmk::cheb_start<T>(st, o.n, dinv, b, r, d, x, (T)a0));
  } else if (its >= 2) {
    FM_CHECK(fmk::resid_d(st, A, b, x, r, d, dinv, (T)a0));
    pending = true;
  } else {
    FM_CHECK(fmk::resid(st, A, b, x, r));
    FM_CHECK(fmk::cheb_first<T>(st, o.n, dinv, r, d, x, (T)a0));
  }
  for (int k = 1; k < its; k++) {
    FM_CHECK(fmk::cheb_step(st, A, r, x, d, d2, dinv, (T)c[k - 1].first, (T)c[k - 1].second, pending ? 1 : 0));
    pending = false;
    std::swap(d, d2);
  }
  return 0;
}

// xz: x is zero on entry (for the Chebyshev smoothers its memory need not hold zeros: the first sweep writes x)
template <typename T>
static bool smoother_writes_x(const fm_mg_opts *opt, int its) {
  return its > 0 && (opt->smoother == FM_SMOOTH_CHEBYSHEV || opt->smoother == FM_SMOOTH_CHEBYSHEV4);
}

template <typename T>
static int smooth(fm_system *s, const fm_mg_opts *opt, int L, const LOps &o, int its, bool xz, bool post) {
  hipStream_t st = s->ctx->stream;
  Level *l = o.v;
  T *x = Vec<T>::x(l), *b = Vec<T>::b(l), *r = Vec<T>::r(l), *d = Vec<T>::d(l), *d2 = Vec<T>::d2(l);
  T *dinv = Vec<T>::dinv(o);
  CsrView<T> A = cview<T>(*o.A);
  if (its <= 0) return 0;
  if (opt->smoother == FM_SMOOTH_CHEBYSHEV) {
    double lo = opt->cheb_lo * o.lam, hi = opt->cheb_hi * o.lam;
    double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta, rho = 1. / sigma;
    std::vector<std::pair<double, double>> c;
    for (int k = 2; k <= its; k++) {
      double rho_new = 1. / (2. * sigma - rho);
      c.emplace_back(rho_new * rho, 2. * rho_new / delta);
      rho = rho_new;
    }
    return cheb_sweeps<T>(st, A, o, x, b, r, d, d2, dinv, its, xz, 1.0 / theta, c);
  }
  if (opt->smoother == FM_SMOOTH_CHEBYSHEV4) {
    // d_0 = 4/(3 rho) D^-1 r_0; r_k = r_{k-1} - A d_{k-1};
    // d_k = (2k-1)/(2k+3) d_{k-1} + (8k+4)/((2k+3) rho) D^-1 r_k; x += d_k   (rho = cheb_hi * lambda_max)
    const double rho = opt->cheb_hi * o.lam;
    std::vector<std::pair<double, double>> c;
    for (int k = 1; k < its; k++)
      c.emplace_back((2.0 * k - 1.0) / (2.0 * k + 3.0), (8.0 * k + 4.0) / ((2.0 * k + 3.0) * rho));
    return cheb_sweeps<T>(st, A, o, x, b, r, d, d2, dinv, its, xz, 4.0 / (3.0 * rho), c);
  }
  if (opt->smoother == FM_SMOOTH_JACOBI) {
    // ping-pong through t: x_{k+1} = x_k + omega D^-1 (b - A x_k)
    T *t = Vec<T>::t(l);
    for (int k = 0; k < its; k++) {
      FM_CHECK(fmk::jacobi_step(st, A, b, x, t, dinv, (T)opt->omega));
      FM_CHECK(fmk::copy<T>(st, o.n, t, x));
    }
    return 0;
  }
  if (opt->smoother == FM_SMOOTH_VANKA) {
    if (sizeof(T) == 8)
      return fm_vanka_sweep64(s, L, o, (double *)x, (const double *)b, its, opt->omega, opt->vanka_colored, xz, post);
    return fm_vanka_sweep32(s, L, o, (float *)x, (const float *)b, its, opt->omega, opt->vanka_colored, xz, post);
  }
  return fm_einval(__func__);
}

// ---- V-cycle (MGSolBase::MGStep) -----------------------------------------
// xz: x[L] is zero on entry (every coarse level; the top level of a correction cycle)
template <typename T>
static int vcycle(fm_system *s, const fm_mg_opts *opt, std::vector<LOps> &ops, int L, bool xz) {
  hipStream_t st = s->ctx->stream;
  const LOps &o = ops[L];
  Level *l = o.v;
  if (L == 0) {
    FM_CHECK(fmk::dense_mv<T>(st, (int)o.n, Vec<T>::Ainv(o), Vec<T>::b(l), Vec<T>::x(l)));
    return 0;
  }
  int top = (int)ops.size() - 1;
  const int mult = L < top ? (opt->coarse_mult > 0 ? opt->coarse_mult : 2) : 1;
  FM_CHECK(smooth<T>(s, opt, L, o, opt->nu_pre * mult, xz, false));
  CsrView<T> A = cview<T>(*o.A);
  T *x = Vec<T>::x(l), *b = Vec<T>::b(l), *r = Vec<T>::r(l);
  FM_CHECK(fmk::resid(st, A, b, x, r));
  const LOps &oc = ops[L - 1];
  Level *lc = oc.v;
  double *cond = nullptr;
  if (opt->correction == 0) {
    // rest = ||r||/||b|| on this level; coarse_rest measured after the coarse cycle
    FM_CHECK(fmk::dot_dev<T>(st, o.n, r, r, s->ctx->scratch, l->nrm + 0, 1));
    FM_CHECK(fmk::dot_dev<T>(st, o.n, b, b, s->ctx->scratch, l->nrm + 1, 1));
    FM_CHECK(fmk::ratio(st, l->nrm + 0, l->nrm + 1, l->nrm + 3));   // fine rest
  }
  FM_CHECK(fmk::spmv(st, cview<T>(*oc.R), r, Vec<T>::b(lc)));
  // the coarse iterate starts at zero: cleared here unless the coarse solve (level 0) or the coarse level's
  // first Chebyshev sweep writes it whole
  const int mult_c = opt->coarse_mult > 0 ? opt->coarse_mult : 2;
  if (L - 1 > 0 && !smoother_writes_x<T>(opt, opt->nu_pre * mult_c)) FM_CHECK(fmk::fill<T>(st, oc.n, Vec<T>::x(lc), (T)0));
  const int gamma = (L - 1 > 0 && opt->gamma > 1) ? opt->gamma : 1;
  for (int g = 0; g < gamma; g++) FM_CHECK(vcycle<T>(s, opt, ops, L - 1, g == 0));
  double alpha = 1.0;
  if (opt->correction == 0) {
    // coarse_rest: relative residual of the coarse level after its cycle
    CsrView<T> Ac = cview<T>(*oc.A);
    FM_CHECK(fmk::resid(st, Ac, Vec<T>::b(lc), Vec<T>::x(lc), Vec<T>::r(lc)));
    FM_CHECK(fmk::dot_dev<T>(st, oc.n, Vec<T>::r(lc), Vec<T>::r(lc), s->ctx->scratch, l->nrm + 4, 1));
    FM_CHECK(fmk::dot_dev<T>(st, oc.n, Vec<T>::b(lc), Vec<T>::b(lc), s->ctx->scratch, l->nrm + 5, 1));
    FM_CHECK(fmk::ratio(st, l->nrm + 4, l->nrm + 5, l->nrm + 2));   // nrm[2] = coarse rest, nrm[3] = fine rest
    cond = l->nrm + 2;
    alpha = 0.95;
  }
  FM_CHECK(fmk::addmul(st, cview<T>(*o.P), Vec<T>::x(lc), x, (T)alpha, cond));
  FM_CHECK(smooth<T>(s, opt, L, o, opt->nu_post * mult, false, true));
  return 0;
}

// ---- top-level cycle, replayed from a captured HIP graph -------------------
template <typename T>
static int cycle_top(fm_system *s, const fm_mg_opts *o, std::vector<LOps> &ops, int adj, bool xz) {
  const int top = (int)ops.size() - 1;
  if (!o->use_graph) return vcycle<T>(s, o, ops, top, xz);
  const int prec = sizeof(T) == 8 ? 64 : 32;
  const int split = fmk::sell_split(-1);
  hipStream_t st = s->ctx->stream;
  for (auto &g : s->graphs)
    if (g.prec == prec && g.adj == adj && g.xz == (int)xz && g.split == split && g.smoother == o->smoother && g.nu_pre == o->nu_pre && g.nu_post == o->nu_post &&
        g.correction == o->correction && g.gamma == o->gamma && g.cmult == o->coarse_mult && g.colored == o->vanka_colored && g.lo == o->cheb_lo &&
        g.hi == o->cheb_hi && g.omega == o->omega)
      return hipGraphLaunch(g.exec, st) == hipSuccess ? 0 : FM_EHIP;
  hipGraph_t graph = nullptr;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return FM_EHIP;
  int rc = vcycle<T>(s, o, ops, top, xz);
  hipError_t e = hipStreamEndCapture(st, &graph);
  if (rc || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    fm_set_error("V-cycle capture failed");
    return rc ? rc : FM_EHIP;
  }
  GraphEntry g;
  g.prec = prec; g.adj = adj; g.xz = (int)xz; g.split = split; g.smoother = o->smoother; g.nu_pre = o->nu_pre; g.nu_post = o->nu_post;
  g.correction = o->correction; g.gamma = o->gamma; g.cmult = o->coarse_mult; g.colored = o->vanka_colored; g.lo = o->cheb_lo; g.hi = o->cheb_hi;
  g.omega = o->omega;
  e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) { fm_set_error("V-cycle graph instantiate failed"); return FM_EHIP; }
  s->graphs.push_back(g);
  return hipGraphLaunch(g.exec, st) == hipSuccess ? 0 : FM_EHIP;
}

// ---- setup: eigen bounds, coarse inverse, Vanka --------------------------
static int mg_setup_ops(fm_system *s, const fm_mg_opts *o, std::vector<LOps> &ops, bool adjoint);

extern "C" int fm_mg_setup(fm_system *s, const fm_mg_opts *o) {
  if (!s || !o || !s->initialized) return FM_ESTATE;
  FM_HIP(hipSetDevice(s->ctx->device));
  fm_drop_graphs(s);
  std::vector<LOps> ops = forward_ops(s);
  FM_CHECK(mg_setup_ops(s, o, ops, false));
  for (size_t L = 0; L < s->lev.size(); L++) {
    s->lev[L].lam = ops[L].lam;
    s->lev[L].Ainv = ops[L].Ainv; s->lev[L].Ainvf = ops[L].Ainvf; s->lev[L].vinv = ops[L].vinv;
    s->lev[L].vinvf = ops[L].vinvf;
  }
  s->mg_ready = true;
  s->mg_smoother = o->smoother;
  return FM_OK;
}

extern "C" int fm_system_set_vanka_store(fm_system *s, int mask) {
  if (!s || mask < 0 || mask > 3) return fm_einval(__func__);
  const int old = s->vanka_store;
  if (mask > 0) s->vanka_store = mask;
  return old;
}

extern "C" int fm_mg_release(fm_system *s, int adjoint) {
  if (!s) return fm_einval(__func__);
  FM_HIP(hipSetDevice(s->ctx->device));
  FM_HIP(hipStreamSynchronize(s->ctx->stream));
  fm_drop_graphs(s);  // captured cycles hold the addresses released here
  fm_krylov_free(s);  // the flexible-GMRES workspace is shared by both hierarchies
  for (Level &l : adjoint ? s->alev : s->lev) {
    for (double **p : {&l.vinv, &l.Ainv}) if (*p) { (void)hipFree(*p); *p = nullptr; }
    for (float **p : {&l.vinvf, &l.Ainvf}) if (*p) { (void)hipFree(*p); *p = nullptr; }
  }
  if (adjoint) s->adj_ready = false;
  else s->mg_ready = false;
  return FM_OK;
}

static int mg_setup_ops(fm_system *s, const fm_mg_opts *o, std::vector<LOps> &ops, bool adjoint) {
  (void)adjoint;
  for (size_t L = 1; L < ops.size(); L++) {
    if ((o->smoother == FM_SMOOTH_CHEBYSHEV || o->smoother == FM_SMOOTH_CHEBYSHEV4) && ops[L].lam == 0.)
      FM_CHECK(lambda_max(s->ctx, ops[L], 20, &ops[L].lam));
    if (o->smoother == FM_SMOOTH_VANKA) FM_CHECK(fm_vanka_setup(s, (int)L, &ops));
  }
  FM_CHECK(coarse_inverse(s, *ops[0].A, *ops[0].v, &ops[0].Ainv, &ops[0].Ainvf));
  return 0;
}

// ---- outer solvers ---------------------------------------------------------
struct Timer {
  fm_ctx *c;
  explicit Timer(fm_ctx *cc) : c(cc) { (void)hipEventRecord(c->ev0, c->stream); }
  double stop() {
    (void)hipEventRecord(c->ev1, c->stream);
    (void)hipEventSynchronize(c->ev1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    return ms * 1e-3;
  }
};

// x[top] (FP64) := solution; b[top] FP64 rhs.  Stationary cycles / IR.
static int solve_stationary(fm_system *s, const fm_mg_opts *o, std::vector<LOps> &ops, fm_mg_report *rep, int adj) {
  fm_ctx *c = s->ctx;
  hipStream_t st = c->stream;
  int top = (int)ops.size() - 1;
  LOps &ot = ops[top];
  Level *l = ot.v;
  double bn = dev_dot<double>(c, ot.n, l->b, l->b, true);
  if (!(bn > 0.)) { rep->iters = 0; rep->converged = 1; rep->rel_resid = 0.; return 0; }
  CsrView<double> A = cview<double>(*ot.A);
  int it = 0;
  double rel = 1.;
  FM_CHECK(fmk::resid(st, A, l->b, l->x, l->res));
  rel = dev_dot<double>(c, ot.n, l->res, l->res, true) / bn;
  while (it < o->max_iter && rel > o->tol) {
    if (o->precision == FM_PREC_FP64) {
      FM_CHECK(cycle_top<double>(s, o, ops, adj, false));
    } else {
      // FP32 correction: A e = r (FP32 V-cycle from e = 0), x += e
      FM_CHECK(fmk::to_f32(st, ot.n, l->res, l->bf, nullptr));
      FM_CHECK(fmk::fill<float>(st, ot.n, l->xf, 0.f));
      FM_CHECK(cycle_top<float>(s, o, ops, adj, true));
      FM_CHECK(fmk::add_f32(st, ot.n, l->x, l->xf, nullptr));
    }
    FM_CHECK(fmk::resid(st, A, l->b, l->x, l->res));
    rel = dev_dot<double>(c, ot.n, l->res, l->res, true) / bn;
    if (it < 256) rep->hist[it] = rel;
    it++;
    if (!std::isfinite(rel)) break;
  }
  rep->iters = it;
  rep->rel_resid = rel;
  rep->converged = rel <= o->tol;
  return 0;
}

// Flexible GMRES(m), right preconditioned by one V-cycle (FP32 or FP64).
// part[j * gridDim.x + block] = sum_i V_j[i] w[i] for j < kk, where V_j is row j of V for j < k and w itself
// for j = k (kk = k + 1: the squared norm of w rides along).  Block row blockIdx.y takes the 8 dot products
// j0 .. j0 + 7 (j0 = 8 blockIdx.y) in one sweep, reading w once for the eight.  With skip non-null the
// launch does nothing when skip[0] == 0.
__global__ void k_multi_dot(int64_t n, int k, int kk, const double *__restrict__ V, const double *__restrict__ w,
                            double *part, const double *skip) {
  if (skip && skip[0] == 0.) return;
  __shared__ double sh[4][8];
  const int j0 = 8 * blockIdx.y;
  const int nj = min(8, kk - j0);
  const double *vp[8];
#pragma unroll
  for (int t = 0; t < 8; t++) vp[t] = (j0 + t < k) ? V + (int64_t)(j0 + t) * n : w;
  double s[8];
#pragma unroll
  for (int t = 0; t < 8; t++) s[t] = 0.;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double wi = w[i];
#pragma unroll
    for (int t = 0; t < 8; t++)
      if (t < nj) s[t] += vp[t][i] * wi;
  }
#pragma unroll
  for (int t = 0; t < 8; t++)
    for (int off = 32; off > 0; off >>= 1) s[t] += __shfl_xor(s[t], off, 64);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int t = 0; t < 8; t++) sh[threadIdx.x >> 6][t] = s[t];
  }
  __syncthreads();
  if ((int)threadIdx.x < nj) {
    double t = 0.;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += sh[q][threadIdx.x];
    part[(int64_t)(j0 + threadIdx.x) * gridDim.x + blockIdx.x] = t;
  }
}
// w = (w - sum_{j < k} h[j] V[j]) scale[0], and its FP32 copy in f when f is non-null (nothing when skip is non-null and
// skip[0] == 0)
__global__ void k_multi_axpy(int64_t n, int k, const double *__restrict__ V, const double *__restrict__ h, double *w,
                             const double *scale, float *f, const double *skip) {
  if (skip && skip[0] == 0.) return;
  const double sc = scale[0];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double s = w[i];
    for (int j = 0; j < k; j++) s -= h[j] * V[(int64_t)j * n + i];
    s *= sc;
    w[i] = s;
    if (f) f[i] = (float)s;
  }
}
// x += sum_j y[j] Z[j]
__global__ void k_multi_add(int64_t n, int k, const double *__restrict__ Z, const double *__restrict__ y, double *x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double s = x[i];
    for (int j = 0; j < k; j++) s += y[j] * Z[(int64_t)j * n + i];
    x[i] = s;
  }
}

// w /= nrm[0] (device scalar; left unchanged when nrm[0] == 0)
__global__ void k_scale_inv(int64_t n, double *w, const double *nrm) {
  const double d = nrm[0];
  if (!(d > 0.)) return;
  const double r = 1.0 / d;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) w[i] *= r;
}

// Hessenberg state of one restart cycle, on the device (offsets into fm_system::kH, m = restart length):
// H [(m + 1) m] row-major, cs [m], sn [m], g [m + 1], y [m], rel [m] (residual estimate after each step),
// st [8] (st[0]: columns before the first breakdown, -1 while there is none), then the Gram-Schmidt scratch
// hd [256]: [0, 64) first pass, [64, 128) second pass, [128] norm of the new vector, [129] reorthogonalisation flag,
// [130] |r| at a restart, [131] scale of the Gram-Schmidt update
struct KryLayout {
  int64_t H, cs, sn, g, y, rel, st, hd, total;
  explicit KryLayout(int m) {
    H = 0; cs = H + (int64_t)(m + 1) * m; sn = cs + m; g = sn + m; y = g + m + 1; rel = y + m; st = rel + m;
    hd = st + 8; total = hd + 256;
  }
};

// start of a restart cycle: g = beta e_1 (beta = |r| from the device scalar), no breakdown yet
__global__ void k_fgmres_start(int m, double *g, double *st, const double *beta) {
  for (int i = threadIdx.x; i <= m; i += blockDim.x) g[i] = i == 0 ? beta[0] : 0.;
  if (threadIdx.x == 0) st[0] = -1.;
}
// Column j of the Hessenberg matrix from the Gram-Schmidt coefficients (two passes) and the new vector's
// norm, the earlier Givens rotations applied to it, the rotation that zeroes its subdiagonal, the rotated
// right-hand side g and the residual estimate rel[j] = |g[j + 1]| / bn.  One thread: at most 60 columns.
__device__ void arnoldi_col(int j, int m, const double *hd, double *H, double *cs, double *sn, double *g,
                            double *rel, double *st, double bn) {
  for (int k = 0; k <= j; k++) H[(int64_t)k * m + j] = hd[k] + hd[64 + k];
  const double hn = hd[128];
  H[(int64_t)(j + 1) * m + j] = hn;
  for (int k = 0; k < j; k++) {
    const double a = H[(int64_t)k * m + j], b = H[(int64_t)(k + 1) * m + j];
    H[(int64_t)k * m + j] = cs[k] * a + sn[k] * b;
    H[(int64_t)(k + 1) * m + j] = -sn[k] * a + cs[k] * b;
  }
  const double a = H[(int64_t)j * m + j], b = H[(int64_t)(j + 1) * m + j];
  const double r = sqrt(a * a + b * b);
  cs[j] = r > 0. ? a / r : 1.;
  sn[j] = r > 0. ? b / r : 0.;
  H[(int64_t)j * m + j] = r;
  H[(int64_t)(j + 1) * m + j] = 0.;
  g[j + 1] = -sn[j] * g[j];
  g[j] = cs[j] * g[j];
  rel[j] = fabs(g[j + 1]) / bn;
  // a zero new vector (the Krylov space is invariant: the solution is in it) ends the usable columns
  if (!(hn > 0.) && st[0] < 0.) st[0] = (double)(j + 1);
}
// One Gram-Schmidt pass of Arnoldi step j, after k_multi_dot has left the nb block partials of the j + 2 dot
// products (w with V_0 .. V_j, then |w|^2): the four waves sum them (coefficients to hd[64 pass + t]); pass 0 takes
// the selective reorthogonalisation test (Daniel, Gragg, Kaufman and Stewart, Math. Comp. 30 (1976), eta = 1/sqrt 2):
// a second pass (hd[129] = 1) only when |w'|^2 = |w|^2 - sum h^2 < |w|^2 / 2.  The pass that completes the step
// (pass 0 without a second pass, or pass 1) takes the new vector's norm from the same identity -- no cancellation:
// the test bounds it below by half of |w|^2 -- forms the Hessenberg column and its rotation (arnoldi_col) and sets
// the update's scale hd[131] = 1 / norm (1 while a second pass follows, or when the norm is zero).  Pass 1 does
// nothing when no second pass was asked for.
__global__ void k_gs_step(int j, int m, int nb, const double *part, double *hd, int pass, double *H, double *cs,
                          double *sn, double *g, double *rel, double *st, double bn) {
  if (pass == 1 && hd[129] == 0.) return;
  double *out = hd + 64 * pass;
  const int kk = j + 2, wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  for (int t = wv; t < kk; t += (int)(blockDim.x >> 6)) {
    double s = 0.;
    for (int i = ln; i < nb; i += 64) s += part[(int64_t)t * nb + i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (ln == 0) out[t] = s;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double proj = 0.;
  for (int t = 0; t <= j; t++) proj += out[t] * out[t];
  const double ww = out[j + 1];
  if (pass == 0) {
    const bool again = ww - proj < 0.5 * ww;
    hd[129] = again ? 1. : 0.;
    if (again) {
      hd[131] = 1.;
      return;
    }
    for (int t = 0; t <= j; t++) hd[64 + t] = 0.;
  }
  const double hn = ww - proj > 0. ? sqrt(ww - proj) : 0.;
  hd[128] = hn;
  hd[131] = hn > 0. ? 1. / hn : 1.;
  arnoldi_col(j, m, hd, H, cs, sn, g, rel, st, bn);
}

// y = H^-1 g over the first jn columns (fewer after a breakdown; the rest of y is zero)
__global__ void k_fgmres_y(int jn, int m, const double *H, const double *g, double *y, const double *st) {
  if (threadIdx.x != 0) return;
  const int j = st[0] > 0. ? min(jn, (int)st[0]) : jn;
  for (int i = jn - 1; i >= j; i--) y[i] = 0.;
  for (int i = j - 1; i >= 0; i--) {
    double acc = g[i];
    for (int k = i + 1; k < j; k++) acc -= H[(int64_t)i * m + k] * y[k];
    const double d = H[(int64_t)i * m + i];
    y[i] = d != 0. ? acc / d : 0.;
  }
}

void fm_krylov_free(fm_system *s) {
  for (double **p : {&s->kV, &s->kZ, &s->kpart, &s->kH})
    if (*p) { (void)hipFree(*p); *p = nullptr; }
  if (s->kpin) { (void)hipHostFree(s->kpin); s->kpin = nullptr; }
  for (hipEvent_t e : s->kev) (void)hipEventDestroy(e);
  s->kev.clear();
  s->kn = 0; s->km = 0; s->knb = 0;
}

// the workspace for restart length m and top-level size n (allocated once, reused while both stay)
static int krylov_ensure(fm_system *s, int64_t n, int m, int nb) {
  if (s->kV && s->kn == n && s->km == m && s->knb == nb) return 0;
  FM_HIP(hipStreamSynchronize(s->ctx->stream));
  fm_krylov_free(s);
  const KryLayout K(m);
  int rc = 0;
  if (hipMalloc((void **)&s->kV, (size_t)(m + 1) * n * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&s->kZ, (size_t)m * n * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&s->kpart, (size_t)(m + 1) * nb * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&s->kH, (size_t)K.total * sizeof(double)) != hipSuccess)
    rc = FM_ENOMEM;
  if (!rc && hipHostMalloc((void **)&s->kpin, (size_t)m * sizeof(double), hipHostMallocDefault) != hipSuccess) rc = FM_EHIP;
  for (int j = 0; j < m && !rc; j++) {
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) { rc = FM_EHIP; break; }
    s->kev.push_back(e);
  }
  if (rc) {
    fm_krylov_free(s);
    fm_set_error("flexible GMRES workspace: allocation failed");
    return rc;
  }
  s->kn = n; s->km = m; s->knb = nb;
  return 0;
}

// Flexible GMRES(m), right preconditioned by one V-cycle (FP32 or FP64).  The Arnoldi steps run without
// host round trips: the Gram-Schmidt coefficients, the Hessenberg column, its Givens rotation and the
// residual estimate are formed on the device, and each step's estimate is copied to pinned memory behind
// an event.  The host waits only at checkpoints: after a checkpoint with estimate r, the next
// ceil(log(tol / 2 r) / log rho) steps are launched without waiting, rho being the fastest reduction per
// step assumed possible: 1/20, or the fastest one seen if that was faster (GMRES accelerates, so the rate
// of the first steps is no bound).  A step launched after the one that converged only costs time: the
// least-squares update over more columns is at least as good.
static int solve_fgmres(fm_system *s, const fm_mg_opts *o, std::vector<LOps> &ops, fm_mg_report *rep, int adj) {
  fm_ctx *c = s->ctx;
  hipStream_t st = c->stream;
  int top = (int)ops.size() - 1;
  LOps &ot = ops[top];
  Level *l = ot.v;
  int64_t n = ot.n;
  const int m = std::max(1, std::min(o->fgmres_restart > 0 ? o->fgmres_restart : 30, 60));
  const int nb = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + 2047) / 2048));
  FM_CHECK(krylov_ensure(s, n, m, nb));
  double *V = s->kV, *Z = s->kZ, *part = s->kpart;
  const KryLayout K(m);
  double *H = s->kH + K.H, *cs = s->kH + K.cs, *sn = s->kH + K.sn, *g = s->kH + K.g, *y = s->kH + K.y;
  double *relh = s->kH + K.rel, *stt = s->kH + K.st, *hd = s->kH + K.hd;
  CsrView<double> A = cview<double>(*ot.A);
  int rc = 0;
  double bn = dev_dot<double>(c, n, l->b, l->b, true);
  int total = 0;
  double rel = 1.;
  double rho = 0.05;  // fastest reduction of the estimate per step assumed possible (lowered if one is faster)
  unsigned gv = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  bool res_fresh = false;  // V holds b - A x of the current x (the exit residual needs no product)
  bool x_zero = dev_dot<double>(c, n, l->x, l->x, true) == 0.;
  if (!(bn > 0.)) { rel = 0.; goto done; }
  while (total < o->max_iter) {
    if (x_zero) {   // r = b without a product
      x_zero = false;
      if (hipMemcpyAsync(V, l->b, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) { rc = FM_EHIP; goto done; }
    } else if ((rc = fmk::resid(st, A, l->b, l->x, V)) != 0) {
      goto done;
    }
    {
      // beta = |r| stays on the device for g; the host needs it for the restart test
      if ((rc = fmk::dot_dev<double>(st, n, V, V, c->scratch, hd + 130, 1)) != 0) goto done;
      if (hipMemcpyAsync(c->h_scratch, hd + 130, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess) { rc = FM_EHIP; goto done; }
      const double beta = c->h_scratch[0];
      rel = beta / bn;
      if (total == 0) rep->hist[0] = rel;
      if (rel <= o->tol || !std::isfinite(rel)) { res_fresh = true; break; }
      hipLaunchKernelGGL(k_scale_inv, dim3(gv), dim3(256), 0, st, n, V, hd + 130);
      hipLaunchKernelGGL(k_fgmres_start, dim3(1), dim3(64), 0, st, m, g, stt, hd + 130);
      // steps from an estimate r that cannot reach tol / 2 at the reduction rho per step, plus one (1 .. m)
      auto lookahead = [&](double r) -> int {
        const double need = std::ceil(std::log(0.5 * o->tol / r) / std::log(rho));
        return need >= 1. ? (int)std::min(need, (double)m) : 1;
      };
      int j = 0, jv = -1, check_at = lookahead(rel) - 1;
      double rv = rel;
      bool stop = false;
      while (j < m && total < o->max_iter && !stop) {
        double *vj = V + (size_t)j * n, *zj = Z + (size_t)j * n, *w = V + (size_t)(j + 1) * n;
        // z_j = M v_j
        if (o->precision == FM_PREC_FP64_FGMRES) {
          // the FP64 cycle works on the level's own b / x: park them in sb / sx
          if (hipMemcpyAsync(l->sb, l->b, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) { rc = FM_EHIP; goto done; }
          if (hipMemcpyAsync(l->b, vj, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) { rc = FM_EHIP; goto done; }
          if (hipMemcpyAsync(l->sx, l->x, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) { rc = FM_EHIP; goto done; }
          if (!smoother_writes_x<double>(o, o->nu_pre) && (rc = fmk::fill<double>(st, n, l->x, 0.0)) != 0) goto done;
          if ((rc = cycle_top<double>(s, o, ops, adj, true)) != 0) goto done;
          if (hipMemcpyAsync(zj, l->x, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) { rc = FM_EHIP; goto done; }
          if (hipMemcpyAsync(l->x, l->sx, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) { rc = FM_EHIP; goto done; }
          if (hipMemcpyAsync(l->b, l->sb, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) { rc = FM_EHIP; goto done; }
        } else {
          // the FP32 copy of v_j: v_0 here, v_j (j > 0) written by the update that formed it
          if (j == 0 && (rc = fmk::to_f32(st, n, vj, l->bf, nullptr)) != 0) goto done;
          if (!smoother_writes_x<float>(o, o->nu_pre) && (rc = fmk::fill<float>(st, n, l->xf, 0.f)) != 0) goto done;
          if ((rc = cycle_top<float>(s, o, ops, adj, true)) != 0) goto done;
          if ((rc = fmk::from_f32(st, n, l->xf, zj)) != 0) goto done;
        }
        // w = A z_j; classical Gram-Schmidt (second pass only when the DGKS test asks), the norm, the Hessenberg
        // column and the normalisation (with the FP32 copy of v_{j+1} for the next FP32 cycle), all on the device:
        // per pass one dot sweep (|w|^2 riding along), one single-block step, one update
        if ((rc = fmk::spmv(st, A, zj, w)) != 0) goto done;
        float *fnext = o->precision == FM_PREC_FP32_FGMRES ? l->bf : nullptr;
        for (int pass = 0; pass < 2; pass++) {
          const double *skip = pass == 0 ? nullptr : hd + 129;
          hipLaunchKernelGGL(k_multi_dot, dim3(nb, (j + 2 + 7) / 8), dim3(256), 0, st, n, j + 1, j + 2, V, w, part, skip);
          hipLaunchKernelGGL(k_gs_step, dim3(1), dim3(256), 0, st, j, m, nb, part, hd, pass, H, cs, sn, g, relh, stt, bn);
          hipLaunchKernelGGL(k_multi_axpy, dim3(gv), dim3(256), 0, st, n, j + 1, V, hd + 64 * pass, w, hd + 131, fnext, skip);
        }
        if (hipMemcpyAsync(s->kpin + j, relh + j, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipEventRecord(s->kev[j], st) != hipSuccess) { rc = FM_EHIP; goto done; }
        j++;
        total++;
        if (j - 1 < check_at && j < m && total < o->max_iter) continue;
        // checkpoint: read the estimates of the steps launched since the last one
        if (hipEventSynchronize(s->kev[j - 1]) != hipSuccess) { rc = FM_EHIP; goto done; }
        for (int q = jv + 1; q < j; q++) {
          const double r = s->kpin[q], prev = q == jv + 1 ? rv : s->kpin[q - 1];
          const int it = total - j + q + 1;
          if (it < 256) rep->hist[it] = r;
          if (prev > 0. && r > 0. && r < prev) rho = std::min(rho, r / prev);
          if (r <= o->tol * 0.5 || !std::isfinite(r)) stop = true;
        }
        jv = j - 1;
        rv = s->kpin[jv];
        rel = rv;
        check_at = jv + 1;
        if (!stop && rv > 0.) check_at = jv + lookahead(rv);
      }
      // y = H^-1 g and x += Z y, on the device
      hipLaunchKernelGGL(k_fgmres_y, dim3(1), dim3(64), 0, st, j, m, H, g, y, stt);
      hipLaunchKernelGGL(k_multi_add, dim3(gv), dim3(256), 0, st, n, j, Z, y, l->x);
    }
  }
  // true FP64 residual at exit (already in V when the convergence check at the loop top ended the solve)
  if (res_fresh) {
    if (hipMemcpyAsync(l->res, V, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) { rc = FM_EHIP; goto done; }
  } else {
    if ((rc = fmk::resid(st, A, l->b, l->x, l->res)) != 0) goto done;
    rel = dev_dot<double>(c, n, l->res, l->res, true) / bn;
  }
done:
  (void)hipStreamSynchronize(st);
  rep->iters = total;
  rep->rel_resid = rel;
  rep->converged = rel <= o->tol;
  return rc;
}

// Vanka smoothing needs the patch inverses in the cycle's precision: 0 when they are there, else 1
static int vanka_store_missing(const fm_mg_opts *o, const std::vector<LOps> &ops) {
  if (o->smoother != FM_SMOOTH_VANKA) return 0;
  const bool fp32 = o->precision == FM_PREC_FP32_IR || o->precision == FM_PREC_FP32_FGMRES;
  for (size_t L = 1; L < ops.size(); L++)
    if (fp32 ? !ops[L].vinvf : !ops[L].vinv) return 1;
  return 0;
}
static int vanka_store_error(const fm_mg_opts *o) {
  const bool fp32 = o->precision == FM_PREC_FP32_IR || o->precision == FM_PREC_FP32_FGMRES;
  fm_set_error(std::string("Vanka inverses for ") + (fp32 ? "FP32" : "FP64") +
               " cycles are not kept (fm_system_set_vanka_store)");
  return FM_ESTATE;
}

extern "C" int fm_mg_solve(fm_system *s, const fm_mg_opts *o, fm_mg_report *rep) {
  if (!s || !o || !rep) return fm_einval(__func__);
  if (!s->initialized) return FM_ESTATE;
  FM_HIP(hipSetDevice(s->ctx->device));
  std::memset(rep, 0, sizeof(*rep));
  if (!s->mg_ready || s->mg_smoother != o->smoother) FM_CHECK(fm_mg_setup(s, o));
  if (vanka_store_missing(o, forward_ops(s))) {
    FM_CHECK(fm_mg_setup(s, o));  // e.g. the store mask changed since the last set-up
    if (vanka_store_missing(o, forward_ops(s))) return vanka_store_error(o);
  }
  std::vector<LOps> ops = forward_ops(s);
  Timer tm(s->ctx);
  int rc;
  if (o->precision == FM_PREC_FP64 || o->precision == FM_PREC_FP32_IR) rc = solve_stationary(s, o, ops, rep, 0);
  else rc = solve_fgmres(s, o, ops, rep, 0);
  rep->seconds = tm.stop();
  if (rc) return rc;
  return rep->converged ? FM_OK : FM_ENOCONV;
}

// ---- Vanka (vanka.hip) -----------------------------------------------------
int fm_vanka_invert(fm_system *s, int L, const double *Av, double **store, float **storef);
int fm_vanka_apply64(fm_system *s, int L, const CsrView<double> &A, const double *vinv, double *x, const double *b,
                     double *r, int its, double omega, int colored, bool xz, bool post);
int fm_vanka_apply32(fm_system *s, int L, const CsrView<float> &A, const float *vinv, float *x, const float *b,
                     float *r, int its, double omega, int colored, bool xz, bool post);
int fm_vanka_setup(fm_system *s, int L, std::vector<LOps> *ops) {
  LOps &o = (*ops)[L];
  int rc = fm_vanka_invert(s, L, o.A->sell.v, &o.v->vinv, &o.v->vinvf);
  if (rc) return rc;
  o.vinv = o.v->vinv;
  o.vinvf = o.v->vinvf;
  return 0;
}
int fm_vanka_sweep64(fm_system *s, int L, const LOps &o, double *x, const double *b, int its, double omega, int colored,
                     bool xz, bool post) {
  return fm_vanka_apply64(s, L, cview<double>(*o.A), o.vinv, x, b, o.v->r, its, omega, colored, xz, post);
}
int fm_vanka_sweep32(fm_system *s, int L, const LOps &o, float *x, const float *b, int its, double omega, int colored,
                     bool xz, bool post) {
  return fm_vanka_apply32(s, L, cview<float>(*o.A), o.vinvf, x, b, o.v->rf, its, omega, colored, xz, post);
}

// Transposed multigrid: A^T lambda = g on the adjoint hierarchy (A^T from the
// transposed assembly, restriction Prl^T, prolongation Rst^T, smoothers and
// coarse inverse of A^T), same outer solvers as the forward problem.
extern "C" int fm_adjoint_solve(fm_system *s, const fm_mg_opts *o, const double *g, double *lam, fm_mg_report *rep) {
  if (!s || !o || !g || !lam || !rep) return fm_einval(__func__);
  if (!s->initialized || s->alev.empty()) { fm_set_error("adjoint: assemble A^T on every level first"); return FM_ESTATE; }
  for (char a : s->a_assembled) if (!a) { fm_set_error("adjoint: A^T missing on a level"); return FM_ESTATE; }
  FM_HIP(hipSetDevice(s->ctx->device));
  std::memset(rep, 0, sizeof(*rep));
  std::vector<LOps> ops(s->alev.size());
  for (size_t L = 0; L < s->alev.size(); L++) {
    Level &a = s->alev[L];
    LOps &op = ops[L];
    op.n = a.n; op.A = &a.A; op.R = &a.R; op.P = &a.P; op.dinv = a.dinv; op.dinvf = a.dinvf; op.lam = a.lam;
    op.Ainv = a.Ainv; op.Ainvf = a.Ainvf; op.v = &a; op.vinv = a.vinv; op.vinvf = a.vinvf;
  }
  if (!s->adj_ready || s->adj_smoother != o->smoother || vanka_store_missing(o, ops)) {
    fm_drop_graphs(s);
    FM_CHECK(mg_setup_ops(s, o, ops, true));
    if (vanka_store_missing(o, ops)) return vanka_store_error(o);
    for (size_t L = 0; L < s->alev.size(); L++) {
      s->alev[L].lam = ops[L].lam; s->alev[L].Ainv = ops[L].Ainv; s->alev[L].Ainvf = ops[L].Ainvf;
    }
    s->adj_ready = true;
    s->adj_smoother = o->smoother;
  }
  int top = (int)ops.size() - 1;
  Level &at = s->alev[top];
  hipStream_t st = s->ctx->stream;
  // g arrives in the reference numbering; the hierarchy works in internal order
  FM_HIP(hipMemcpyAsync(at.t, g, at.n * sizeof(double), hipMemcpyHostToDevice, st));
  FM_CHECK(fmk::gather(st, at.n, at.ord, at.t, at.b));
  FM_CHECK(fmk::fill<double>(st, at.n, at.x, 0.0));
  Timer tm(s->ctx);
  int rc;
  if (o->precision == FM_PREC_FP64 || o->precision == FM_PREC_FP32_IR) rc = solve_stationary(s, o, ops, rep, 1);
  else rc = solve_fgmres(s, o, ops, rep, 1);
  rep->seconds = tm.stop();
  if (rc) return rc;
  FM_CHECK(fmk::gather(st, at.n, at.rank, at.x, at.t));
  FM_HIP(hipMemcpyAsync(lam, at.t, at.n * sizeof(double), hipMemcpyDeviceToHost, st));
  FM_HIP(hipStreamSynchronize(st));
  return rep->converged ? FM_OK : FM_ENOCONV;
}
