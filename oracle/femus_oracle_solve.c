/*
 * femus_oracle_solve.c -- TEST INFRASTRUCTURE (oracle) and CPU baseline.
 * Restates:
 *   contrib/laspack/operats.c:782-1047 (Mul_QV), contrib/matrix/laspack_vectorM.C:147-189 (resid)
 *   contrib/laspack/factor.c:25-320 (ILUFactor, non-symmetric branch)
 *   contrib/laspack/precond.c:75-110 (ILUPrecond)
 *   contrib/laspack/itersolv.c:1341-1510 (GMRESIter), rtc.c:48-72 (RTCResult)
 *   contrib/matrix/laspack_linear_solverM.C:90-93 (GMRES restart 30)
 *   src/MGSolverBase.C:163-289 (MGSolve / MGStep)
 * plus the build's GPU smoother algorithms (Chebyshev-Jacobi, damped Jacobi)
 * so that GPU multigrid iterates can be checked cycle by cycle.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "femus_oracle.h"

void om_spmv(int64_t n, const int64_t *rp, const int *ci, const double *v, const double *x, double *y) {
  for (int64_t i = 0; i < n; i++) {
    double s = 0.;
    for (int64_t q = rp[i]; q < rp[i + 1]; q++) s += v[q] * x[ci[q]];
    y[i] = s;
  }
}
void om_resid(int64_t n, const int64_t *rp, const int *ci, const double *v, const double *b, const double *x, double *r) {
  for (int64_t i = 0; i < n; i++) {
    double s = 0.;
    for (int64_t q = rp[i]; q < rp[i + 1]; q++) s += v[q] * x[ci[q]];
    r[i] = b[i] - s;
  }
}
static double nrm2(int64_t n, const double *x) { double s = 0.; for (int64_t i = 0; i < n; i++) s += x[i] * x[i]; return sqrt(s); }
static double dot(int64_t n, const double *x, const double *y) { double s = 0.; for (int64_t i = 0; i < n; i++) s += x[i] * y[i]; return s; }

/* ---- ILU in LASPACK's (D+L) D^-1 (D+U) form, computed row/column-wise ---- */
typedef struct { int64_t n; const int64_t *rp; const int *ci; double *B; int64_t *diag; } ilu_t;

static int64_t find(const int64_t *rp, const int *ci, int64_t row, int col) {
  int64_t lo = rp[row], hi = rp[row + 1] - 1;
  while (lo <= hi) { int64_t mid = (lo + hi) >> 1; if (ci[mid] == col) return mid; if (ci[mid] < col) lo = mid + 1; else hi = mid - 1; }
  return -1;
}

static int ilu_factor(ilu_t *f, int64_t n, const int64_t *rp, const int *ci, const double *v) {
  f->n = n; f->rp = rp; f->ci = ci;
  f->B = (double *)malloc(sizeof(double) * rp[n]);
  f->diag = (int64_t *)malloc(sizeof(int64_t) * n);
  memcpy(f->B, v, sizeof(double) * rp[n]);
  int maxlen = 0;
  for (int64_t i = 0; i < n; i++) { int l = (int)(rp[i + 1] - rp[i]); if (l > maxlen) maxlen = l; f->diag[i] = find(rp, ci, i, (int)i); if (f->diag[i] < 0) return -1; }
  double *Lm = (double *)malloc(sizeof(double) * (maxlen + 1) * (maxlen + 1));
  int *S = (int *)malloc(sizeof(int) * (maxlen + 1));
  for (int64_t i = 0; i < n; i++) {
    int LD = 0;
    for (int64_t q = rp[i]; q < rp[i + 1] && ci[q] <= i; q++) S[LD++] = ci[q];
    /* dense L[a][b] = B(S[a], S[b]) for entries present */
    for (int a = 0; a < LD; a++) for (int b = 0; b < LD; b++) Lm[a * LD + b] = 0.;
    for (int a = 0; a < LD; a++) {
      int r = S[a];
      int bidx = 0;
      for (int64_t q = rp[r]; q < rp[r + 1] && ci[q] <= i; q++) {
        while (bidx < LD && S[bidx] < ci[q]) bidx++;
        if (bidx < LD && S[bidx] == ci[q]) Lm[a * LD + bidx] = f->B[q];
      }
    }
    int l = LD - 1;
    for (int j = 0; j < l; j++) for (int k = j + 1; k < l; k++) Lm[l * LD + k] -= Lm[l * LD + j] * Lm[j * LD + k] / Lm[j * LD + j];
    for (int j = 0; j < l; j++) for (int k = j + 1; k < l; k++) Lm[k * LD + l] -= Lm[j * LD + l] * Lm[k * LD + j] / Lm[j * LD + j];
    for (int j = 0; j < l; j++) Lm[l * LD + l] -= Lm[j * LD + l] * Lm[l * LD + j] / Lm[j * LD + j];
    if (Lm[l * LD + l] == 0.) { free(Lm); free(S); return -2; }
    /* write back row i (cols <= i) and column i (rows < i) */
    for (int a = 0; a < LD; a++) f->B[find(rp, ci, i, S[a])] = Lm[l * LD + a];
    for (int a = 0; a < l; a++) {
      int64_t q = find(rp, ci, S[a], (int)i);
      if (q < 0) { free(Lm); free(S); return -3; }
      f->B[q] = Lm[a * LD + l];
    }
  }
  free(Lm); free(S);
  return 0;
}
static void ilu_free(ilu_t *f) { free(f->B); free(f->diag); }
/* y = (D+U)^-1 D (D+L)^-1 c  (precond.c:99-101) */
static void ilu_apply(const ilu_t *f, const double *c, double *y) {
  int64_t n = f->n;
  for (int64_t i = 0; i < n; i++) {
    double s = c[i];
    for (int64_t q = f->rp[i]; q < f->diag[i]; q++) s -= f->B[q] * y[f->ci[q]];
    y[i] = s / f->B[f->diag[i]];
  }
  for (int64_t i = 0; i < n; i++) y[i] *= f->B[f->diag[i]];
  for (int64_t i = n - 1; i >= 0; i--) {
    double s = y[i];
    for (int64_t q = f->diag[i] + 1; q < f->rp[i + 1]; q++) s -= f->B[q] * y[f->ci[q]];
    y[i] = s / f->B[f->diag[i]];
  }
}

/* ---- GMRES(restart) with left preconditioning (itersolv.c:1341-1510) ---- */
static int gmres_core(int64_t n, const int64_t *rp, const int *ci, const double *v, const ilu_t *P,
                      const double *b, double *x, int maxit, int m, double tol, double *acc) {
  double **V = (double **)malloc(sizeof(double *) * (m + 2));
  for (int i = 1; i <= m + 1; i++) V[i] = (double *)malloc(sizeof(double) * n);
  double *h = (double *)calloc((size_t)(m + 1) * (m + 2), sizeof(double));
#define H(i, k) h[(size_t)(i) * (m + 2) + (k)]
  double *y = (double *)malloc(sizeof(double) * (m + 1)), *s = (double *)malloc(sizeof(double) * (m + 2));
  double *c1 = (double *)malloc(sizeof(double) * (m + 1)), *c2 = (double *)malloc(sizeof(double) * (m + 1));
  double *w = (double *)malloc(sizeof(double) * n);
  double bn = nrm2(n, b);
  int Iter = 0;
  om_resid(n, rp, ci, v, b, x, V[1]);
  double rn = nrm2(n, V[1]);
  *acc = bn != 0. ? rn / bn : 1.;
  while (!(rn < tol * bn) && Iter < maxit) {
    if (P) { ilu_apply(P, V[1], w); memcpy(V[1], w, sizeof(double) * n); }
    s[1] = nrm2(n, V[1]);
    for (int64_t q = 0; q < n; q++) V[1][q] *= 1.0 / s[1];
    int i = 0;
    while ((P ? 1 : !(fabs(s[i + 1]) < tol * bn)) && i < m && Iter < maxit) {
      i++; Iter++;
      om_spmv(n, rp, ci, v, V[i], w);
      if (P) ilu_apply(P, w, V[i + 1]); else memcpy(V[i + 1], w, sizeof(double) * n);
      for (int k = 1; k <= i; k++) {
        H(i, k) = dot(n, V[i + 1], V[k]);
        for (int64_t q = 0; q < n; q++) V[i + 1][q] -= H(i, k) * V[k][q];
      }
      H(i, i + 1) = nrm2(n, V[i + 1]);
      for (int64_t q = 0; q < n; q++) V[i + 1][q] *= 1.0 / H(i, i + 1);
      for (int k = 1; k < i; k++) {
        double h1 = c1[k] * H(i, k) + c2[k] * H(i, k + 1);
        double h2 = -c2[k] * H(i, k) + c1[k] * H(i, k + 1);
        H(i, k) = h1; H(i, k + 1) = h2;
      }
      double r = sqrt(H(i, i) * H(i, i) + H(i, i + 1) * H(i, i + 1));
      c1[i] = H(i, i) / r; c2[i] = H(i, i + 1) / r;
      H(i, i) = r; H(i, i + 1) = 0.;
      s[i + 1] = -c2[i] * s[i];
      s[i] = c1[i] * s[i];
    }
    for (int j = i; j > 0; j--) {
      y[j] = s[j] / H(j, j);
      for (int k = j - 1; k > 0; k--) s[k] -= H(j, k) * y[j];
    }
    for (int j = i; j > 0; j--) for (int64_t q = 0; q < n; q++) x[q] += y[j] * V[j][q];
    om_resid(n, rp, ci, v, b, x, V[1]);
    rn = nrm2(n, V[1]);
    *acc = bn != 0. ? rn / bn : 1.;
  }
#undef H
  for (int i = 1; i <= m + 1; i++) free(V[i]);
  free(V); free(h); free(y); free(s); free(c1); free(c2); free(w);
  return Iter;
}

int om_gmres_ilu(int64_t n, const int64_t *rp, const int *ci, const double *v, const double *b, double *x,
                 int maxit, int restart, double tol, int use_ilu, double *acc) {
  ilu_t P;
  if (use_ilu && ilu_factor(&P, n, rp, ci, v) != 0) return -1;
  int it = gmres_core(n, rp, ci, v, use_ilu ? &P : NULL, b, x, maxit, restart, tol, acc);
  if (use_ilu) ilu_free(&P);
  return it;
}

/* ---- Chebyshev / Jacobi smoothers (the GPU algorithms) ---- */
static void diag_inv(int64_t n, const int64_t *rp, const int *ci, const double *v, double *dinv) {
  for (int64_t i = 0; i < n; i++) {
    double d = 0.;
    for (int64_t q = rp[i]; q < rp[i + 1]; q++) if (ci[q] == i) d = v[q];
    dinv[i] = d != 0. ? 1. / d : 0.;
  }
}

double om_lambda_max(int64_t n, const int64_t *rp, const int *ci, const double *v, int iters) {
  double *x = (double *)malloc(sizeof(double) * n), *y = (double *)malloc(sizeof(double) * n);
  double *dinv = (double *)malloc(sizeof(double) * n);
  diag_inv(n, rp, ci, v, dinv);
  for (int64_t i = 0; i < n; i++) x[i] = 1.0 + (double)(i % 7) / 7.0;
  double lam = 0.;
  for (int it = 0; it < iters; it++) {
    double xn = nrm2(n, x);
    for (int64_t i = 0; i < n; i++) x[i] /= xn;
    om_spmv(n, rp, ci, v, x, y);
    for (int64_t i = 0; i < n; i++) y[i] *= dinv[i];
    lam = nrm2(n, y);
    memcpy(x, y, sizeof(double) * n);
  }
  free(x); free(y); free(dinv);
  return lam;
}

/* Chebyshev iteration on D^-1 A over [lo, hi]; returns ||b-Ax||/||b||. */
static double cheb_smooth(int64_t n, const int64_t *rp, const int *ci, const double *v, const double *dinv,
                          const double *b, double *x, int its, double lo, double hi) {
  double *r = (double *)malloc(sizeof(double) * n), *d = (double *)malloc(sizeof(double) * n);
  double *t = (double *)malloc(sizeof(double) * n);
  double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta, rho = 1. / sigma;
  om_resid(n, rp, ci, v, b, x, r);
  for (int64_t i = 0; i < n; i++) d[i] = dinv[i] * r[i] / theta;
  for (int k = 1; k <= its; k++) {
    for (int64_t i = 0; i < n; i++) x[i] += d[i];
    if (k == its) break;
    om_spmv(n, rp, ci, v, d, t);
    for (int64_t i = 0; i < n; i++) r[i] -= t[i];
    double rho_new = 1. / (2. * sigma - rho);
    for (int64_t i = 0; i < n; i++) d[i] = rho_new * rho * d[i] + (2. * rho_new / delta) * dinv[i] * r[i];
    rho = rho_new;
  }
  om_resid(n, rp, ci, v, b, x, r);
  double bn = nrm2(n, b), rn = nrm2(n, r);
  free(r); free(d); free(t);
  return bn != 0. ? rn / bn : 1.;
}
static double jacobi_smooth(int64_t n, const int64_t *rp, const int *ci, const double *v, const double *dinv,
                            const double *b, double *x, int its, double omega) {
  double *r = (double *)malloc(sizeof(double) * n);
  for (int k = 0; k < its; k++) {
    om_resid(n, rp, ci, v, b, x, r);
    for (int64_t i = 0; i < n; i++) x[i] += omega * dinv[i] * r[i];
  }
  om_resid(n, rp, ci, v, b, x, r);
  double bn = nrm2(n, b), rn = nrm2(n, r);
  free(r);
  return bn != 0. ? rn / bn : 1.;
}

/* ---- dense LU ---- */
int om_dense_lu_solve(int n, double *A, double *b) {
  for (int k = 0; k < n; k++) {
    int p = k; double mx = fabs(A[(size_t)k * n + k]);
    for (int i = k + 1; i < n; i++) if (fabs(A[(size_t)i * n + k]) > mx) { mx = fabs(A[(size_t)i * n + k]); p = i; }
    if (mx == 0.) return -1;
    if (p != k) {
      for (int j = 0; j < n; j++) { double t = A[(size_t)k * n + j]; A[(size_t)k * n + j] = A[(size_t)p * n + j]; A[(size_t)p * n + j] = t; }
      double t = b[k]; b[k] = b[p]; b[p] = t;
    }
    for (int i = k + 1; i < n; i++) {
      double l = A[(size_t)i * n + k] / A[(size_t)k * n + k];
      if (l == 0.) continue;
      for (int j = k + 1; j < n; j++) A[(size_t)i * n + j] -= l * A[(size_t)k * n + j];
      b[i] -= l * b[k];
    }
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = b[i];
    for (int j = i + 1; j < n; j++) s -= A[(size_t)i * n + j] * b[j];
    b[i] = s / A[(size_t)i * n + i];
  }
  return 0;
}

/* ---- multigrid hierarchy ---- */
typedef struct { int64_t nr, nc; int64_t *rp; int *ci; double *v; } csr_t;
struct om_hier {
  int nlev;
  csr_t *A, *R, *P;
  double **x, **b, **res, **dinv, *lam;
  ilu_t *ilu; int *ilu_ok;
  double *lu; int lu_n;
};

static void csr_set(csr_t *c, int64_t nr, int64_t nc, const int64_t *rp, const int *ci, const double *v) {
  free(c->rp); free(c->ci); free(c->v);
  c->nr = nr; c->nc = nc;
  c->rp = (int64_t *)malloc(sizeof(int64_t) * (nr + 1)); memcpy(c->rp, rp, sizeof(int64_t) * (nr + 1));
  c->ci = (int *)malloc(sizeof(int) * (rp[nr] + 1)); memcpy(c->ci, ci, sizeof(int) * rp[nr]);
  c->v = (double *)malloc(sizeof(double) * (rp[nr] + 1)); memcpy(c->v, v, sizeof(double) * rp[nr]);
}

om_hier *om_hier_new(int nlev) {
  om_hier *h = (om_hier *)calloc(1, sizeof(om_hier));
  h->nlev = nlev;
  h->A = (csr_t *)calloc(nlev, sizeof(csr_t)); h->R = (csr_t *)calloc(nlev, sizeof(csr_t)); h->P = (csr_t *)calloc(nlev, sizeof(csr_t));
  h->x = (double **)calloc(nlev, sizeof(double *)); h->b = (double **)calloc(nlev, sizeof(double *));
  h->res = (double **)calloc(nlev, sizeof(double *)); h->dinv = (double **)calloc(nlev, sizeof(double *));
  h->lam = (double *)calloc(nlev, sizeof(double));
  h->ilu = (ilu_t *)calloc(nlev, sizeof(ilu_t)); h->ilu_ok = (int *)calloc(nlev, sizeof(int));
  return h;
}
void om_hier_free(om_hier *h) {
  for (int L = 0; L < h->nlev; L++) {
    free(h->A[L].rp); free(h->A[L].ci); free(h->A[L].v);
    free(h->R[L].rp); free(h->R[L].ci); free(h->R[L].v);
    free(h->P[L].rp); free(h->P[L].ci); free(h->P[L].v);
    free(h->x[L]); free(h->b[L]); free(h->res[L]); free(h->dinv[L]);
    if (h->ilu_ok[L]) ilu_free(&h->ilu[L]);
  }
  free(h->A); free(h->R); free(h->P); free(h->x); free(h->b); free(h->res); free(h->dinv); free(h->lam);
  free(h->ilu); free(h->ilu_ok); free(h->lu); free(h);
}
void om_hier_set_A(om_hier *h, int L, int64_t n, const int64_t *rp, const int *ci, const double *v) {
  csr_set(&h->A[L], n, n, rp, ci, v);
  free(h->x[L]); free(h->b[L]); free(h->res[L]); free(h->dinv[L]);
  h->x[L] = (double *)calloc(n, sizeof(double)); h->b[L] = (double *)calloc(n, sizeof(double));
  h->res[L] = (double *)calloc(n, sizeof(double)); h->dinv[L] = (double *)calloc(n, sizeof(double));
  diag_inv(n, rp, ci, v, h->dinv[L]);
  h->lam[L] = 0.;
  if (h->ilu_ok[L]) { ilu_free(&h->ilu[L]); h->ilu_ok[L] = 0; }
  if (L == 0) { free(h->lu); h->lu = NULL; }
}
void om_hier_set_R(om_hier *h, int L, int64_t nr, int64_t nc, const int64_t *rp, const int *ci, const double *v) { csr_set(&h->R[L], nr, nc, rp, ci, v); }
void om_hier_set_P(om_hier *h, int L, int64_t nr, int64_t nc, const int64_t *rp, const int *ci, const double *v) { csr_set(&h->P[L], nr, nc, rp, ci, v); }

static double smooth(om_hier *h, const om_mg_opts *o, int L, int its) {
  csr_t *A = &h->A[L];
  int64_t n = A->nr;
  double acc = 1.;
  if (o->smoother == 0) {
    if (!h->ilu_ok[L]) { if (ilu_factor(&h->ilu[L], n, A->rp, A->ci, A->v) != 0) return -1.; h->ilu_ok[L] = 1; }
    gmres_core(n, A->rp, A->ci, A->v, &h->ilu[L], h->b[L], h->x[L], its, 30, 1.e-20, &acc);
    return acc;
  }
  if (o->smoother == 1) {
    if (h->lam[L] == 0.) h->lam[L] = om_lambda_max(n, A->rp, A->ci, A->v, 20);
    return cheb_smooth(n, A->rp, A->ci, A->v, h->dinv[L], h->b[L], h->x[L], its, o->cheb_lo * h->lam[L], o->cheb_hi * h->lam[L]);
  }
  return jacobi_smooth(n, A->rp, A->ci, A->v, h->dinv[L], h->b[L], h->x[L], its, o->jacobi_omega);
}

static double coarse_solve(om_hier *h, const om_mg_opts *o) {
  csr_t *A = &h->A[0];
  int n = (int)A->nr;
  if (o->coarse == 0) return smooth(h, o, 0, 60);
  double *M = (double *)calloc((size_t)n * n, sizeof(double));
  for (int i = 0; i < n; i++) for (int64_t q = A->rp[i]; q < A->rp[i + 1]; q++) M[(size_t)i * n + A->ci[q]] = A->v[q];
  memcpy(h->x[0], h->b[0], sizeof(double) * n);
  om_dense_lu_solve(n, M, h->x[0]);
  free(M);
  double *r = (double *)malloc(sizeof(double) * n);
  om_resid(n, A->rp, A->ci, A->v, h->b[0], h->x[0], r);
  double bn = nrm2(n, h->b[0]), rn = nrm2(n, r);
  free(r);
  return bn != 0. ? rn / bn : 1.;
}

/* MGSolBase::MGStep (src/MGSolverBase.C:212-289) */
static double mg_step(om_hier *h, const om_mg_opts *o, int L) {
  double rest;
  csr_t *A = &h->A[L];
  if (L == 0) {
    rest = coarse_solve(h, o);
    om_resid(A->nr, A->rp, A->ci, A->v, h->b[0], h->x[0], h->res[0]);
    return rest;
  }
  int top = h->nlev - 1;
  int npre = o->nu_pre * (L < top ? 2 : 1);
  rest = smooth(h, o, L, npre);
  om_resid(A->nr, A->rp, A->ci, A->v, h->b[L], h->x[L], h->res[L]);
  csr_t *R = &h->R[L - 1];
  om_spmv(R->nr, R->rp, R->ci, R->v, h->res[L], h->b[L - 1]);
  memset(h->x[L - 1], 0, sizeof(double) * h->A[L - 1].nr);
  double crest = 0.;
  for (int g = 1; g <= o->gamma; g++) crest = mg_step(h, o, L - 1);
  csr_t *P = &h->P[L];
  om_spmv(P->nr, P->rp, P->ci, P->v, h->x[L - 1], h->res[L]);
  if (o->correction == 0) {
    if (crest < rest) for (int64_t i = 0; i < A->nr; i++) h->x[L][i] += 0.95 * h->res[L][i];
  } else {
    for (int64_t i = 0; i < A->nr; i++) h->x[L][i] += h->res[L][i];
  }
  int npost = o->nu_post * (L < top ? 2 : 1);
  rest = smooth(h, o, L, npost);
  om_resid(A->nr, A->rp, A->ci, A->v, h->b[L], h->x[L], h->res[L]);
  return rest;
}

/* MGSolBase::MGSolve (src/MGSolverBase.C:163-208) */
int om_mg_solve(om_hier *h, const om_mg_opts *o, const double *b, double *x, double *hist) {
  int top = h->nlev - 1;
  int64_t n = h->A[top].nr;
  memcpy(h->b[top], b, sizeof(double) * n);
  memcpy(h->x[top], x, sizeof(double) * n);
  double bn = nrm2(n, b);
  int cycle = 0;
  double *r = (double *)malloc(sizeof(double) * n);
  while (cycle < o->max_cycles) {
    double rest = mg_step(h, o, top);
    cycle++;
    csr_t *A = &h->A[top];
    om_resid(n, A->rp, A->ci, A->v, h->b[top], h->x[top], r);
    double rel = bn != 0. ? nrm2(n, r) / bn : 0.;
    if (hist) hist[cycle - 1] = rel;
    if (o->ref_exit ? (rest < o->eps * (1. + bn)) : (rel <= o->eps)) break;
  }
  memcpy(x, h->x[top], sizeof(double) * n);
  free(r);
  return cycle;
}
