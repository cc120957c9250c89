/*
 * laspack_ref_driver.c -- TEST INFRASTRUCTURE.  A thin C driver, written for
 * this build, that calls the reference's vendored LASPACK library (compiled
 * from contrib/laspack/*.c where it lies, into oracle/_ref/) so that the
 * oracle's SpMV / ILU / GMRES restatement can be checked against the real
 * implementation on small systems.  Entry points take CSR arrays.
 * LASPACK API used: qmatrix.h:59 Q_Constr, :67-69 Q_SetLen/Q_SetEntry,
 * operats.h:47 Mul_QV, precond.h:36 ILUPrecond, itersolv.h:65 GMRESIter,
 * itersolv.h:75 SetGMRESRestart, rtc.h:68-72 SetRTCAccuracy/GetLast*.
 */
#include <stdint.h>
#include <stdlib.h>
#include "qmatrix.h"
#include "qvector.h"
#include "operats.h"
#include "precond.h"
#include "itersolv.h"
#include "rtc.h"
#include "errhandl.h"

static void build_q(QMatrix *Q, int64_t n, const int64_t *rp, const int *ci, const double *v) {
  Q_Constr(Q, "A", (size_t)n, _LPFalse, Rowws, Normal, _LPTrue);
  for (int64_t i = 0; i < n; i++) {
    size_t len = (size_t)(rp[i + 1] - rp[i]);
    Q_SetLen(Q, (size_t)(i + 1), len);
    for (size_t k = 0; k < len; k++)
      Q_SetEntry(Q, (size_t)(i + 1), k, (size_t)(ci[rp[i] + k] + 1), v[rp[i] + k]);
  }
}
static void build_v(QVector *V, int64_t n, const double *x) {
  V_Constr(V, "v", (size_t)n, Normal, _LPTrue);
  for (int64_t i = 0; i < n; i++) V_SetCmp(V, (size_t)(i + 1), x ? x[i] : 0.);
}
static void get_v(QVector *V, int64_t n, double *x) {
  for (int64_t i = 0; i < n; i++) x[i] = V_GetCmp(V, (size_t)(i + 1));
}

int lr_spmv(int64_t n, const int64_t *rp, const int *ci, const double *v, const double *x, double *y) {
  QMatrix Q; QVector X, Y;
  build_q(&Q, n, rp, ci, v); build_v(&X, n, x); build_v(&Y, n, NULL);
  Asgn_VV(&Y, Mul_QV(&Q, &X));
  get_v(&Y, n, y);
  int err = LASResult() != LASOK;
  Q_Destr(&Q); V_Destr(&X); V_Destr(&Y);
  return err ? -1 : 0;
}

int lr_ilu_apply(int64_t n, const int64_t *rp, const int *ci, const double *v, const double *c, double *y) {
  QMatrix Q; QVector C, Y;
  build_q(&Q, n, rp, ci, v); build_v(&C, n, c); build_v(&Y, n, NULL);
  ILUPrecond(&Q, &Y, &C, 1.0);
  get_v(&Y, n, y);
  int err = LASResult() != LASOK;
  Q_Destr(&Q); V_Destr(&C); V_Destr(&Y);
  return err ? -1 : 0;
}

int lr_gmres(int64_t n, const int64_t *rp, const int *ci, const double *v, const double *b, double *x,
             int maxit, int restart, int use_ilu, double tol, double *acc, int *iters) {
  QMatrix Q; QVector B, X;
  build_q(&Q, n, rp, ci, v); build_v(&B, n, b); build_v(&X, n, x);
  SetRTCAccuracy(tol);
  SetGMRESRestart(restart);
  GMRESIter(&Q, &X, &B, maxit, use_ilu ? ILUPrecond : NULL, 1.0);
  get_v(&X, n, x);
  *acc = GetLastAccuracy();
  *iters = GetLastNoIter();
  int err = LASResult() != LASOK;
  Q_Destr(&Q); V_Destr(&B); V_Destr(&X);
  return err ? -1 : 0;
}
