// fm_kernels.h -- launchers of the gfx950 kernels (kernels.hip, assemble.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

template <typename T>
struct CsrView {
  int64_t nr = 0;
  const int64_t *rp = nullptr;
  const int *ci = nullptr;
  const T *v = nullptr;
  int lpr = 16;  // lanes per row
  // SELL-64 copy in internal DoF order (used when nsl > 0)
  int64_t nsl = 0, nslot = 0;
  const int64_t *sl_off = nullptr, *sl_c32 = nullptr, *sl_c8 = nullptr, *sl_c16 = nullptr, *sl_c2 = nullptr;
  const int *sl_w = nullptr, *perm = nullptr, *base = nullptr, *c32 = nullptr;
  const unsigned *c2 = nullptr;
  const unsigned short *c16 = nullptr;
  const unsigned char *c8 = nullptr;
  const T *sv = nullptr;
  // optional restriction of the SELL product to the slices slist[0 .. nlist-1]
  const int *slist = nullptr;
  int64_t nlist = 0;
};

// SELL slot <-> CSR entry correspondence of one matrix: slot (s, k, l) of internal row i = perm[64 s + l] holds
// CSR entry rbase[i] + pos[slot] (pos 0xFFFF: padding)
struct SellMap {
  int64_t nsl = 0;
  const int64_t *sl_off = nullptr;
  const int *sl_w = nullptr, *perm = nullptr;
  const int64_t *rbase = nullptr;
  const unsigned short *pos = nullptr;
};

// Reference-element tables (src/MGFE.C:195-924 layout: phi[s*ng+qp],
// dphi[(s*ng+qp) + d*ns*ng]); face tables are the (dim-1) Q2 tables.
struct FeTab {
  double phi2[27 * 27];
  double dphi2[3 * 27 * 27];
  double w[27];
  double phi1[8 * 27];
  double bphi2[9 * 9];
  double bdphi2[2 * 9 * 9];
  double bw[9];
};

struct HeatArgs {
  double alpha0, rhocp, dt, tc, source, upwind;
  double vel_const[3];
};
struct NsArgs {
  double IRe, dt, tc, upwind, IFr;
  double dirg[3], src[3];
  double buoy[3];          // Boussinesq body force buoy * T (T: nodal temperature field), zero when unused
  const double *Tn;        // nodal temperature [Ntop] or nullptr
  int newton = 0;          // add the Newton term phi_i phi_j du_iv/dx_jv of the advection (steady Jacobian)
};

namespace fmk {
int lanes_for(int64_t nnz, int64_t n);
int sell_split(int W);  // SELL split threshold (slices); W < 0 reads it; returns the previous value
int spmv(hipStream_t s, const CsrView<double> &A, const double *x, double *y);
int spmv(hipStream_t s, const CsrView<float> &A, const float *x, float *y);
int resid(hipStream_t s, const CsrView<double> &A, const double *b, const double *x, double *r);
int resid(hipStream_t s, const CsrView<float> &A, const float *b, const float *x, float *r);
// Chebyshev step: r -= A dold; dnew = c1 dold + c2 dinv r; x += dnew (x += dold + dnew with xold = 1)
int cheb_step(hipStream_t s, const CsrView<double> &A, double *r, double *x, const double *dold, double *dnew,
              const double *dinv, double c1, double c2, int xold = 0);
int cheb_step(hipStream_t s, const CsrView<float> &A, float *r, float *x, const float *dold, float *dnew,
              const float *dinv, float c1, float c2, int xold = 0);
// r = b - A x and d = a dinv r in one product (x unchanged)
int resid_d(hipStream_t s, const CsrView<double> &A, const double *b, const double *x, double *r, double *d,
            const double *dinv, double a);
int resid_d(hipStream_t s, const CsrView<float> &A, const float *b, const float *x, float *r, float *d,
            const float *dinv, float a);
int addmul(hipStream_t s, const CsrView<double> &P, const double *xc, double *x, double alpha, const double *cond);
int addmul(hipStream_t s, const CsrView<float> &P, const float *xc, float *x, float alpha, const double *cond);
int jacobi_step(hipStream_t s, const CsrView<double> &A, const double *b, const double *x, double *xn,
                const double *dinv, double omega);
int jacobi_step(hipStream_t s, const CsrView<float> &A, const float *b, const float *x, float *xn,
                const float *dinv, float omega);
template <typename T> int fill(hipStream_t s, int64_t n, T *x, T a);
template <typename T> int copy(hipStream_t s, int64_t n, const T *x, T *y);  // y = x
template <typename T> int axpby(hipStream_t s, int64_t n, T a, T *x, T b, const T *y);
template <typename T> int cheb_first(hipStream_t s, int64_t n, const T *dinv, const T *r, T *d, T *x, T a);
template <typename T> int cheb_start(hipStream_t s, int64_t n, const T *dinv, const T *b, T *r, T *d, T *x, T a);
int from_f32(hipStream_t s, int64_t n, const float *a, double *b);
int to_f32(hipStream_t s, int64_t n, const double *a, float *b, const double *scale);
int add_f32(hipStream_t s, int64_t n, double *x, const float *e, const double *scale);
template <typename T> int dot_dev(hipStream_t s, int64_t n, const T *x, const T *y, double *part, double *out, int do_sqrt);
int ratio(hipStream_t s, const double *a, const double *b, double *out);
template <typename T> int dense_mv(hipStream_t s, int n, const T *M, const T *b, T *y);
int gather(hipStream_t s, int64_t n, const int *idx, const double *src, double *dst);
int node_gather(hipStream_t s, int64_t N, int nvar, const int *dof, const int *rank, const double *x, double *out);
int dinv_slot(hipStream_t s, int64_t n, const int64_t *dslot, const double *sv, double *d64, float *d32);
int sell2csr(hipStream_t s, const SellMap &m, const double *sv, double *cv);
int csr2sell(hipStream_t s, const SellMap &m, const double *csr_v, double *sv, float *svf);
int sell_kof(hipStream_t s, const SellMap &m, unsigned short *kof);
int diag_slot(hipStream_t s, int64_t n, const int64_t *rp, const int *ci, const int *ord, const int64_t *rslot,
              const unsigned short *kof, int64_t *dslot);
int elslot(hipStream_t s, int64_t n, const int *eldofi, const int64_t *rslot, int64_t *out);
int cast_vals(hipStream_t s, int64_t n, const double *v, float *vf);
int elpos(hipStream_t s, int nel, int nloc, const int *eldof, const int64_t *rp, const int *ci,
          const unsigned short *kof, unsigned short *pos);
int copy16(hipStream_t s, int64_t n16, const void *a, void *b);

// assemble.hip
int heat_assemble(hipStream_t s, int dim, int nel_color, const int *elems, const int *conn, const int *eldof,
                  const unsigned short *elpos, const int64_t *elslot, const unsigned char *bmask, const double *xyz, int Ntop, const int *dof_top,
                  const int *bc0, const int *bc1, const double *x_old, const double *xbc, const double *coef,
                  const double *vel, const FeTab *tab, HeatArgs a, int mode, int trans, int bd_only, double *val,
                  double *rhs);
int ns_assemble(hipStream_t s, int dim, int nel_color, const int *elems, const int *conn, const int *eldof,
                const unsigned short *elpos, const int64_t *elslot, const unsigned char *bmask, const double *xyz, int Ntop, const int *dof_top,
                const int *bc0, const int *bc1, const double *x_old, const double *xbc, const double *coef,
                const FeTab *tab, NsArgs a, int mode, int trans, int bd_only, double *val, double *rhs);
// element-buffer assembly: element matrices / rhs of every element to E [nel][nl][nl] and Fe [nel][nl] (face
// terms included), then egather sums them into the SELL values (FP64 and FP32) and the rhs
int heat_assemble_e(hipStream_t s, int dim, int nel, const int *conn, const unsigned char *bmask, const double *xyz, int Ntop,
                    const int *dof_top, const int *bc0, const int *bc1, const double *x_old, const double *xbc,
                    const double *coef, const double *vel, const FeTab *tab, HeatArgs a, int mode, int trans, double *E,
                    double *Fe);
int ns_assemble_e(hipStream_t s, int dim, int nel, const int *conn, const unsigned char *bmask, const double *xyz, int Ntop,
                  const int *dof_top, const int *bc0, const int *bc1, const double *x_old, const double *xbc,
                  const double *coef, const FeTab *tab, NsArgs a, int mode, int trans, double *E, double *Fe);
int egather(hipStream_t s, int64_t nblk, int lg, int nl, int oc, int os, int on, size_t lds, const int64_t *sl_off,
            const int *sl_w, const int *perm, const int64_t *lseg, const int *seg, const int64_t *bbase,
            const unsigned char *cnt, const unsigned short *cst, const unsigned short *cidx, const double *E,
            const double *Fe, double *v, float *vf, double *rhs);
int coupling(hipStream_t s, int dim, int nel, const int *elems, const int *conn, const double *xyz, int Ntop,
             const int *dof_top, const int *bc0, const double *T, const double *lam, const double *m, double scale, int mode,
             const double *vel, double upwind, double rhocp, double alpha0, const FeTab *tab, double *out);
int heat_nusselt(hipStream_t s, int dim, int nel, const int *elems, const int *conn, const int *eldof, const double *xyz,
                 int Ntop, const int *dof_top, const double *vel, const double *T, const double *coef, double kappa,
                 int axis, const FeTab *tab, double *Jel, double *dJdk, double *y);
int ns_visc(hipStream_t s, int dim, int nel, const int *elems, const int *conn, const int *eldof, const double *xyz,
            int Ntop, const int *dof_top, const int *bc0, const double *U, const double *lam, const double *coef,
            double nu0, const FeTab *tab, double *diss, double *adj, double *y, double ysc);
int heat_lap(hipStream_t s, int dim, int nel, const int *elems, const int *conn, const int *eldof, const double *xyz,
             int Ntop, const int *dof_top, const int *bc0, const double *T, const double *lam, const double *coef,
             const FeTab *tab, double *diss, double *adj, double *y, double ysc);
}  // namespace fmk
