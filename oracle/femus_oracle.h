/*
 * femus_oracle.h -- CPU (FP64, single-thread) restatement of the FEMuS hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the CPU
 * baseline.  It may be loaded by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py -- never by the product library (exfemus_amd).
 *
 * Each function cites the reference file:line it restates.  Layouts follow the
 * reference:  coordinates SoA  xyz[node + d*Ntop];  element connectivity
 * conn[e*n2 + k] in libMesh local node order (Quad9 / Hex27);  DoF map
 * node_dof[node + ivar*Ntop] (-1 if absent);  element matrices row-major.
 */
#ifndef FEMUS_ORACLE_H
#define FEMUS_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct om_mesh om_mesh;

/* ---- FE tables (src/MGFE.C:195-924) ------------------------------------ */
/* kind: 2 = Q2 (Quad9/Hex27), 1 = Q1.  dim = 1,2,3.  Fills
 *   phi[ish*ng+qp], dphi[(ish*ng+qp) + idir*ns*ng], w[qp]; returns ns, sets *ng. */
int om_fe_tables(int kind, int dim, double *phi, double *dphi, double *w, int *ng);
/* Child embedding tables  E[c][i][j] = phi_j(child node i)  (src/MGGeomEl.C:182-690). */
int om_embedding(int kind, int dim, double *E);

/* ---- mesh hierarchy (USER_APPL/gencase/src/MGGenCase.C:63-758) --------- */
om_mesh *om_mesh_build(int dim, const int nx0[3], const double lo[3],
                       const double hi[3], int nlev);
void om_mesh_free(om_mesh *m);
/* info[0..5] = dim, nlev, n2, n1, nsides, n_nodes_top */
void om_mesh_info(const om_mesh *m, int *info);
/* per level: n_el, n_nodes(Q2), n_nodes_q1 */
void om_mesh_level(const om_mesh *m, int L, int *n_el, int *nq2, int *nq1);
void om_mesh_coords(const om_mesh *m, double *xyz);
void om_mesh_set_coords(om_mesh *m, const double *xyz);
void om_mesh_conn(const om_mesh *m, int L, int *conn, int *neigh, int *parent);
/* libMesh-style id and (lev, levp) keys of every top node (ordering audit) */
void om_mesh_node_keys(const om_mesh *m, int *libmesh_id, int *lev, int *levp);

/* ---- DoF map and BC (src/MGSolverDA.C:880-948, 696-862) ---------------- */
int om_init_dof(const om_mesh *m, int L, int nq2, int nq1, int *node_dof);
/* bc_read callback: xp[dim] -> bc_neum[nvars], bc_flag[nvars] (both preset to 1) */
typedef void (*om_bc_fn)(const double *xp, int *bc_neum, int *bc_flag, void *user);
void om_genbc(const om_mesh *m, int nq2, int nq1, const int *node_dof_top,
              om_bc_fn fn, void *user, int *bc0, int *bc1);

/* ---- gencase operators (MGGenCase.C:1781-1853, 2308-2390, 2452-2566) --- */
/* Block sparsity, first-occurrence column order.  rkind/ckind: 2 = Q2, 1 = Q1.
 * len[nrows+1]; pos sized by caller (query with pos==NULL -> returns nnz).   */
int64_t om_mat_block(const om_mesh *m, int L, int rkind, int ckind, int *len, int *pos);
/* Prolongation block (fine L <- coarse L-1), restriction block (coarse L <- fine L+1) */
int64_t om_prol_block(const om_mesh *m, int L, int kind, int *len, int *pos, double *val);
int64_t om_rest_block(const om_mesh *m, int L, int kind, int *len, int *pos, double *val);

/* Assembled system pattern at level L: sorted CSR over all variable blocks. */
int64_t om_system_pattern(const om_mesh *m, int L, int nq2, int nq1,
                          const int *node_dof_L, int64_t *rowptr, int *colind);
/* Assembled R (coarse L <- fine L+1, masked by bc0%2 of the coarse row) and
 * P (fine L <- coarse L-1) in DoF space, reference column order. */
int64_t om_system_rest(const om_mesh *m, int L, int nq2, int nq1,
                       const int *dof_c, const int *dof_f, const int *dof_top,
                       const int *bc0, int64_t *rowptr, int *colind, double *val);
int64_t om_system_prol(const om_mesh *m, int L, int nq2, int nq1,
                       const int *dof_f, const int *dof_c,
                       int64_t *rowptr, int *colind, double *val);

/* ---- assembly (template_appl/nse/SRC/MGSolverT.C:93-525, MGSolverNS.C:108-717) */
typedef struct {
  double alpha0;     /* _IPrdl*_IRe                                        */
  double rhocp;      /* nondimensional rho*cp (CONST -> 1)                  */
  double dt;         /* pseudo time step (penalty det/dt and time terms)     */
  double time_coeff; /* 1 = reference transient terms, 0 = steady           */
  double source;     /* volume heat source f (the src_value[3] slot)        */
  double upwind;     /* UP_WIND_T                                           */
  double vel_const[3];
} om_heat_params;

typedef struct {
  double IRe;        /* _IRe = mu/(rho L U)                                 */
  double dt;
  double time_coeff;
  double upwind;     /* UP_WIND_NS                                          */
  double IFr;        /* _IFr                                                */
  double dirg[3];    /* gravity direction                                   */
  double src[3];     /* constant volume source                              */
  /* Boussinesq body force (no reference counterpart: MGSolverNS.C:344 keeps the
   * temperature term commented out): RHS of velocity row i of component iv gains
   * int buoy[iv] T phi_i, T interpolated from the nodal field Tn (top-level
   * nodes; NULL = none) */
  double buoy[3];
  const double *Tn;
} om_ns_params;

/* Assemble level L into a CSR with the given (sorted) pattern.  x_old_top is
 * the top-level solution (DoF order of node_dof_top).  k_elem (per element of
 * level L, or NULL = 1) scales alpha_eff / IRe_eff.  vel_top: Q2 nodal
 * velocity [dim*Ntop] (heat only) or NULL for vel_const.  mode 1 = matrix+rhs. */
int om_heat_genmatrhs(const om_mesh *m, int L, int mode, const om_heat_params *p,
                      const int *node_dof_L, const int *node_dof_top,
                      const int *bc0, const int *bc1, const double *x_old_top,
                      const double *k_elem, const double *vel_top,
                      const int64_t *rowptr, const int *colind, double *val, double *rhs);
int om_ns_genmatrhs(const om_mesh *m, int L, int mode, const om_ns_params *p,
                    const int *node_dof_L, const int *node_dof_top,
                    const int *bc0, const int *bc1, const double *x_old_top,
                    const double *nu_elem,
                    const int64_t *rowptr, const int *colind, double *val, double *rhs);
/* One element matrix (row-major, nloc x nloc) + rhs, for unit tests. */
int om_heat_element(const om_mesh *m, int L, int iel, int mode, const om_heat_params *p,
                    const int *node_dof_top, const int *bc0, const int *bc1,
                    const double *x_old_top, double k_e, const double *vel_top,
                    double *KeM, double *FeM, int *dofs, const int *node_dof_L);
int om_ns_element(const om_mesh *m, int L, int iel, int mode, const om_ns_params *p,
                  const int *node_dof_top, const int *bc0, const int *bc1,
                  const double *x_old_top, double nu_e,
                  double *KeM, double *FeM, int *dofs, const int *node_dof_L);

/* ---- linear algebra (contrib/laspack/operats.c, itersolv.c, precond.c) -- */
void om_spmv(int64_t n, const int64_t *rp, const int *ci, const double *v,
             const double *x, double *y);
void om_resid(int64_t n, const int64_t *rp, const int *ci, const double *v,
              const double *b, const double *x, double *r);
/* GMRES(restart) with left ILU(0) preconditioning, LASPACK GMRESIter semantics.
 * Returns iterations; *acc = ||b-Ax||/||b|| at exit. */
int om_gmres_ilu(int64_t n, const int64_t *rp, const int *ci, const double *v,
                 const double *b, double *x, int maxit, int restart, double tol,
                 int use_ilu, double *acc);

/* ---- multigrid (src/MGSolverBase.C:163-289) ----------------------------- */
typedef struct om_hier om_hier;
om_hier *om_hier_new(int nlev);
void om_hier_free(om_hier *h);
void om_hier_set_A(om_hier *h, int L, int64_t n, const int64_t *rp, const int *ci, const double *v);
void om_hier_set_R(om_hier *h, int L, int64_t nr, int64_t nc, const int64_t *rp, const int *ci, const double *v);
void om_hier_set_P(om_hier *h, int L, int64_t nr, int64_t nc, const int64_t *rp, const int *ci, const double *v);
/* smoother: 0 = GMRES(30)+ILU(0) (reference), 1 = Chebyshev-Jacobi, 2 = damped Jacobi
 * coarse:   0 = smoother with 60 its (reference), 1 = dense LU
 * correction: 0 = reference (x += 0.95 P xc only if coarse_rest < rest), 1 = plain x += P xc */
typedef struct {
  int smoother, coarse, correction, nu_pre, nu_post, gamma, max_cycles;
  double eps;            /* MGSolve exit: rest < eps*(1+||b||)             */
  double cheb_lo, cheb_hi; /* Chebyshev interval as fractions of lambda_max  */
  double jacobi_omega;
  int ref_exit;          /* 1 = reference exit test on smoother residual;
                            0 = true relative residual ||b-Ax||/||b|| <= eps */
} om_mg_opts;
/* Returns cycles performed; resid_hist[cycle] = ||b-Ax||/||b|| after the cycle. */
int om_mg_solve(om_hier *h, const om_mg_opts *o, const double *b, double *x,
                double *resid_hist);
/* Power-iteration estimate of lambda_max(D^-1 A) with a fixed start vector. */
double om_lambda_max(int64_t n, const int64_t *rp, const int *ci, const double *v, int iters);

/* Dense LU with partial pivoting; returns 0 on success (singular -> -1). */
int om_dense_lu_solve(int n, double *A, double *b);

#ifdef __cplusplus
}
#endif
#endif
