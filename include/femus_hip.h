/*
 * femus_hip.h -- C ABI of the MI355X (gfx950) FEMuS hot path.
 *
 * One context = one HIP device + one HIP stream; calls on a context are not
 * thread-safe.  Status codes: 0 = OK, <0 = error (see FM_E*); no exceptions
 * cross this boundary.  After a failing call, fm_last_error() returns that
 * call's message (kept per thread until the next failure on the thread).
 * Host arrays are copied in; device memory is owned by
 * the library.  Vectors exchanged with the caller are in the reference DoF
 * order (MGSolDA::_node_dof, src/MGSolverDA.C:880-948).
 *
 * The entry points replace the reference's operator interface:
 *   MGSolBase / MGSolDA   include/MGSolverBase.h:24-389, include/MGSolverDA.h:55-376
 *   SparseMatrixM / SparseMMatrixM / NumericVectorM / LinearSolverM
 *                         contrib/matrix/sparse_matrixM.h:49-186,
 *                         contrib/matrix/numeric_vectorM.h:284-293,
 *                         contrib/matrix/linear_solverM.h:150-162
 * Each declaration names the reference function it stands in for.
 */
#ifndef FEMUS_HIP_H
#define FEMUS_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define FM_OK 0
#define FM_EINVAL -1      /* invalid argument                                 */
#define FM_EHIP -2        /* HIP runtime error                                */
#define FM_ENOCONV -3     /* solver did not reach the tolerance               */
#define FM_ESINGULAR -4   /* singular coarse or local (Vanka) matrix          */
#define FM_ESTATE -5      /* call out of order (e.g. solve before assembly)   */
#define FM_ENOMEM -6      /* device allocation failed                         */

typedef struct fm_ctx fm_ctx;
typedef struct fm_mesh fm_mesh;
typedef struct fm_system fm_system;

/* ---- context ---------------------------------------------------------- */
int fm_ctx_create(int device, fm_ctx **out);
int fm_ctx_destroy(fm_ctx *ctx);
const char *fm_last_error(void);
/* Launch-shape knob of the level products (no counterpart in the reference, whose
 * products are sequential): products over a SELL matrix with fewer than W slices of
 * 64 rows let 2, 4 or 8 wavefronts share a slice (default W = 16384, or the
 * FEMUS_SELL_SPLIT environment variable; 0 = never).  W < 0 only reads the value.
 * Returns the previous W.  Captured V-cycle graphs are keyed on W: a solve after a change captures
 * (and keeps) a graph of the new shape. */
int fm_set_sell_split(int W);
/* Column-offset width knob of the SELL copies (no counterpart in the reference, whose LASPACK rows store
 * full column indices): each slice keeps its columns as 2-bit residuals from base + lane (when every column of
 * a position is base + lane + 0..3, as on the rows of structured grids), as 8-bit or 16-bit offsets from a
 * per-position base when its span allows, else as absolute 32-bit columns.  bits = 2 (default; FEMUS_SELL_P2=0
 * starts at 8, FEMUS_SELL_C8=0 at 16) lets every class be used, 8 never uses the lane-predicted class, 16 no
 * 8-bit offsets either, 32 stores absolute columns in every slice.
 * Applies to matrices set up afterwards (fm_system_init); bits < 0 only reads it.  Returns the previous
 * value, or FM_EINVAL for another width. */
int fm_set_sell_offsets(int bits);
/* Assembly route of GenMatRhs and of the transposed / Newton assemblies: on = 1 (default) writes the element
 * matrices of every element to an element buffer and sums them into the SELL values with one gather per level
 * (deterministic, element order) wherever the buffers of the system stay within a 12 GiB budget; on = 0 adds the
 * element matrices colour by colour straight into the SELL values (also the route of the levels over budget).
 * Returns the previous setting, or FM_EINVAL. */
int fm_set_assembly_gather(int on);

/* ---- mesh hierarchy: gencase GenCase + printMesh -----------------------
 * USER_APPL/gencase/src/MGGenCase.C:63-229 (build_square/build_cube +
 * uniformly_refine), :234-758 (element/node ordering, offsets, node maps).
 * dim = 2 (Quad9) or 3 (Hex27); nx0 = coarse intervals; nlev levels.      */
int fm_mesh_create_box(int dim, const int nx0[3], const double lo[3], const double hi[3],
                       int nlev, fm_mesh **out);
/* A caller's hierarchy in MGMesh's layout (MGMesh::read_c, src/MGMesh.C:493-899; include/MGMesh.h:65-80),
 * one subdomain: n_nodes top-level nodes with coordinates xyz[node + d*n_nodes] (_xyz); level L holds the
 * elements off_el[L] .. off_el[L+1]-1 (_off_el) of el_map (n2 = 9 / 27 top-level node ids per element in
 * libMesh Quad9 / Hex27 local order, _el_map[0]) and el_neighbor (2*dim neighbours per element in libMesh
 * side order as global element ids off_el[L] .. off_el[L+1]-1 of the same level, -1 on the boundary:
 * _el_neighbor[0], the EL_NEIG dataset of gencase, MGGenCase.C:1178-1194); nodes_q2[L] / nodes_q1[L] are the Q2 / Q1 node counts of level L
 * (_NoNodes, _off_nd[0] / [1]: the level's nodes are the first ids of the top numbering); parent[e] and
 * child[e] give each element of levels L >= 1 its level L-1 parent (level-local index) and its index among that parent's 2^dim
 * children (corner bits (c >> d) & 1 along axis d of the parent's reference element: the embedding tables of
 * MGGeomEl, src/MGGeomEl.C:182-690), from which the transfers are built unless the caller hands them in
 * (fm_system_set_transfer).  With parent = child = NULL both are inferred from the nested node sets (the
 * reference's mesh file does not hold them: a child's vertices are its parent's nodes at the child's corner).
 * Elements may be distorted or curved (the Jacobians are general).  FM_EINVAL with a message for an
 * inconsistent hierarchy. */
int fm_mesh_create(int dim, int nlev, int n_nodes, const double *xyz, const int *off_el, const int *el_map,
                   const int *el_neighbor, const int *nodes_q2, const int *nodes_q1, const int *parent,
                   const int *child, fm_mesh **out);
int fm_mesh_destroy(fm_mesh *m);
/* info[0..5] = dim, nlev, n2, n1, nsides, n_nodes_top   (MGMesh::read_c) */
int fm_mesh_info(const fm_mesh *m, int *info);
/* per level: elements, Q2 nodes (_NoNodes[L]), Q1 nodes                    */
int fm_mesh_level(const fm_mesh *m, int L, int *n_el, int *n_q2, int *n_q1);
/* coordinates SoA xyz[node + d*Ntop]  (MGMesh::_xyz, include/MGMesh.h:65-80) */
int fm_mesh_coords(const fm_mesh *m, double *xyz);
/* connectivity / neighbours / parents of level L (MGMesh::get_el_nod_conn
 * src/MGMesh.C:216-236, get_el_neighbor :261-275), neighbours and parents as level-local
 * indices; any pointer may be NULL */
int fm_mesh_conn(const fm_mesh *m, int L, int *conn, int *neigh, int *parent);

/* ---- preprocessor output: the datasets gencase writes ---------------------
 * Kinds: 2 = the level's Q2 nodes, 1 = its Q1 nodes, 0 = its elements (one piecewise-constant DoF each,
 * NDOF_K = 1).  Indices are level-local (gencase's node maps MAP<l> are the identity on the level's nodes).
 * Block (rkind, ckind) of level L's matrix pattern, in gencase's column order (MGGenCase::compute_matrix
 * :1370-1551 with compute_mat_qq / _qk / _kq / _lk / _kl / _kk :1781-2307; the datasets LEN<rc> = len[rows+1]
 * and POS<rc> = pos[*nnz] of Matrix<L>).  With pos = NULL only len and *nnz are filled. */
int fm_gencase_matrix_block(const fm_mesh *m, int L, int rkind, int ckind, int *len, int *pos, int64_t *nnz);
/* Prolongation block of one kind (rest = 0: fine L <- coarse L-1, Prol<L-1>_<L>, compute_prol_qq / _kk
 * :2308-2448) or restriction block (rest = 1: coarse L <- fine L+1, Rest<L+1>_<L>, compute_res_qq / _kk
 * :2452-2631): LEN = len[rows+1], POS = pos[*nnz], VAL = val[*nnz], unmasked.  pos and val both NULL: count
 * only. */
int fm_gencase_transfer_block(const fm_mesh *m, int L, int kind, int rest, int *len, int *pos, double *val,
                              int64_t *nnz);
/* Boundary mesh of level L (printMesh :358-380 and :555-570, print_mesh_h5 /ELEMS/FEM1 :1246-1300): every side
 * without a neighbour, over the level's elements in order and their sides in libMesh order, as its nb2 = 3 / 9
 * nodes in the order of libMesh's build_side; *n_bd gets the count, nodes[n_bd * nb2] may be NULL. */
int fm_gencase_boundary(const fm_mesh *m, int L, int *nodes, int *n_bd);

/* ---- system (one MGSolDA equation) --------------------------------------
 * nq2 Q2 variables (u,v,w or T) followed by nq1 Q1 variables (p):
 * MGSolDA ctor + init_dof (src/MGSolverDA.C:880-948).                      */
int fm_system_create(fm_ctx *ctx, const fm_mesh *m, int nq2, int nq1, fm_system **out);
int fm_system_destroy(fm_system *s);
/* _Dim[L] */
int64_t fm_system_ndof(const fm_system *s, int L);
/* _node_dof[L]: copies (nq2+nq1)*Ntop ints (-1 where absent)               */
int fm_system_node_dof(const fm_system *s, int L, int *node_dof);

/* bc_read callback (MGSolDA::bc_read, template_appl/nse/SRC/User_NS.C):
 * xp[dim] -> bc_neum[nvars] (0 = Dirichlet), bc_flag[nvars]; both preset 1. */
typedef void (*fm_bc_fn)(const double *xp, int *bc_neum, int *bc_flag, void *user);
/* MGSolDA::GenBc / GenBc_loop (src/MGSolverDA.C:696-862)                   */
int fm_system_genbc(fm_system *s, fm_bc_fn fn, void *user);
/* read back bc[0], bc[1] over the top-level DoFs                           */
int fm_system_get_bc(const fm_system *s, int *bc0, int *bc1);

/* MGSolDA::init(L) for all L: sparsity (ReadMatrix, MGGenCase.C:1370-1551),
 * restriction with Dirichlet mask (ReadRest, src/MGSolverDA.C:2236-2444),
 * prolongation (ReadProl, :2011-2230).  Builds device CSR + element maps.  */
int fm_system_init(fm_system *s);
/* Caller-supplied operators for fm_system_init, in the reference DoF numbering of each level (call before
 * fm_system_init; NULL rowptr drops one again).
 * fm_system_set_pattern: MGSolDA::ReadMatrix (src/MGSolverDA.C:1785-2004; SparseMatrixM::update_sparsity_pattern,
 *   contrib/matrix/sparse_matrixM.h:49-186): the pattern of A[L]; it must hold every coupling of the level's
 *   elements (fm_system_init fails with FM_EINVAL naming the missing entry), extra entries stay zero.
 * fm_system_set_transfer: ReadRest / ReadProl (src/MGSolverDA.C:2236-2444, 2011-2230): which 1 = Rst[L] (coarse
 *   L <- fine L+1; its rows are multiplied by the Dirichlet mask as ReadRest does), 2 = Prl[L] (fine L <-
 *   coarse L-1), with values.  Repeated columns in a row are refused.                                        */
int fm_system_set_pattern(fm_system *s, int L, const int64_t *rowptr, const int *colind);
int fm_system_set_transfer(fm_system *s, int which, int L, const int64_t *rowptr, const int *colind,
                           const double *val);
/* Pattern / transfer export for parity checks (host arrays).
 * which: 0 = A[L] (sorted CSR), 1 = Rst[L] (coarse L <- fine L+1), 2 = Prl[L] */
int64_t fm_system_csr_nnz(const fm_system *s, int which, int L);
int fm_system_csr_get(const fm_system *s, int which, int L, int64_t *rowptr, int *colind,
                      double *val);
/* SELL storage of matrix `which` (0 A, 1 R, 2 P) on level L (no counterpart in the reference, whose
 * LASPACK rows are per-row arrays): info = {slices of 64 rows, slots (entries + padding), slots with 8-bit,
 * 16-bit and 32-bit column offsets, nonzeros, slots with 2-bit lane-predicted columns}; the per-class counts
 * are of the stored arrays (8-bit offsets padded to 4 positions, 16-bit to 2, 2-bit to 16 per lane). */
int fm_system_sell_info(const fm_system *s, int which, int L, int64_t info[7]);
/* A caller-assembled A[L]: SparseMatrixM::zero + set(i, j, v) over the whole pattern
 * (contrib/matrix/sparse_matrixM.h:113-116), as a matrix handed to LinearSolverM::solve
 * (contrib/matrix/linear_solverM.h:150-154).  val = the nnz values in the CSR order of fm_system_csr_get
 * (which 0: reference numbering, sorted columns, the level's own or fm_system_set_pattern's pattern).  They
 * replace the level's FP64 / FP32 values and Jacobi diagonal; fm_mg_solve rebuilds its set-up.  The _dev form
 * takes the values in device memory of the context's device (a host pointer is refused).  Complete on return. */
int fm_system_set_matrix(fm_system *s, int L, const double *val);
int fm_system_set_matrix_dev(fm_system *s, int L, const double *val_dev);

/* x_old (top level, DoF order): SparseMatrix fields read by GenMatRhs      */
int fm_system_set_xold(fm_system *s, const double *x_old_top);
int fm_system_get_x(const fm_system *s, int L, double *x);
int fm_system_set_x(fm_system *s, int L, const double *x);
int fm_system_get_rhs(const fm_system *s, int L, double *b);
int fm_system_set_rhs(fm_system *s, int L, const double *b);
/* The same four exchanges with vectors in device memory of the context's device (NumericVectorM data already
 * on the GPU: no host copy, no PCIe); a host pointer is refused with FM_EINVAL.  Complete on return.      */
int fm_system_get_x_dev(const fm_system *s, int L, double *x_dev);
int fm_system_set_x_dev(fm_system *s, int L, const double *x_dev);
int fm_system_get_rhs_dev(const fm_system *s, int L, double *b_dev);
int fm_system_set_rhs_dev(fm_system *s, int L, const double *b_dev);

/* ---- assembly (GenMatRhs) ---------------------------------------------- */
typedef struct {
  double alpha0;      /* _IPrdl*_IRe                                        */
  double rhocp;       /* 1 (CONST)                                          */
  double dt;          /* penalty det/dt and time terms                      */
  double time_coeff;  /* 1 = reference transient terms, 0 = steady          */
  double source;      /* volume source                                      */
  double upwind;      /* UP_WIND_T                                          */
  double vel_const[3];/* advection velocity when no field is set            */
} fm_heat_params;

typedef struct {
  double IRe, dt, time_coeff, upwind, IFr;
  double dirg[3], src[3];
} fm_ns_params;

/* per-element conductivity / viscosity of level L (NULL = 1 / IRe)        */
int fm_system_set_elem_coef(fm_system *s, int L, const double *coef);
/* nodal Q2 advection velocity for heat [dim*Ntop] (NULL = vel_const)       */
int fm_system_set_velocity(fm_system *s, const double *vel_top);
/* Coupling and time stepping without host round trips (no counterpart in the reference, whose
 * systems exchange external fields through host arrays, src/MGSolverDA.C:101-543).
 * fm_system_couple: dst's coupling field from src's top-level solution, gathered on the device (same
 * context and mesh): which 0 = the temperature of a heat system as dst's buoyancy field (coef as in
 * fm_ns_set_buoyancy), 1 = the velocity of a Navier-Stokes system as dst's advection field.
 * fm_system_advance: x_old = x on the top level (MGTimeStep's x_old <- x); *rel_change (if not NULL)
 * = |x - x_old| / |x| before the copy. */
int fm_system_couple(fm_system *dst, const fm_system *src, int which, const double coef[3]);
int fm_system_advance(fm_system *s, double *rel_change);
/* Dirichlet data of the boundary-face penalty RHS, top-level DoFs in the
 * reference numbering.  The reference reads it from x_old
 * (template_appl/nse/SRC/MGSolverNS.C:532-601, MGSolverT.C:401-438); a
 * separate vector lets a parameter act on the boundary data alone.
 * NULL = x_old again.                                                      */
int fm_system_set_bcdata(fm_system *s, const double *g);
/* The linear map g -> boundary-face penalty RHS of the top level (the RHS
 * part of the same face loops), matrix untouched; reference numbering.     */
int fm_heat_bc_rhs(fm_system *s, const fm_heat_params *p, const double *g, double *rhs);
int fm_ns_bc_rhs(fm_system *s, const fm_ns_params *p, const double *g, double *rhs);
/* MGSolT::GenMatRhs (template_appl/nse/SRC/MGSolverT.C:93-525)             */
int fm_heat_genmatrhs(fm_system *s, int L, int mode, const fm_heat_params *p);
/* MGSolNS::GenMatRhs (template_appl/nse/SRC/MGSolverNS.C:108-717)          */
int fm_ns_genmatrhs(fm_system *s, int L, int mode, const fm_ns_params *p);

/* ---- operators on level vectors ----------------------------------------
 * LaspackVectorM::resid (contrib/matrix/laspack_vectorM.C:147-168),
 * Mul_QV (contrib/laspack/operats.c:782-1047),
 * NumericVectorM::matrix_mult with Rst / Prl (laspack_vectorM.C:170-189).   */
int fm_spmv(fm_system *s, int L, const double *x_host, double *y_host);
int fm_residual(fm_system *s, int L);            /* res[L] = b[L] - A[L] x[L] */
int fm_restrict(fm_system *s, int L);            /* b[L-1] = Rst[L-1] res[L]   */
int fm_prolong(fm_system *s, int L);             /* res[L] = Prl[L] x[L-1]     */
int fm_get_res(const fm_system *s, int L, double *r);

/* ---- multigrid solve (MGSolBase::MGSolve / MGStep, src/MGSolverBase.C:163-289) */
#define FM_PREC_FP64 0        /* all-FP64 V-cycle                               */
#define FM_PREC_FP32_IR 1     /* FP32 V-cycle inside FP64 iterative refinement  */
#define FM_PREC_FP32_FGMRES 2 /* FP32 V-cycle preconditioning FP64 FGMRES       */
#define FM_PREC_FP64_FGMRES 3 /* FP64 V-cycle preconditioning FP64 FGMRES       */
#define FM_SMOOTH_CHEBYSHEV 0
#define FM_SMOOTH_JACOBI 1
#define FM_SMOOTH_VANKA 2
/* fourth-kind Chebyshev polynomial in D^-1 A (Lottes, "Optimal polynomial
 * smoothers for multigrid V-cycles", 2023): needs only lambda_max            */
#define FM_SMOOTH_CHEBYSHEV4 3
typedef struct {
  int precision;      /* FM_PREC_*                                          */
  int smoother;       /* FM_SMOOTH_*                                        */
  int nu_pre, nu_post;/* sweeps; doubled below the top level (MGStep)       */
  int correction;     /* 0 = reference (x += 0.95 P xc if coarse_rest < rest), 1 = plain */
  int max_iter;       /* cycles (or outer iterations)                        */
  int fgmres_restart;
  double tol;         /* ||b - A x||_2 / ||b||_2 target (FP64 residual)      */
  double cheb_lo, cheb_hi;   /* Chebyshev interval, fractions of lambda_max  */
  double omega;       /* Jacobi / Vanka damping                              */
  int vanka_colored;  /* 1 = multiplicative over 2^dim element colours       */
  int gamma;          /* coarse-grid visits per level: 1 = V-cycle, 2 = W (MGStep Gamma) */
  int use_graph;      /* 1 = capture each cycle once into a HIP graph and replay it */
  int coarse_mult;    /* sweep multiplier below the top level; 0 = 2, the reference's MGStep rule */
} fm_mg_opts;

typedef struct {
  int iters;          /* cycles / outer iterations done                      */
  int converged;
  double rel_resid;   /* final ||b-Ax||/||b|| (FP64)                         */
  double hist[256];   /* relative residual after each iteration              */
  double seconds;     /* device time of the solve (HIP events)               */
} fm_mg_report;

int fm_mg_setup(fm_system *s, const fm_mg_opts *o);   /* eigen bounds, coarse LU, Vanka inverses */
int fm_mg_solve(fm_system *s, const fm_mg_opts *o, fm_mg_report *rep);   /* x[top] in/out, b[top] */
/* Memory controls of the multigrid data (no counterpart in the reference, whose smoother
 * keeps no patch inverses).  fm_system_set_vanka_store: precisions in which set-up keeps the
 * Vanka patch inverses, bit 0 FP64 (FP64 cycles), bit 1 FP32 (FP32 cycles), default 3; 0 only
 * reads it; returns the previous mask.  A solve whose cycles need a precision left out fails
 * with FM_ESTATE.  fm_mg_release: frees the patch and coarse inverses of the forward
 * (adjoint = 0) or transposed (adjoint = 1) hierarchy, and the flexible-GMRES workspace that solves
 * keep between calls ((2 m + 1) top-level vectors for restart length m); the next solve sets them up
 * again. */
int fm_system_set_vanka_store(fm_system *s, int mask);
int fm_mg_release(fm_system *s, int adjoint);
/* Diagnostic: the FP64 local inverse K_e^-1 the Vanka smoother applies for element e on level L
 * (after fm_mg_setup with FM_SMOOTH_VANKA), row-major nloc x nloc, and the patch DoFs in reference
 * numbering, local order [u, v, (w), p] (the local solves of VankaIter, contrib/laspack/itersolv.c:723-798). */
int fm_vanka_patch(const fm_system *s, int L, int e, int *dofs, double *kinv);

/* ---- discrete adjoint -------------------------------------------------- */
/* Transpose-Jacobian assembly: the GenMatRhs kernels scatter K_e^T, giving
 * A[L]^T of the linear (heat) / Oseen-Picard (NS) system on level L into the
 * adjoint hierarchy.  No reference counterpart (the reference has only the
 * continuous adjoints of template_appl/tbcont/SRC/MGSolverT_ADJ.C:88-412 and
 * template_appl/turb-ck/SRC/MGSolverNS-A.C:356-371).                        */
int fm_heat_genmat_transpose(fm_system *s, int L, const fm_heat_params *p);
int fm_ns_genmat_transpose(fm_system *s, int L, const fm_ns_params *p);
/* values of the assembled A[L]^T (pattern of A[L]) for checks              */
/* Newton Jacobian of the steady Navier-Stokes residual R(U) = A(U) U - b at U = x_old: the GenMatRhs matrix
 * plus the derivatives with respect to u_old of the advection term (u_old . grad) u (velocity blocks gain
 * int phi_i phi_j du_iv/dx_jv) and of the upwind coefficients f_upwind(|u_old|) u_old,k u_old,l
 * (MGSolverNS.C:330-435 differentiated; the reference itself only has the Picard matrix).  trans = 0: into the
 * level matrix (a Newton step solves J dU = b - A(U) U); trans = 1: its transpose into the adjoint matrices
 * (the adjoint of a functional of the steady state).  time_coeff = 0 gives the steady Jacobian. */
int fm_ns_genjac(fm_system *s, int L, int trans, const fm_ns_params *p);
/* y = A^T x with the matrix of the transposed assembly of level L (reference DoF order): the products of a
 * Krylov solve of a coupled adjoint system whose diagonal blocks are these transposed matrices. */
int fm_adjoint_spmv(fm_system *s, int L, const double *x_host, double *y_host);
int fm_get_adjoint_csr(const fm_system *s, int L, double *val);
/* Transposed multigrid solve A^T lambda = g on the assembled hierarchy:
 * restriction and prolongation swap roles (Prl^T, Rst^T), smoothers and the
 * coarse inverse act on A^T; outer loop and precision as in fm_mg_solve.   */
int fm_adjoint_solve(fm_system *s, const fm_mg_opts *o, const double *g_host,
                     double *lambda_host, fm_mg_report *rep);
/* Heat dissipation functional J = alpha0 sum_e k_e int |grad T|^2 over the
 * top elements, its gradient dJ/dT (DoF order) and partials dJ/dk_e.       */
int fm_heat_dissipation(fm_system *s, const fm_heat_params *p, const double *T_host, double *J,
                        double *dJdT_host, double *dJdk_host);
/* Adjoint term of dJ/dk_e: -lambda^T (dA/dk_e) T per top element          */
int fm_heat_grad_coef(fm_system *s, const fm_heat_params *p, const double *lambda_host,
                      const double *x_host, double *grad_host);
/* NS viscous dissipation J = sum_e nu_e int_e |grad u|^2, dJ/dU, dJ/dnu_e  */
int fm_ns_dissipation(fm_system *s, const fm_ns_params *p, const double *U_host, double *J,
                      double *dJdU_host, double *dJdnu_host);
/* Adjoint term of dJ/dnu_e for the Oseen (Picard) system: -lambda^T (dA/dnu_e) U */
int fm_ns_grad_coef(fm_system *s, const fm_ns_params *p, const double *lambda_host,
                    const double *U_host, double *grad_host);
/* Boussinesq coupling (new): body force coef * T in the NS momentum RHS,
 * T_node a nodal temperature field on the top-level nodes [Ntop]; NULL
 * removes it.  (The reference's nse application leaves the temperature
 * dependence of the density commented out, MGSolverNS.C:344.)            */
int fm_ns_set_buoyancy(fm_system *s, const double *T_node, const double coef[3]);
/* Nusselt number of a differentially heated box along `axis`:
 * J = sum_e int_e (u_a T / alpha0 - k_e dT/dx_a) (advection field of
 * fm_system_set_velocity), dJ/dT and the partials dJ/dk_e               */
int fm_heat_nusselt(fm_system *s, const fm_heat_params *p, int axis, const double *T_host, double *J,
                    double *dJdT_host, double *dJdk_host);
/* fm_heat_nusselt without allocating: the same functional, gradient and partials, with the per-element
 * scratch (contributions and conductivity partials) in a fixed device array of the library (16 MB per
 * device, resolved once per device at fm_ctx_create and held until the process ends) instead of device
 * memory allocated per call.  No counterpart in the reference: for callers that cannot allocate device
 * memory (HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  At most 1,048,576
 * top-level elements (FM_EINVAL beyond); calls that overlap on one device race on the shared array.     */
int fm_heat_nusselt_noalloc(fm_system *s, const fm_heat_params *p, int axis, const double *T_host, double *J,
                            double *dJdT_host, double *dJdk_host);
/* Coupling terms of the Boussinesq system (heat system s with parameters p, top level; node-space outputs of
 * Ntop entries, dim * Ntop for mode 1):
 *   mode 0: out_j = scale sum_e int phi_j T                 (dNu/du_axis with scale = 1/kappa; T_host needed)
 *   mode 1: out_(m,j) = (dR_T/du_(m,j))^T lam: scale sum_e int phi_j lam dT/dx_m (advection, scale = rho c_p) plus
 *           the derivative of the upwind diffusion f_upwind(|u|) u_d^2 (MGSolverT.C:290-300; p->upwind, p->rhocp,
 *           p->alpha0, velocity of fm_system_set_velocity); lam on the Dirichlet rows is ignored; T_host, lam_host
 *   mode 2: out_j = scale sum_e int phi_j m                 (Q2 mass matrix times the nodal field m_host)
 * The transposes of the off-diagonal blocks of the coupled Jacobian: the NS body force coef T (fm_ns_set_buoyancy)
 * gives (dR_NS/dT)^T lam_U = -coef M (lam_U on the non-Dirichlet rows of the buoyant component). */
int fm_heat_coupling(fm_system *s, const fm_heat_params *p, int mode, double scale, const double *T_host,
                     const double *lam_host, const double *m_host, double *out_host);
/* Pressure drop J = mean p(x = lo_x) - mean p(x = hi_x) on the top level
 * (face means of the Q1 pressure) and dJ/dU (linear functional)            */
int fm_ns_pressure_drop(fm_system *s, const double *U_host, double *J, double *dJdU_host);

/* ---- timing helpers for benchmarks ------------------------------------- */
/* Run `reps` products of A[L] on device buffers: prec 0 / 1 = y = A x on the SELL copy in FP64 / FP32,
 * 2 / 3 = y = A x on the CSR copy, 4 / 5 = the SELL Chebyshev smoothing step (r -= A d,
 * d' = c1 d + c2 D^-1 r, x += d') in FP64 / FP32; returns the average kernel time in seconds via *sec
 * (HIP events on the library stream).  Overwrites the level's work vectors. */
int fm_bench_spmv(fm_system *s, int L, int prec, int reps, double *sec);
int fm_sync(fm_ctx *ctx);
/* STREAM copy of `bytes` between two fresh device buffers (read + write),
 * average kernel seconds over reps: the achievable-bandwidth yardstick.    */
int fm_bench_stream(fm_ctx *ctx, int64_t bytes, int reps, double *sec);

#ifdef __cplusplus
}
#endif
#endif
