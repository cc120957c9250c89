/*
 * femus_oracle_mesh.c -- TEST INFRASTRUCTURE (oracle).  FE tables, mesh
 * hierarchy + node/element ordering, DoF map, boundary conditions and the
 * gencase operator patterns (sparsity, prolongation, restriction).
 *
 * Restates:
 *   src/MGFE.C:195-924                    Q2/Q1 shapes, 3-point Gauss rule
 *   src/MGGeomEl.C:160-690                 _surf_top, embedding matrices
 *   USER_APPL/gencase/src/MGGenCase.C:234-758  element/node ordering
 *   USER_APPL/gencase/src/MGGenCase.C:1781-1853 compute_mat_qq
 *   USER_APPL/gencase/src/MGGenCase.C:2308-2390 compute_prol_qq
 *   USER_APPL/gencase/src/MGGenCase.C:2452-2566 compute_res_qq
 *   src/MGSolverDA.C:880-948 (init_dof), 786-862 (GenBc_loop),
 *   2236-2444 (ReadRest mask), 2011-2230 (ReadProl)
 *
 * The libMesh node/element ids that gencase sorts by are emulated:
 * build_square/build_cube element order (i fastest), uniform refinement
 * appending children parent-by-parent, and node ids assigned by first
 * encounter in that element order (libMesh renumbering).  libMesh 0.9.4 is
 * not vendored, so this emulation is "parity unpinned"; it is pinned by
 * invariants in the tests.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "femus_oracle.h"
#include "femus_oracle_internal.h"

/* ---- reference-element node coordinates (libMesh Quad9/Hex27 order) ---- */
const int om_quad9_ref[9][2] = {{-1,-1},{1,-1},{1,1},{-1,1},{0,-1},{1,0},{0,1},{-1,0},{0,0}};
const int om_hex27_ref[27][3] = {
  {-1,-1,-1},{1,-1,-1},{1,1,-1},{-1,1,-1},{-1,-1,1},{1,-1,1},{1,1,1},{-1,1,1},
  {0,-1,-1},{1,0,-1},{0,1,-1},{-1,0,-1},{-1,-1,0},{1,-1,0},{1,1,0},{-1,1,0},
  {0,-1,1},{1,0,1},{0,1,1},{-1,0,1},{0,0,-1},{0,-1,0},{1,0,0},{0,1,0},{-1,0,0},{0,0,1},{0,0,0}};
/* face -> local nodes (src/MGGeomEl.C:172-177, 310-317) */
const int om_surf2d[4][3] = {{0,1,4},{1,2,5},{2,3,6},{3,0,7}};
const int om_surf3d[6][9] = {
  {0,1,2,3,8,9,10,11,20},{0,1,5,4,8,13,16,12,21},{1,2,6,5,9,14,17,13,22},
  {2,3,7,6,10,15,18,14,23},{3,0,4,7,11,12,19,15,24},{4,5,6,7,16,17,18,19,25}};
/* side -> (axis, direction) for the libMesh side numbering */
static const int side2d[4][2] = {{1,-1},{0,1},{1,1},{0,-1}};
static const int side3d[6][2] = {{2,-1},{1,-1},{0,1},{1,1},{0,-1},{2,1}};

int om_ref_coord(int dim, int k, int d) { return dim == 2 ? om_quad9_ref[k][d] : om_hex27_ref[k][d]; }
int om_face_node(int dim, int s, int k) { return dim == 2 ? om_surf2d[s][k] : om_surf3d[s][k]; }
void om_side_dir(int dim, int s, int *axis, int *sgn) {
  if (dim == 2) { *axis = side2d[s][0]; *sgn = side2d[s][1]; }
  else { *axis = side3d[s][0]; *sgn = side3d[s][1]; }
}

/* 1D Lagrange factors (src/MGFE.C:61-71) */
static double l2(int node, double x) { return node < 0 ? .5*x*(x-1.) : (node > 0 ? .5*x*(x+1.) : 1.-x*x); }
static double dl2(int node, double x) { return node < 0 ? x-.5 : (node > 0 ? x+.5 : -2.*x); }
static double l1(int node, double x) { return node < 0 ? .5*(1.-x) : .5*(1.+x); }
static double dl1(int node, double x) { return node < 0 ? -.5 : .5; }

static int nshape(int kind, int dim) {
  if (kind == 2) return dim == 1 ? 3 : (dim == 2 ? 9 : 27);
  return dim == 1 ? 2 : (dim == 2 ? 4 : 8);
}
/* node -> reference coordinate in 1D index order {-1,+1,0} (MGFE.C 1D order) */
static int ref1d(int k) { return k == 0 ? -1 : (k == 1 ? 1 : 0); }
static int refc(int kind, int dim, int k, int d) {
  if (dim == 1) return ref1d(k);
  return om_ref_coord(dim, k, d);
}

/* Shape values at a reference point (used for tables and embeddings). */
double om_shape(int kind, int dim, int k, const double *xi) {
  double v = 1.;
  for (int d = 0; d < dim; d++) {
    int c = refc(kind, dim, k, d);
    v *= (kind == 2) ? l2(c, xi[d]) : l1(c, xi[d]);
  }
  return v;
}
double om_dshape(int kind, int dim, int k, int dir, const double *xi) {
  double v = 1.;
  for (int d = 0; d < dim; d++) {
    int c = refc(kind, dim, k, d);
    if (d == dir) v *= (kind == 2) ? dl2(c, xi[d]) : dl1(c, xi[d]);
    else v *= (kind == 2) ? l2(c, xi[d]) : l1(c, xi[d]);
  }
  return v;
}

/* Gauss point qp -> reference coordinates (MGFE.C:19-36: x slowest). */
void om_gauss_point(int dim, int qp, double *xi, double *w) {
  const double a = -sqrt(3. / 5.);
  const double x1[3] = {a, 0., -a};
  const double w1[3] = {5. / 9., 8. / 9., 5. / 9.};
  int idx[3] = {0, 0, 0};
  if (dim == 1) idx[0] = qp;
  else if (dim == 2) { idx[0] = qp / 3; idx[1] = qp % 3; }
  else { idx[0] = qp / 9; idx[1] = (qp / 3) % 3; idx[2] = qp % 3; }
  double ww = 1.;
  for (int d = 0; d < dim; d++) { xi[d] = x1[idx[d]]; ww *= w1[idx[d]]; }
  if (w) *w = ww;
}

int om_fe_tables(int kind, int dim, double *phi, double *dphi, double *w, int *ng_out) {
  int ns = nshape(kind, dim);
  int ng = dim == 1 ? 3 : (dim == 2 ? 9 : 27);
  for (int qp = 0; qp < ng; qp++) {
    double xi[3], ww;
    om_gauss_point(dim, qp, xi, &ww);
    if (w) w[qp] = ww;
    for (int s = 0; s < ns; s++) {
      if (phi) phi[s * ng + qp] = om_shape(kind, dim, s, xi);
      if (dphi) for (int d = 0; d < dim; d++) dphi[(s * ng + qp) + d * ns * ng] = om_dshape(kind, dim, s, d, xi);
    }
  }
  if (ng_out) *ng_out = ng;
  return ns;
}

int om_embedding(int kind, int dim, double *E) {
  int ns = nshape(kind, dim);
  int nch = 1 << dim;
  for (int c = 0; c < nch; c++)
    for (int i = 0; i < ns; i++) {
      double xp[3];
      for (int d = 0; d < dim; d++) {
        int bit = (c >> d) & 1;
        double lo = bit ? 0. : -1.;
        xp[d] = lo + (refc(kind, dim, i, d) + 1) * 0.5;
      }
      for (int j = 0; j < ns; j++) E[(c * ns + i) * ns + j] = om_shape(kind, dim, j, xp);
    }
  return ns;
}

/* ========================================================================= */
/* Mesh hierarchy                                                            */
/* ========================================================================= */
om_mesh *om_mesh_build(int dim, const int nx0[3], const double lo[3], const double hi[3], int nlev) {
  om_mesh *m = (om_mesh *)calloc(1, sizeof(om_mesh));
  m->dim = dim; m->nlev = nlev;
  m->n2 = dim == 2 ? 9 : 27; m->n1 = dim == 2 ? 4 : 8;
  m->nsides = 2 * dim; m->nb2 = dim == 2 ? 3 : 9; m->nch = 1 << dim;
  int64_t gtot = 1;
  for (int d = 0; d < 3; d++) {
    m->nx0[d] = d < dim ? nx0[d] : 1;
    m->lo[d] = d < dim ? lo[d] : 0.; m->hi[d] = d < dim ? hi[d] : 0.;
    m->G[d] = d < dim ? 2 * nx0[d] * (1 << (nlev - 1)) + 1 : 1;
    gtot *= m->G[d];
  }
  m->nel = (int *)calloc(nlev, sizeof(int));
  m->corner = (int **)calloc(nlev, sizeof(int *));
  m->parent = (int **)calloc(nlev, sizeof(int *));
  m->conn = (int **)calloc(nlev, sizeof(int *));
  m->neigh = (int **)calloc(nlev, sizeof(int *));
  /* level 0: build_square / build_cube order (k, j, i with i fastest) */
  int n0 = m->nx0[0] * m->nx0[1] * m->nx0[2];
  int S0 = 1 << nlev;
  m->nel[0] = n0;
  m->corner[0] = (int *)malloc(sizeof(int) * 3 * n0);
  m->parent[0] = (int *)malloc(sizeof(int) * n0);
  int e = 0;
  for (int k = 0; k < m->nx0[2]; k++)
    for (int j = 0; j < m->nx0[1]; j++)
      for (int i = 0; i < m->nx0[0]; i++, e++) {
        m->corner[0][3 * e] = i * S0; m->corner[0][3 * e + 1] = j * S0; m->corner[0][3 * e + 2] = k * S0;
        m->parent[0][e] = -1;
      }
  /* uniform refinement: children appended parent by parent */
  for (int L = 1; L < nlev; L++) {
    int np = m->nel[L - 1], nc = np * m->nch, S = 1 << (nlev - L);
    m->nel[L] = nc;
    m->corner[L] = (int *)malloc(sizeof(int) * 3 * nc);
    m->parent[L] = (int *)malloc(sizeof(int) * nc);
    for (int p = 0; p < np; p++)
      for (int c = 0; c < m->nch; c++) {
        int ce = p * m->nch + c;
        for (int d = 0; d < 3; d++)
          m->corner[L][3 * ce + d] = m->corner[L - 1][3 * p + d] + (d < dim ? ((c >> d) & 1) * S : 0);
        m->parent[L][ce] = p;
      }
  }
  /* libMesh-style ids: first encounter over the element list (levels in order) */
  int *lid = (int *)malloc(sizeof(int) * gtot);
  int *glev = (int *)malloc(sizeof(int) * gtot);
  int *glevp = (int *)malloc(sizeof(int) * gtot);
  for (int64_t g = 0; g < gtot; g++) { lid[g] = -1; glev[g] = nlev; glevp[g] = nlev; }
  int cnt = 0;
  for (int L = 0; L < nlev; L++) {
    int S = 1 << (nlev - L);
    for (e = 0; e < m->nel[L]; e++)
      for (int k = 0; k < m->n2; k++) {
        int64_t g = 0, stride = 1;
        for (int d = 0; d < dim; d++) {
          int I = m->corner[L][3 * e + d] + (om_ref_coord(dim, k, d) + 1) * S / 2;
          g += I * stride; stride *= m->G[d];
        }
        if (lid[g] < 0) { lid[g] = cnt++; glev[g] = L; }
        if (k < m->n1 && glevp[g] > L) glevp[g] = L;
      }
  }
  m->Ntop = cnt;
  /* gencase node sort: key = levp + (nlev+1)*lev, ties by libMesh id */
  int *g_by_lid = (int *)malloc(sizeof(int) * cnt);
  for (int64_t g = 0; g < gtot; g++) if (lid[g] >= 0) g_by_lid[lid[g]] = (int)g;
  int nkeys = (nlev + 1) * (nlev + 1);
  int *kcount = (int *)calloc(nkeys + 1, sizeof(int));
  for (int i = 0; i < cnt; i++) { int g = g_by_lid[i]; kcount[glevp[g] + (nlev + 1) * glev[g] + 1]++; }
  for (int k = 0; k < nkeys; k++) kcount[k + 1] += kcount[k];
  m->gnew = (int *)malloc(sizeof(int) * gtot);
  for (int64_t g = 0; g < gtot; g++) m->gnew[g] = -1;
  m->node_lid = (int *)malloc(sizeof(int) * cnt);
  m->node_lev = (int *)malloc(sizeof(int) * cnt);
  m->node_levp = (int *)malloc(sizeof(int) * cnt);
  m->xyz = (double *)malloc(sizeof(double) * dim * cnt);
  for (int i = 0; i < cnt; i++) {
    int g = g_by_lid[i];
    int key = glevp[g] + (nlev + 1) * glev[g];
    int nn = kcount[key]++;
    m->gnew[g] = nn;
    m->node_lid[nn] = i; m->node_lev[nn] = glev[g]; m->node_levp[nn] = glevp[g];
    int I[3] = {g % m->G[0], (g / m->G[0]) % m->G[1], g / (m->G[0] * m->G[1])};
    for (int d = 0; d < dim; d++)
      m->xyz[nn + d * cnt] = m->lo[d] + (m->hi[d] - m->lo[d]) * (double)I[d] / (double)(m->G[d] - 1);
  }
  m->nq2 = (int *)calloc(nlev, sizeof(int));
  m->nq1 = (int *)calloc(nlev, sizeof(int));
  for (int i = 0; i < cnt; i++)
    for (int L = 0; L < nlev; L++) {
      if (m->node_lev[i] <= L) m->nq2[L]++;
      if (m->node_levp[i] <= L) m->nq1[L]++;
    }
  /* connectivity and neighbours per level */
  for (int L = 0; L < nlev; L++) {
    int S = 1 << (nlev - L), ne = m->nel[L];
    m->conn[L] = (int *)malloc(sizeof(int) * ne * m->n2);
    m->neigh[L] = (int *)malloc(sizeof(int) * ne * m->nsides);
    int nE[3] = {1, 1, 1};
    for (int d = 0; d < dim; d++) nE[d] = m->nx0[d] << L;
    int *eidx = (int *)malloc(sizeof(int) * nE[0] * nE[1] * nE[2]);
    for (e = 0; e < ne; e++) {
      int ci = m->corner[L][3 * e] / S, cj = m->corner[L][3 * e + 1] / S, ck = dim == 3 ? m->corner[L][3 * e + 2] / S : 0;
      eidx[ci + nE[0] * (cj + nE[1] * ck)] = e;
      for (int k = 0; k < m->n2; k++) {
        int64_t g = 0, stride = 1;
        for (int d = 0; d < dim; d++) {
          int I = m->corner[L][3 * e + d] + (om_ref_coord(dim, k, d) + 1) * S / 2;
          g += I * stride; stride *= m->G[d];
        }
        m->conn[L][e * m->n2 + k] = m->gnew[g];
      }
    }
    for (e = 0; e < ne; e++) {
      int c[3] = {m->corner[L][3 * e] / S, m->corner[L][3 * e + 1] / S, dim == 3 ? m->corner[L][3 * e + 2] / S : 0};
      for (int s = 0; s < m->nsides; s++) {
        int ax, sg; om_side_dir(dim, s, &ax, &sg);
        int cc[3] = {c[0], c[1], c[2]};
        cc[ax] += sg;
        int nb = -1;
        if (cc[ax] >= 0 && cc[ax] < nE[ax]) nb = eidx[cc[0] + nE[0] * (cc[1] + nE[1] * cc[2])];
        m->neigh[L][e * m->nsides + s] = nb;
      }
    }
    free(eidx);
  }
  free(lid); free(glev); free(glevp); free(g_by_lid); free(kcount);
  return m;
}

void om_mesh_free(om_mesh *m) {
  if (!m) return;
  for (int L = 0; L < m->nlev; L++) { free(m->corner[L]); free(m->parent[L]); free(m->conn[L]); free(m->neigh[L]); }
  free(m->corner); free(m->parent); free(m->conn); free(m->neigh); free(m->nel);
  free(m->gnew); free(m->node_lid); free(m->node_lev); free(m->node_levp); free(m->xyz);
  free(m->nq2); free(m->nq1); free(m);
}

void om_mesh_info(const om_mesh *m, int *info) {
  info[0] = m->dim; info[1] = m->nlev; info[2] = m->n2; info[3] = m->n1; info[4] = m->nsides; info[5] = m->Ntop;
}
void om_mesh_level(const om_mesh *m, int L, int *n_el, int *nq2, int *nq1) {
  *n_el = m->nel[L]; *nq2 = m->nq2[L]; *nq1 = m->nq1[L];
}
void om_mesh_coords(const om_mesh *m, double *xyz) { memcpy(xyz, m->xyz, sizeof(double) * m->dim * m->Ntop); }
/* moved nodes (a distorted or curved hierarchy with the same connectivity): the element routines read the
 * geometry from xyz only, so every level follows */
void om_mesh_set_coords(om_mesh *m, const double *xyz) { memcpy(m->xyz, xyz, sizeof(double) * m->dim * m->Ntop); }
void om_mesh_conn(const om_mesh *m, int L, int *conn, int *neigh, int *parent) {
  if (conn) memcpy(conn, m->conn[L], sizeof(int) * m->nel[L] * m->n2);
  if (neigh) memcpy(neigh, m->neigh[L], sizeof(int) * m->nel[L] * m->nsides);
  if (parent) memcpy(parent, m->parent[L], sizeof(int) * m->nel[L]);
}
void om_mesh_node_keys(const om_mesh *m, int *libmesh_id, int *lev, int *levp) {
  memcpy(libmesh_id, m->node_lid, sizeof(int) * m->Ntop);
  memcpy(lev, m->node_lev, sizeof(int) * m->Ntop);
  memcpy(levp, m->node_levp, sizeof(int) * m->Ntop);
}

/* ========================================================================= */
/* DoF map and boundary conditions                                           */
/* ========================================================================= */
int om_init_dof(const om_mesh *m, int L, int nq2, int nq1, int *node_dof) {
  int N = m->Ntop, nv = nq2 + nq1, count = 0;
  for (int i = 0; i < N * nv; i++) node_dof[i] = -1;
  for (int iv = 0; iv < nq2; iv++)
    for (int k = 0; k < m->nq2[L]; k++) node_dof[k + iv * N] = count++;
  for (int iv = 0; iv < nq1; iv++)
    for (int k = 0; k < m->nq1[L]; k++) node_dof[k + (nq2 + iv) * N] = count++;
  return count;
}

void om_genbc(const om_mesh *m, int nq2, int nq1, const int *node_dof_top,
              om_bc_fn fn, void *user, int *bc0, int *bc1) {
  int top = m->nlev - 1, N = m->Ntop, nv = nq2 + nq1, dim = m->dim;
  int dimtop = nq2 * m->nq2[top] + nq1 * m->nq1[top];
  for (int i = 0; i < dimtop; i++) { bc0[i] = 1; bc1[i] = 1; }
  int neum[8], flag[8];
  /* volume pass: bc_intern_read default (all 1) -- src/MGSolverDA.C:867-878 */
  /* boundary pass over top-level boundary faces -- src/MGSolverDA.C:786-862 */
  int nb1 = dim == 2 ? 2 : 4;
  for (int e = 0; e < m->nel[top]; e++)
    for (int s = 0; s < m->nsides; s++) {
      if (m->neigh[top][e * m->nsides + s] != -1) continue;
      for (int i = 0; i < m->nb2; i++) {
        int k = m->conn[top][e * m->n2 + om_face_node(dim, s, i)];
        double xp[3];
        for (int d = 0; d < dim; d++) xp[d] = m->xyz[k + d * N];
        for (int iv = 0; iv < nv; iv++) { neum[iv] = 1; flag[iv] = 1; }
        fn(xp, neum, flag, user);
        for (int iv = 0; iv < nq2; iv++) {
          int kd = node_dof_top[k + iv * N];
          bc0[kd] = neum[iv]; bc1[kd] = flag[iv];
        }
        if (i < nb1)
          for (int iv = nq2; iv < nv; iv++) {
            int kd = node_dof_top[k + iv * N];
            bc0[kd] = neum[iv]; bc1[kd] = flag[iv];
          }
      }
    }
}

/* ========================================================================= */
/* gencase operators                                                          */
/* ========================================================================= */
int64_t om_mat_block(const om_mesh *m, int L, int rkind, int ckind, int *len, int *pos) {
  int nr = rkind == 2 ? m->n2 : m->n1, nc = ckind == 2 ? m->n2 : m->n1;
  int nrows = rkind == 2 ? m->nq2[L] : m->nq1[L];
  int ncols_tot = ckind == 2 ? m->nq2[L] : m->nq1[L];
  int ne = m->nel[L];
  const int *conn = m->conn[L];
  /* row lists of raw appended columns in element order (MGGenCase.C:1806-1822) */
  int64_t *off = (int64_t *)calloc(nrows + 1, sizeof(int64_t));
  for (int e = 0; e < ne; e++) for (int kk = 0; kk < nr; kk++) off[conn[e * m->n2 + kk] + 1] += nc;
  for (int i = 0; i < nrows; i++) off[i + 1] += off[i];
  int *raw = (int *)malloc(sizeof(int) * (off[nrows] + 1));
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (nrows + 1));
  memcpy(fill, off, sizeof(int64_t) * (nrows + 1));
  for (int e = 0; e < ne; e++)
    for (int kk = 0; kk < nr; kk++) {
      int row = conn[e * m->n2 + kk];
      for (int kj = 0; kj < nc; kj++) raw[fill[row]++] = conn[e * m->n2 + kj];
    }
  /* compress: first occurrence (MGGenCase.C:1826-1853) */
  char *aux = (char *)calloc(ncols_tot > 0 ? ncols_tot : 1, 1);
  int64_t count = 0;
  if (len) len[0] = 0;
  for (int i = 0; i < nrows; i++) {
    for (int64_t q = off[i]; q < off[i + 1]; q++) aux[raw[q]] = 1;
    for (int64_t q = off[i]; q < off[i + 1]; q++)
      if (aux[raw[q]]) { aux[raw[q]] = 0; if (pos) pos[count] = raw[q]; count++; }
    if (len) len[i + 1] = (int)count;
  }
  free(aux); free(raw); free(fill); free(off);
  return count;
}

int64_t om_prol_block(const om_mesh *m, int L, int kind, int *len, int *pos, double *val) {
  int ns = kind == 2 ? m->n2 : m->n1, NF = m->n2;
  int nrows = kind == 2 ? m->nq2[L] : m->nq1[L];
  double *E = (double *)malloc(sizeof(double) * m->nch * ns * ns);
  om_embedding(kind, m->dim, E);
  int *slot = (int *)malloc(sizeof(int) * (size_t)nrows * NF);
  double *sval = (double *)malloc(sizeof(double) * (size_t)nrows * NF);
  for (int64_t i = 0; i < (int64_t)nrows * NF; i++) { slot[i] = -1; sval[i] = 0.; }
  for (int e = 0; e < m->nel[L]; e++) {
    int par = m->parent[L][e], ch = e - par * m->nch; /* child id: position in parent's list */
    for (int kk = 0; kk < ns; kk++) {
      int row = m->conn[L][e * m->n2 + kk];
      for (int kj = 0; kj < ns; kj++) {
        slot[(int64_t)row * NF + kj] = m->conn[L - 1][par * m->n2 + kj];
        sval[(int64_t)row * NF + kj] = E[(ch * ns + kk) * ns + kj];
      }
    }
  }
  int64_t count = 0;
  if (len) len[0] = 0;
  for (int i = 0; i < nrows; i++) {
    for (int kj = 0; kj < NF; kj++) {
      int64_t q = (int64_t)i * NF + kj;
      if (fabs(sval[q]) > 1.e-8) { if (pos) { pos[count] = slot[q]; val[count] = sval[q]; } count++; }
    }
    if (len) len[i + 1] = (int)count;
  }
  free(E); free(slot); free(sval);
  return count;
}

int64_t om_rest_block(const om_mesh *m, int L, int kind, int *len, int *pos, double *val) {
  int ns = kind == 2 ? m->n2 : m->n1;
  int nrows = kind == 2 ? m->nq2[L] : m->nq1[L];
  int maxr = 125 * 2; /* >= max row length (5^3 Q2 vertex rows) */
  double *E = (double *)malloc(sizeof(double) * m->nch * ns * ns);
  om_embedding(kind, m->dim, E);
  int *rp = (int *)malloc(sizeof(int) * (size_t)nrows * maxr);
  double *rv = (double *)malloc(sizeof(double) * (size_t)nrows * maxr);
  int *mem = (int *)calloc(nrows, sizeof(int));
  for (int e = 0; e < m->nel[L]; e++)
    for (int ich = 0; ich < m->nch; ich++) {
      int ch = e * m->nch + ich;  /* children of e at level L+1 */
      for (int kk = 0; kk < ns; kk++) {
        int irow = m->conn[L][e * m->n2 + kk];
        for (int kj = 0; kj < ns; kj++) {
          int col = m->conn[L + 1][ch * m->n2 + kj];
          double v = E[(ich * ns + kj) * ns + kk];
          if (fabs(v) > 1.e-6) {
            int found = 0;
            for (int ks = 0; ks < mem[irow]; ks++) if (rp[(int64_t)irow * maxr + ks] == col) found = 1;
            if (!found) {
              rp[(int64_t)irow * maxr + mem[irow]] = col;
              rv[(int64_t)irow * maxr + mem[irow]] = v;
              mem[irow]++;
            }
          }
        }
      }
    }
  int64_t count = 0;
  if (len) len[0] = 0;
  for (int i = 0; i < nrows; i++) {
    for (int k = 0; k < mem[i]; k++) {
      if (pos) { pos[count] = rp[(int64_t)i * maxr + k]; val[count] = rv[(int64_t)i * maxr + k]; }
      count++;
    }
    if (len) len[i + 1] = (int)count;
  }
  free(E); free(rp); free(rv); free(mem);
  return count;
}

static int cmp_int(const void *a, const void *b) { int x = *(const int *)a, y = *(const int *)b; return x < y ? -1 : x > y; }

/* element DoF list at level L: [q2 vars x n2 nodes, q1 vars x n1 vertices] */
int om_elem_dofs(const om_mesh *m, int L, int e, int nq2, int nq1, const int *node_dof, int *dofs) {
  int N = m->Ntop, c = 0;
  for (int iv = 0; iv < nq2; iv++)
    for (int k = 0; k < m->n2; k++) dofs[c++] = node_dof[m->conn[L][e * m->n2 + k] + iv * N];
  for (int iv = 0; iv < nq1; iv++)
    for (int k = 0; k < m->n1; k++) dofs[c++] = node_dof[m->conn[L][e * m->n2 + k] + (nq2 + iv) * N];
  return c;
}

int64_t om_system_pattern(const om_mesh *m, int L, int nq2, int nq1, const int *node_dof_L,
                          int64_t *rowptr, int *colind) {
  int nloc = nq2 * m->n2 + nq1 * m->n1;
  int n = nq2 * m->nq2[L] + nq1 * m->nq1[L];
  int ne = m->nel[L];
  int *dofs = (int *)malloc(sizeof(int) * nloc);
  int64_t *off = (int64_t *)calloc(n + 1, sizeof(int64_t));
  for (int e = 0; e < ne; e++) {
    om_elem_dofs(m, L, e, nq2, nq1, node_dof_L, dofs);
    for (int a = 0; a < nloc; a++) off[dofs[a] + 1] += nloc;
  }
  for (int i = 0; i < n; i++) off[i + 1] += off[i];
  int *raw = (int *)malloc(sizeof(int) * (off[n] + 1));
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
  memcpy(fill, off, sizeof(int64_t) * (n + 1));
  for (int e = 0; e < ne; e++) {
    om_elem_dofs(m, L, e, nq2, nq1, node_dof_L, dofs);
    for (int a = 0; a < nloc; a++) for (int b = 0; b < nloc; b++) raw[fill[dofs[a]]++] = dofs[b];
  }
  int64_t count = 0;
  if (rowptr) rowptr[0] = 0;
  for (int i = 0; i < n; i++) {
    int64_t l = off[i + 1] - off[i];
    qsort(raw + off[i], l, sizeof(int), cmp_int);
    int last = -1;
    for (int64_t q = off[i]; q < off[i + 1]; q++)
      if (raw[q] != last) { last = raw[q]; if (colind) colind[count] = last; count++; }
    if (rowptr) rowptr[i + 1] = count;
  }
  free(dofs); free(off); free(raw); free(fill);
  return count;
}

/* Assemble R or P blocks into DoF space (src/MGSolverDA.C:2400-2430, 2180-2215) */
static int64_t system_transfer(const om_mesh *m, int L, int nq2, int nq1, int is_rest,
                               const int *dof_row, const int *dof_col, const int *dof_top,
                               const int *bc0, int64_t *rowptr, int *colind, double *val) {
  int N = m->Ntop;
  int Lr = L;
  int nrows = nq2 * m->nq2[Lr] + nq1 * m->nq1[Lr];
  int *blen[2] = {0, 0}, *bpos[2] = {0, 0};
  double *bval[2] = {0, 0};
  int64_t bnnz[2] = {0, 0};
  for (int t = 0; t < 2; t++) {
    int kind = t == 0 ? 2 : 1;
    if ((t == 0 && nq2 == 0) || (t == 1 && nq1 == 0)) continue;
    int nr = kind == 2 ? m->nq2[Lr] : m->nq1[Lr];
    blen[t] = (int *)malloc(sizeof(int) * (nr + 1));
    bnnz[t] = is_rest ? om_rest_block(m, L, kind, blen[t], NULL, NULL) : om_prol_block(m, L, kind, blen[t], NULL, NULL);
    bpos[t] = (int *)malloc(sizeof(int) * (bnnz[t] + 1));
    bval[t] = (double *)malloc(sizeof(double) * (bnnz[t] + 1));
    if (is_rest) om_rest_block(m, L, kind, blen[t], bpos[t], bval[t]);
    else om_prol_block(m, L, kind, blen[t], bpos[t], bval[t]);
  }
  /* row lengths */
  int64_t *rl = (int64_t *)calloc(nrows + 1, sizeof(int64_t));
  for (int iv = 0; iv < nq2 + nq1; iv++) {
    int t = iv < nq2 ? 0 : 1;
    int nr = t == 0 ? m->nq2[Lr] : m->nq1[Lr];
    for (int i = 0; i < nr; i++) rl[dof_row[i + iv * N] + 1] = blen[t][i + 1] - blen[t][i];
  }
  for (int i = 0; i < nrows; i++) rl[i + 1] += rl[i];
  if (rowptr) memcpy(rowptr, rl, sizeof(int64_t) * (nrows + 1));
  int64_t total = rl[nrows];
  if (colind) {
    for (int iv = 0; iv < nq2 + nq1; iv++) {
      int t = iv < nq2 ? 0 : 1;
      int nr = t == 0 ? m->nq2[Lr] : m->nq1[Lr];
      for (int i = 0; i < nr; i++) {
        int row = dof_row[i + iv * N];
        double mask = 1.;
        if (is_rest) mask = (double)(bc0[dof_top[i + iv * N]] % 2);
        int64_t q0 = rl[row];
        for (int j = blen[t][i]; j < blen[t][i + 1]; j++) {
          colind[q0 + j - blen[t][i]] = dof_col[bpos[t][j] + iv * N];
          val[q0 + j - blen[t][i]] = mask * bval[t][j];
        }
      }
    }
  }
  for (int t = 0; t < 2; t++) { free(blen[t]); free(bpos[t]); free(bval[t]); }
  free(rl);
  return total;
}

int64_t om_system_rest(const om_mesh *m, int L, int nq2, int nq1, const int *dof_c, const int *dof_f,
                       const int *dof_top, const int *bc0, int64_t *rowptr, int *colind, double *val) {
  return system_transfer(m, L, nq2, nq1, 1, dof_c, dof_f, dof_top, bc0, rowptr, colind, val);
}
int64_t om_system_prol(const om_mesh *m, int L, int nq2, int nq1, const int *dof_f, const int *dof_c,
                       int64_t *rowptr, int *colind, double *val) {
  return system_transfer(m, L, nq2, nq1, 0, dof_f, dof_c, NULL, NULL, rowptr, colind, val);
}
