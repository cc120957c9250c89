// gencase.cpp -- host construction of the multigrid mesh hierarchy and of the
// per-level operators' patterns for the MI355X FEMuS hot path.
//
// Mirrors the reference preprocessor and DoF setup:
//   USER_APPL/gencase/src/MGGenCase.C:98-181   build_square/build_cube + uniformly_refine
//   USER_APPL/gencase/src/MGGenCase.C:534-620  element order (lev+nlev*subdom, libMesh id),
//                                              node order (var, levp, lev, subdom, libMesh id)
//   USER_APPL/gencase/src/MGGenCase.C:1781-1853 sparsity (element couplings)
//   USER_APPL/gencase/src/MGGenCase.C:2308-2390 prolongation (last fine element wins)
//   USER_APPL/gencase/src/MGGenCase.C:2452-2566 restriction (first found, |E|>1e-6)
//   src/MGSolverDA.C:880-948 init_dof; :2400-2430 ReadRest Dirichlet mask
// libMesh ids are emulated: level-0 elements k,j,i (i fastest), children
// appended parent by parent, node ids by first encounter over that element
// list.  Rows of A are stored with sorted columns (the order PETSc's AIJ and
// LASPACK's Q_SortEl both impose before any product).
#include <algorithm>
#include <cmath>
#include <cstring>
#include "fm_mesh.h"

const int fm_quad9_ref[9][2] = {{-1,-1},{1,-1},{1,1},{-1,1},{0,-1},{1,0},{0,1},{-1,0},{0,0}};
const int fm_hex27_ref[27][3] = {
  {-1,-1,-1},{1,-1,-1},{1,1,-1},{-1,1,-1},{-1,-1,1},{1,-1,1},{1,1,1},{-1,1,1},
  {0,-1,-1},{1,0,-1},{0,1,-1},{-1,0,-1},{-1,-1,0},{1,-1,0},{1,1,0},{-1,1,0},
  {0,-1,1},{1,0,1},{0,1,1},{-1,0,1},{0,0,-1},{0,-1,0},{1,0,0},{0,1,0},{-1,0,0},{0,0,1},{0,0,0}};
// side -> local face nodes (src/MGGeomEl.C:172-177, 310-317)
const int fm_face2d[4][3] = {{0,1,4},{1,2,5},{2,3,6},{3,0,7}};
const int fm_face3d[6][9] = {
  {0,1,2,3,8,9,10,11,20},{0,1,5,4,8,13,16,12,21},{1,2,6,5,9,14,17,13,22},
  {2,3,7,6,10,15,18,14,23},{3,0,4,7,11,12,19,15,24},{4,5,6,7,16,17,18,19,25}};

static inline int refc(int dim, int k, int d) { return dim == 2 ? fm_quad9_ref[k][d] : fm_hex27_ref[k][d]; }

int fm_build_box(int dim, const int nx0[3], const double lo[3], const double hi[3], int nlev, fm_mesh *m) {
  if ((dim != 2 && dim != 3) || nlev < 1 || nlev > 12) return FM_EINVAL;
  m->dim = dim; m->nlev = nlev;
  m->n2 = dim == 2 ? 9 : 27; m->n1 = dim == 2 ? 4 : 8; m->nsides = 2 * dim;
  m->nb2 = dim == 2 ? 3 : 9; m->nch = 1 << dim;
  int64_t G[3] = {1, 1, 1}, gtot = 1;
  for (int d = 0; d < 3; d++) {
    m->nx0[d] = d < dim ? nx0[d] : 1;
    if (m->nx0[d] < 1) return FM_EINVAL;
    m->lo[d] = d < dim ? lo[d] : 0.; m->hi[d] = d < dim ? hi[d] : 0.;
    if (d < dim) G[d] = (int64_t)2 * m->nx0[d] * (1LL << (nlev - 1)) + 1;
    gtot *= G[d];
  }
  if (gtot > (int64_t)1 << 31) return FM_EINVAL;
  m->nel.assign(nlev, 0);
  m->egrid.assign(nlev, {});
  m->parent.assign(nlev, {});
  // element lists: level 0 lexicographic, then children parent by parent
  {
    int n0 = m->nx0[0] * m->nx0[1] * m->nx0[2];
    m->nel[0] = n0;
    m->egrid[0].resize(3 * (size_t)n0);
    m->parent[0].assign(n0, -1);
    int e = 0;
    for (int k = 0; k < m->nx0[2]; k++)
      for (int j = 0; j < m->nx0[1]; j++)
        for (int i = 0; i < m->nx0[0]; i++, e++) {
          m->egrid[0][3 * e] = i; m->egrid[0][3 * e + 1] = j; m->egrid[0][3 * e + 2] = k;
        }
  }
  m->kids.assign(nlev, {});
  m->chid.assign(nlev, {});
  for (int L = 1; L < nlev; L++) {
    int np = m->nel[L - 1], nc = np * m->nch;
    m->nel[L] = nc;
    m->egrid[L].resize(3 * (size_t)nc);
    m->parent[L].resize(nc);
    m->kids[L].resize(nc);
    m->chid[L].resize(nc);
    for (int p = 0; p < np; p++)
      for (int c = 0; c < m->nch; c++) {
        int ce = p * m->nch + c;
        for (int d = 0; d < 3; d++)
          m->egrid[L][3 * ce + d] = d < dim ? 2 * m->egrid[L - 1][3 * p + d] + ((c >> d) & 1) : 0;
        m->parent[L][ce] = p;
        m->kids[L][ce] = ce;
        m->chid[L][ce] = (unsigned char)c;
      }
  }
  // node of element e (level L), local k -> top-grid linear index
  auto gidx = [&](int L, int e, int k) -> int64_t {
    int S = 1 << (nlev - L);
    int64_t g = 0, stride = 1;
    for (int d = 0; d < dim; d++) {
      int64_t I = (int64_t)m->egrid[L][3 * e + d] * S + (refc(dim, k, d) + 1) * (S / 2);
      g += I * stride; stride *= G[d];
    }
    return g;
  };
  // first-encounter ids, level of first appearance, first level as a vertex
  std::vector<int> lid(gtot, -1);
  std::vector<unsigned char> glev(gtot, (unsigned char)nlev), glevp(gtot, (unsigned char)nlev);
  std::vector<int64_t> g_of_lid;
  g_of_lid.reserve(gtot);
  for (int L = 0; L < nlev; L++)
    for (int e = 0; e < m->nel[L]; e++)
      for (int k = 0; k < m->n2; k++) {
        int64_t g = gidx(L, e, k);
        if (lid[g] < 0) { lid[g] = (int)g_of_lid.size(); g_of_lid.push_back(g); glev[g] = (unsigned char)L; }
        if (k < m->n1 && glevp[g] > L) glevp[g] = (unsigned char)L;
      }
  int N = (int)g_of_lid.size();
  m->Ntop = N;
  // stable bucket sort by key = levp + (nlev+1)*lev over nodes in id order
  int nk = (nlev + 1) * (nlev + 1);
  std::vector<int> cnt(nk + 1, 0);
  for (int i = 0; i < N; i++) { int64_t g = g_of_lid[i]; cnt[glevp[g] + (nlev + 1) * glev[g] + 1]++; }
  for (int k = 0; k < nk; k++) cnt[k + 1] += cnt[k];
  std::vector<int> gnew(gtot, -1);
  m->xyz.assign((size_t)dim * N, 0.);
  m->nq2.assign(nlev, 0); m->nq1.assign(nlev, 0);
  for (int i = 0; i < N; i++) {
    int64_t g = g_of_lid[i];
    int nn = cnt[glevp[g] + (nlev + 1) * glev[g]]++;
    gnew[g] = nn;
    int64_t I[3] = {g % G[0], (g / G[0]) % G[1], g / (G[0] * G[1])};
    for (int d = 0; d < dim; d++)
      m->xyz[nn + (size_t)d * N] = m->lo[d] + (m->hi[d] - m->lo[d]) * (double)I[d] / (double)(G[d] - 1);
    for (int L = glev[g]; L < nlev; L++) m->nq2[L]++;
    for (int L = glevp[g]; L < nlev; L++) m->nq1[L]++;
  }
  // connectivity and face neighbours per level
  m->conn.assign(nlev, {});
  m->neigh.assign(nlev, {});
  static const int s2[4][2] = {{1, -1}, {0, 1}, {1, 1}, {0, -1}};
  static const int s3[6][2] = {{2, -1}, {1, -1}, {0, 1}, {1, 1}, {0, -1}, {2, 1}};
  for (int L = 0; L < nlev; L++) {
    int ne = m->nel[L];
    m->conn[L].resize((size_t)ne * m->n2);
    m->neigh[L].resize((size_t)ne * m->nsides);
    int nE[3] = {1, 1, 1};
    for (int d = 0; d < dim; d++) nE[d] = m->nx0[d] << L;
    std::vector<int> at((size_t)nE[0] * nE[1] * nE[2]);
    for (int e = 0; e < ne; e++) {
      const int *c = &m->egrid[L][3 * e];
      at[c[0] + (size_t)nE[0] * (c[1] + (size_t)nE[1] * c[2])] = e;
      for (int k = 0; k < m->n2; k++) m->conn[L][(size_t)e * m->n2 + k] = gnew[gidx(L, e, k)];
    }
    for (int e = 0; e < ne; e++) {
      const int *c = &m->egrid[L][3 * e];
      for (int s = 0; s < m->nsides; s++) {
        int ax = dim == 2 ? s2[s][0] : s3[s][0], sg = dim == 2 ? s2[s][1] : s3[s][1];
        int cc[3] = {c[0], c[1], c[2]};
        cc[ax] += sg;
        int nb = -1;
        if (cc[ax] >= 0 && cc[ax] < nE[ax]) nb = at[cc[0] + (size_t)nE[0] * (cc[1] + (size_t)nE[1] * cc[2])];
        m->neigh[L][(size_t)e * m->nsides + s] = nb;
      }
    }
  }
  return FM_OK;
}

void fm_init_dof(const fm_mesh &m, int L, int nq2, int nq1, std::vector<int> &node_dof, int64_t &ndof) {
  int N = m.Ntop;
  node_dof.assign((size_t)N * (nq2 + nq1), -1);
  int64_t c = 0;
  for (int iv = 0; iv < nq2; iv++)
    for (int k = 0; k < m.nq2[L]; k++) node_dof[k + (size_t)iv * N] = (int)c++;
  for (int iv = 0; iv < nq1; iv++)
    for (int k = 0; k < m.nq1[L]; k++) node_dof[k + (size_t)(nq2 + iv) * N] = (int)c++;
  ndof = c;
}

void fm_elem_dofs(const fm_mesh &m, int L, int e, int nq2, int nq1, const std::vector<int> &dof, int *out) {
  int N = m.Ntop, c = 0;
  const int *cn = &m.conn[L][(size_t)e * m.n2];
  for (int iv = 0; iv < nq2; iv++) for (int k = 0; k < m.n2; k++) out[c++] = dof[cn[k] + (size_t)iv * N];
  for (int iv = 0; iv < nq1; iv++) for (int k = 0; k < m.n1; k++) out[c++] = dof[cn[k] + (size_t)(nq2 + iv) * N];
}

// node -> incident elements (in element order) for level L
static void node_elems(const fm_mesh &m, int L, int nodes, std::vector<int64_t> &off, std::vector<int> &els) {
  off.assign(nodes + 1, 0);
  int ne = m.nel[L];
  for (int e = 0; e < ne; e++)
    for (int k = 0; k < m.n2; k++) off[m.conn[L][(size_t)e * m.n2 + k] + 1]++;
  for (int i = 0; i < nodes; i++) off[i + 1] += off[i];
  els.resize(off[nodes]);
  std::vector<int64_t> f(off.begin(), off.end() - 1);
  for (int e = 0; e < ne; e++)
    for (int k = 0; k < m.n2; k++) els[f[m.conn[L][(size_t)e * m.n2 + k]]++] = e;
}

void fm_pattern(const fm_mesh &m, int L, int nq2, int nq1, const std::vector<int> &dofL,
                std::vector<int64_t> &rp, std::vector<int> &ci) {
  int N = m.Ntop, nv = nq2 + nq1, nloc = nq2 * m.n2 + nq1 * m.n1;
  int64_t n = (int64_t)nq2 * m.nq2[L] + (int64_t)nq1 * m.nq1[L];
  std::vector<int64_t> off;
  std::vector<int> els;
  node_elems(m, L, m.nq2[L], off, els);
  // dof -> node
  std::vector<int> dnode(n);
  for (int iv = 0; iv < nv; iv++) {
    int nn = iv < nq2 ? m.nq2[L] : m.nq1[L];
    for (int k = 0; k < nn; k++) dnode[dofL[k + (size_t)iv * N]] = k;
  }
  // all rows of one node share the same column set
  std::vector<int64_t> len(n + 1, 0);
  int maxlen = 0;
  std::vector<std::vector<int>> cols_of_node_tmp;  // unused when streaming
  // first pass: count (per node) by building the sorted column list
  int nnodes = m.nq2[L];
  std::vector<int> node_len(nnodes, 0);
#pragma omp parallel
  {
    std::vector<int> buf;
    std::vector<int> ed(nloc);
#pragma omp for schedule(dynamic, 1024)
    for (int k = 0; k < nnodes; k++) {
      buf.clear();
      for (int64_t q = off[k]; q < off[k + 1]; q++) {
        fm_elem_dofs(m, L, els[q], nq2, nq1, dofL, ed.data());
        buf.insert(buf.end(), ed.begin(), ed.end());
      }
      std::sort(buf.begin(), buf.end());
      node_len[k] = (int)(std::unique(buf.begin(), buf.end()) - buf.begin());
    }
  }
  for (int64_t i = 0; i < n; i++) {
    int k = dnode[i];
    len[i + 1] = node_len[k];
    if (node_len[k] > maxlen) maxlen = node_len[k];
  }
  for (int64_t i = 0; i < n; i++) len[i + 1] += len[i];
  rp = len;
  ci.resize(rp[n]);
  // second pass: fill each dof row of node k with the node's column set
#pragma omp parallel
  {
    std::vector<int> buf;
    std::vector<int> ed(nloc);
#pragma omp for schedule(dynamic, 1024)
    for (int k = 0; k < nnodes; k++) {
      buf.clear();
      for (int64_t q = off[k]; q < off[k + 1]; q++) {
        fm_elem_dofs(m, L, els[q], nq2, nq1, dofL, ed.data());
        buf.insert(buf.end(), ed.begin(), ed.end());
      }
      std::sort(buf.begin(), buf.end());
      buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
      for (int iv = 0; iv < nv; iv++) {
        int nn = iv < nq2 ? m.nq2[L] : m.nq1[L];
        if (k >= nn) continue;
        int row = dofL[k + (size_t)iv * N];
        std::copy(buf.begin(), buf.end(), ci.begin() + rp[row]);
      }
    }
  }
}

// Child embedding value E[c][i][j] = phi_j(child node i) for Q2 (kind 2) / Q1 (kind 1)
static double lag2(int node, double x) { return node < 0 ? .5 * x * (x - 1.) : (node > 0 ? .5 * x * (x + 1.) : 1. - x * x); }
static double lag1(int node, double x) { return node < 0 ? .5 * (1. - x) : .5 * (1. + x); }
static void embed_tables(int dim, int kind, std::vector<double> &E) {
  int ns = kind == 2 ? (dim == 2 ? 9 : 27) : (dim == 2 ? 4 : 8), nch = 1 << dim;
  E.assign((size_t)nch * ns * ns, 0.);
  for (int c = 0; c < nch; c++)
    for (int i = 0; i < ns; i++) {
      double xp[3];
      for (int d = 0; d < dim; d++) xp[d] = (((c >> d) & 1) ? 0. : -1.) + (refc(dim, i, d) + 1) * 0.5;
      for (int j = 0; j < ns; j++) {
        double v = 1.;
        for (int d = 0; d < dim; d++) v *= kind == 2 ? lag2(refc(dim, j, d), xp[d]) : lag1(refc(dim, j, d), xp[d]);
        E[((size_t)c * ns + i) * ns + j] = v;
      }
    }
}

// Node-level transfer block of one kind (2: Q2 nodes, 1: Q1 nodes) on level L: per row node the (column node,
// value) pairs in the order gencase emits them.  rest = false: prolongation (fine L <- coarse L-1), the last fine
// element containing the node defines its row (MGGenCase.C:2308-2390, entries |E| <= 1e-8 dropped); rest = true:
// restriction (coarse L <- fine L+1), over the coarse elements containing the row node in element order, their
// children's nodes in order, the first occurrence kept (MGGenCase.C:2452-2566, |E| <= 1e-6 dropped).  Unmasked.
void fm_transfer_block(const fm_mesh &m, int L, int kind, bool rest, std::vector<std::vector<std::pair<int, double>>> &blk) {
  const int dim = m.dim, nch = m.nch, ns = kind == 2 ? m.n2 : m.n1;
  const int nr = kind == 2 ? m.nq2[L] : m.nq1[L];
  std::vector<double> E;
  embed_tables(dim, kind, E);
  blk.assign(nr, {});
  if (!rest) {
    std::vector<int> last(nr, -1), lastk(nr, -1);
    for (int e = 0; e < m.nel[L]; e++)
      for (int kk = 0; kk < ns; kk++) { int r = m.conn[L][(size_t)e * m.n2 + kk]; last[r] = e; lastk[r] = kk; }
    for (int r = 0; r < nr; r++) {
      int e = last[r], kk = lastk[r], par = m.parent[L][e], ch = m.chid[L][e];
      for (int kj = 0; kj < ns; kj++) {
        double val = E[((size_t)ch * ns + kk) * ns + kj];
        if (std::fabs(val) > 1.e-8) blk[r].push_back({m.conn[L - 1][(size_t)par * m.n2 + kj], val});
      }
    }
    return;
  }
  std::vector<int64_t> off;
  std::vector<int> els;
  node_elems(m, L, m.nq2[L], off, els);
  const int nfine = m.nq2[L + 1];
#pragma omp parallel
  {
    std::vector<int> mark(nfine, -1);
#pragma omp for schedule(dynamic, 256)
    for (int r = 0; r < nr; r++) {
      auto &row = blk[r];
      for (int64_t q = off[r]; q < off[r + 1]; q++) {
        int e = els[q], kk = -1;
        for (int k = 0; k < ns; k++) if (m.conn[L][(size_t)e * m.n2 + k] == r) kk = k;
        if (kk < 0) continue;  // r is not among this element's first ns nodes (Q1 block)
        for (int ich = 0; ich < nch; ich++) {
          int ch = m.kids[L + 1][(size_t)e * nch + ich];
          for (int kj = 0; kj < ns; kj++) {
            double val = E[((size_t)ich * ns + kj) * ns + kk];
            if (std::fabs(val) <= 1.e-6) continue;
            int col = m.conn[L + 1][(size_t)ch * m.n2 + kj];
            if (mark[col] == r) continue;
            mark[col] = r;
            row.push_back({col, val});
          }
        }
      }
    }
  }
}

// Transfer operators in DoF space.  rest=true: R (coarse L <- fine L+1), rows
// masked by bc0 % 2 of the coarse node's top DoF; rest=false: P (fine L <- coarse L-1).
void fm_transfer(const fm_mesh &m, int L, int nq2, int nq1, bool rest, const std::vector<int> &dof_row,
                 const std::vector<int> &dof_col, const std::vector<int> *dof_top, const std::vector<int> *bc0,
                 std::vector<int64_t> &rp, std::vector<int> &ci, std::vector<double> &v) {
  int N = m.Ntop, nv = nq2 + nq1;
  int64_t nrows = (int64_t)nq2 * m.nq2[L] + (int64_t)nq1 * m.nq1[L];
  // per node-kind block rows: list of (col node, value)
  std::vector<std::vector<std::pair<int, double>>> blk[2];
  for (int t = 0; t < 2; t++) {
    if ((t == 0 && nq2 == 0) || (t == 1 && nq1 == 0)) continue;
    fm_transfer_block(m, L, t == 0 ? 2 : 1, rest, blk[t]);
  }
  rp.assign(nrows + 1, 0);
  for (int iv = 0; iv < nv; iv++) {
    int t = iv < nq2 ? 0 : 1;
    int nr = t == 0 ? m.nq2[L] : m.nq1[L];
    for (int i = 0; i < nr; i++) rp[dof_row[i + (size_t)iv * N] + 1] = (int64_t)blk[t][i].size();
  }
  for (int64_t i = 0; i < nrows; i++) rp[i + 1] += rp[i];
  ci.resize(rp[nrows]);
  v.resize(rp[nrows]);
  for (int iv = 0; iv < nv; iv++) {
    int t = iv < nq2 ? 0 : 1;
    int nr = t == 0 ? m.nq2[L] : m.nq1[L];
    for (int i = 0; i < nr; i++) {
      int row = dof_row[i + (size_t)iv * N];
      double mask = 1.;
      if (rest) mask = (double)((*bc0)[(*dof_top)[i + (size_t)iv * N]] % 2);
      int64_t q = rp[row];
      for (auto &pr : blk[t][i]) { ci[q] = dof_col[pr.first + (size_t)iv * N]; v[q] = mask * pr.second; q++; }
    }
  }
}

// ---------------------------------------------------------------------------------------------------------------
// A hierarchy handed in by the caller in MGMesh's layout (src/MGMesh.C:493-899, include/MGMesh.h:65-80)
// ---------------------------------------------------------------------------------------------------------------
// Greedy colouring of level L's elements (in element order, the smallest colour no earlier element sharing a node
// holds); returns the number of colours
static int greedy_colours(const fm_mesh &m, int L, std::vector<unsigned char> &col) {
  const int ne = m.nel[L], n2 = m.n2;
  std::vector<int64_t> off;
  std::vector<int> els;
  node_elems(m, L, m.nq2[L], off, els);
  col.assign(ne, 0);
  int nc = 0;
  for (int e = 0; e < ne; e++) {
    unsigned used = 0;
    for (int k = 0; k < n2; k++) {
      const int nd = m.conn[L][(size_t)e * n2 + k];
      for (int64_t q = off[nd]; q < off[nd + 1]; q++)
        if (els[q] < e) used |= 1u << col[els[q]];
    }
    int c = 0;
    while (c < 31 && ((used >> c) & 1)) c++;
    col[e] = (unsigned char)c;
    nc = std::max(nc, c + 1);
  }
  return nc;
}

// true when no two elements of one colour share a node
static bool colouring_valid(const fm_mesh &m, int L, const std::vector<unsigned char> &col) {
  std::vector<int> seen((size_t)m.nq2[L] * 8, -1);
  for (int e = 0; e < m.nel[L]; e++)
    for (int k = 0; k < m.n2; k++) {
      const int nd = m.conn[L][(size_t)e * m.n2 + k];
      int &s = seen[(size_t)nd * 8 + col[e]];
      if (s >= 0 && s != e) return false;
      s = e;
    }
  return true;
}

// Parent and child index of every element of level L from the nested node sets alone: the vertices of a child
// are Q2 nodes of its parent, at the parent's reference points (b_d - 1) + (r_d + 1) / 2 for child c with corner
// bits b_d = (c >> d) & 1 and vertex reference coordinates r; the first (coarse element, child) pair that holds
// every vertex there is taken.  Fails when an element has none (a child not oriented like its parent).
static int infer_parents(const fm_mesh &m, int L, std::vector<int> &par, std::vector<int> &ch, std::string &err) {
  const int dim = m.dim, n1 = m.n1, n2 = m.n2, nch = m.nch;
  // at[c][i]: the parent's local node at vertex i of child c
  std::vector<int> at((size_t)nch * n1, -1);
  for (int c = 0; c < nch; c++)
    for (int i = 0; i < n1; i++) {
      int q[3] = {0, 0, 0};
      for (int d = 0; d < dim; d++) q[d] = (((c >> d) & 1) - 1) + (refc(dim, i, d) + 1) / 2;
      for (int k = 0; k < n2; k++) {
        bool same = true;
        for (int d = 0; d < dim; d++) same = same && refc(dim, k, d) == q[d];
        if (same) at[(size_t)c * n1 + i] = k;
      }
    }
  std::vector<int64_t> off;
  std::vector<int> els;
  node_elems(m, L - 1, m.nq2[L - 1], off, els);
  const int ne = m.nel[L];
  par.assign(ne, -1);
  ch.assign(ne, -1);
  for (int e = 0; e < ne; e++) {
    const int *cf = &m.conn[L][(size_t)e * n2];
    const int v0 = cf[0];
    if (v0 >= m.nq2[L - 1]) { err = "fm_mesh_create: a vertex of a level " + std::to_string(L) + " element is no coarse node"; return -1; }
    for (int64_t q = off[v0]; q < off[v0 + 1] && par[e] < 0; q++) {
      const int p = els[q];
      const int *cc = &m.conn[L - 1][(size_t)p * n2];
      for (int c = 0; c < nch && par[e] < 0; c++) {
        bool ok = true;
        for (int i = 0; i < n1 && ok; i++) ok = cf[i] == cc[at[(size_t)c * n1 + i]];
        if (ok) { par[e] = p; ch[e] = c; }
      }
    }
    if (par[e] < 0) {
      err = "fm_mesh_create: no parent of level " + std::to_string(L) + " element " + std::to_string(e) +
            " holds its vertices at a child's positions (pass parent and child)";
      return -1;
    }
  }
  return 0;
}

int fm_mesh_from_arrays(int dim, int nlev, int Ntop, const double *xyz, const int *off_el, const int *el_map,
                        const int *el_neigh, const int *nq2, const int *nq1, const int *parent, const int *child,
                        fm_mesh *m, std::string &err) {
  auto bad = [&](const std::string &w) { err = "fm_mesh_create: " + w; return -1; };
  if ((dim != 2 && dim != 3) || nlev < 1 || nlev > 12 || Ntop < 1) return bad("dim, levels or node count out of range");
  if (!xyz || !off_el || !el_map || !el_neigh || !nq2 || !nq1) return bad("missing array");
  if ((parent == nullptr) != (child == nullptr)) return bad("parent and child arrays go together");
  m->dim = dim; m->nlev = nlev;
  m->n2 = dim == 2 ? 9 : 27; m->n1 = dim == 2 ? 4 : 8; m->nsides = 2 * dim;
  m->nb2 = dim == 2 ? 3 : 9; m->nch = 1 << dim;
  m->Ntop = Ntop;
  m->lattice = false;
  if (off_el[0] != 0) return bad("off_el[0] must be 0");
  m->nel.assign(nlev, 0); m->nq2.assign(nlev, 0); m->nq1.assign(nlev, 0);
  for (int L = 0; L < nlev; L++) {
    if (off_el[L + 1] <= off_el[L]) return bad("level " + std::to_string(L) + " has no elements");
    m->nel[L] = off_el[L + 1] - off_el[L];
    m->nq2[L] = nq2[L]; m->nq1[L] = nq1[L];
    if (nq1[L] < 1 || nq1[L] > nq2[L] || nq2[L] > Ntop || (L > 0 && nq2[L] < nq2[L - 1]))
      return bad("node counts of level " + std::to_string(L) + " inconsistent");
  }
  if (nq2[nlev - 1] != Ntop) return bad("the top level must hold every node");
  m->xyz.assign(xyz, xyz + (size_t)dim * Ntop);
  for (int d = 0; d < 3; d++) { m->nx0[d] = 1; m->lo[d] = 0.; m->hi[d] = 0.; }
  for (int d = 0; d < dim; d++) {
    m->lo[d] = *std::min_element(m->xyz.begin() + (size_t)d * Ntop, m->xyz.begin() + (size_t)(d + 1) * Ntop);
    m->hi[d] = *std::max_element(m->xyz.begin() + (size_t)d * Ntop, m->xyz.begin() + (size_t)(d + 1) * Ntop);
  }
  m->conn.assign(nlev, {}); m->neigh.assign(nlev, {}); m->parent.assign(nlev, {}); m->egrid.assign(nlev, {});
  m->kids.assign(nlev, {}); m->chid.assign(nlev, {}); m->ecol.assign(nlev, {});
  for (int L = 0; L < nlev; L++) {
    const int ne = m->nel[L];
    const int *cm = el_map + (size_t)off_el[L] * m->n2, *nb = el_neigh + (size_t)off_el[L] * m->nsides;
    m->conn[L].assign(cm, cm + (size_t)ne * m->n2);
    m->neigh[L].assign(nb, nb + (size_t)ne * m->nsides);
    std::vector<char> used2(nq2[L], 0), used1(nq1[L], 0);
    for (int e = 0; e < ne; e++) {
      for (int k = 0; k < m->n2; k++) {
        const int nd = m->conn[L][(size_t)e * m->n2 + k];
        if (nd < 0 || nd >= nq2[L]) return bad("level " + std::to_string(L) + " element node outside the level's nodes");
        if (k < m->n1) {
          if (nd >= nq1[L]) return bad("level " + std::to_string(L) + " element vertex outside the level's Q1 nodes");
          used1[nd] = 1;
        }
        used2[nd] = 1;
      }
      // neighbours are global element ids (gencase's EL_NEIG numbering) of the same level, -1 on the boundary
      for (int sd = 0; sd < m->nsides; sd++) {
        int &q = m->neigh[L][(size_t)e * m->nsides + sd];
        if (q != -1 && (q < off_el[L] || q >= off_el[L + 1]))
          return bad("level " + std::to_string(L) + " neighbour outside the level's elements");
        if (q != -1) q -= off_el[L];
      }
    }
    for (char u : used2) if (!u) return bad("a node of level " + std::to_string(L) + " belongs to no element");
    for (char u : used1) if (!u) return bad("a Q1 node of level " + std::to_string(L) + " is no element vertex");
    m->parent[L].assign(ne, -1);
    if (L > 0) {
      const int np = m->nel[L - 1];
      m->kids[L].assign((size_t)np * m->nch, -1);
      m->chid[L].assign(ne, 0);
      std::vector<int> ipar, ich;
      if (!parent) {
        std::string e2;
        if (infer_parents(*m, L, ipar, ich, e2)) return bad(e2);
      }
      for (int e = 0; e < ne; e++) {
        const int p = parent ? parent[off_el[L] + e] : ipar[e], c = child ? child[off_el[L] + e] : ich[e];
        if (p < 0 || p >= np || c < 0 || c >= m->nch) return bad("parent or child index out of range");
        int &slot = m->kids[L][(size_t)p * m->nch + c];
        if (slot >= 0) return bad("two elements are the same child of one parent");
        slot = e;
        m->parent[L][e] = p;
        m->chid[L][e] = (unsigned char)c;
      }
      for (int k : m->kids[L]) if (k < 0) return bad("a parent misses a child");
    }
    // colours: the child index (a consistently oriented refinement gives 2^dim colours of non-touching
    // elements), else greedy
    bool ok = false;
    if (L > 0) {
      m->ecol[L] = m->chid[L];
      ok = colouring_valid(*m, L, m->ecol[L]);
    }
    if (!ok) {
      const int nc = greedy_colours(*m, L, m->ecol[L]);
      if (nc > 8) return bad("level " + std::to_string(L) + " needs more than 8 element colours");
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------------------------------------------
// Preprocessor output of gencase (the datasets of mesh.h5 and of the Matrix / Prol / Rest files)
// ---------------------------------------------------------------------------------------------------------------
// Rows / columns of one kind on level L: 2 = the level's Q2 nodes, 1 = its Q1 nodes, 0 = its elements (one
// piecewise-constant DoF each, NDOF_K = 1)
static int kind_count(const fm_mesh &m, int L, int kind) {
  return kind == 2 ? m.nq2[L] : (kind == 1 ? m.nq1[L] : m.nel[L]);
}

// Block of the matrix pattern of level L in gencase's column order (compute_mat_qq, MGGenCase.C:1781-1853, and the
// piecewise blocks compute_mat_qk / _kq / _lk / _kl / _kk, :1855-2307).  Node rows (rkind 2 / 1) list, over the
// elements holding the row node among their first n(rkind) local nodes in element order, the elements' first
// n(ckind) local nodes in local order (ckind 2 / 1) or the element itself (ckind 0), each column at its first
// occurrence.  Element rows (rkind 0) list the element's nodes that are level-L nodes of kind ckind in local order
// (all n2 local nodes are scanned and mapped through the level's node map, so for ckind 1 only the vertices remain)
// or, for ckind 0, the element itself.  len: rows + 1 prefix counts; pos may be NULL (count only).
int64_t fm_gencase_block(const fm_mesh &m, int L, int rkind, int ckind, std::vector<int> &len, std::vector<int> *pos) {
  const int nrows = kind_count(m, L, rkind), ncols = kind_count(m, L, ckind);
  len.assign(nrows + 1, 0);
  if (pos) pos->clear();
  int64_t count = 0;
  if (rkind == 0) {
    for (int e = 0; e < nrows; e++) {
      const int *c = &m.conn[L][(size_t)e * m.n2];
      if (ckind == 0) {
        if (pos) pos->push_back(e);
        count++;
      } else {
        for (int k = 0; k < m.n2; k++)
          if (c[k] < ncols) { if (pos) pos->push_back(c[k]); count++; }
      }
      len[e + 1] = (int)count;
    }
    return count;
  }
  const int nr = rkind == 2 ? m.n2 : m.n1, nc = ckind == 2 ? m.n2 : (ckind == 1 ? m.n1 : 0);
  std::vector<int64_t> off;
  std::vector<int> els;
  node_elems(m, L, m.nq2[L], off, els);
  std::vector<int> seen(ncols, -1), row;
  for (int r = 0; r < nrows; r++) {
    row.clear();
    for (int64_t q = off[r]; q < off[r + 1]; q++) {
      const int e = els[q];
      const int *c = &m.conn[L][(size_t)e * m.n2];
      bool holds = false;
      for (int k = 0; k < nr; k++) holds = holds || c[k] == r;
      if (!holds) continue;
      if (ckind == 0) {
        if (seen[e] != r) { seen[e] = r; row.push_back(e); }
        continue;
      }
      for (int k = 0; k < nc; k++)
        if (seen[c[k]] != r) { seen[c[k]] = r; row.push_back(c[k]); }
    }
    count += (int64_t)row.size();
    len[r + 1] = (int)count;
    if (pos) pos->insert(pos->end(), row.begin(), row.end());
  }
  return count;
}

// Piecewise-constant transfers (compute_prol_kk, MGGenCase.C:2392-2448: a fine element takes its parent's value;
// compute_res_kk, :2568-2631: a coarse element sums its children in child order), value 1 each.
void fm_gencase_transfer_kk(const fm_mesh &m, int L, bool rest, std::vector<std::vector<std::pair<int, double>>> &blk) {
  blk.assign(m.nel[L], {});
  for (int e = 0; e < m.nel[L]; e++) {
    if (!rest) { blk[e].push_back({m.parent[L][e], 1.}); continue; }
    for (int ich = 0; ich < m.nch; ich++) blk[e].push_back({m.kids[L + 1][(size_t)e * m.nch + ich], 1.});
  }
}

// Boundary elements of level L as gencase lists them (MGGenCase.C:330-380, 550-570): over the level's elements in
// order, every side without a neighbour, its nodes in the side node order of libMesh's build_side (libMesh 0.9.4
// Quad9 / Hex27 side_nodes_map, which orients every side outwards).
static const int g_side2d[4][3] = {{0, 1, 4}, {1, 2, 5}, {2, 3, 6}, {3, 0, 7}};
static const int g_side3d[6][9] = {{0, 3, 2, 1, 11, 10, 9, 8, 20}, {0, 1, 5, 4, 8, 13, 16, 12, 21},
                                   {1, 2, 6, 5, 9, 14, 17, 13, 22}, {2, 3, 7, 6, 10, 15, 18, 14, 23},
                                   {3, 0, 4, 7, 11, 12, 19, 15, 24}, {4, 5, 6, 7, 16, 17, 18, 19, 25}};
void fm_boundary_mesh(const fm_mesh &m, int L, std::vector<int> &nodes) {
  nodes.clear();
  for (int e = 0; e < m.nel[L]; e++)
    for (int sd = 0; sd < m.nsides; sd++) {
      if (m.neigh[L][(size_t)e * m.nsides + sd] != -1) continue;
      for (int k = 0; k < m.nb2; k++) {
        const int ln = m.dim == 2 ? g_side2d[sd][k] : g_side3d[sd][k];
        nodes.push_back(m.conn[L][(size_t)e * m.n2 + ln]);
      }
    }
}
