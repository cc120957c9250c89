// femus.cpp -- host side of the MI355X FEMuS hot path: context, system setup
// (MGSolDA::init_dof / GenBc / init), GenMatRhs drivers, the multigrid
// V-cycle (MGSolBase::MGStep / MGSolve) in FP64 and FP32, the FP64 outer
// loops (iterative refinement, flexible GMRES) and the C ABI.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "fm_internal.h"
#include "fm_kernels.h"
#include "sell_host.h"

static thread_local std::string g_err;
static thread_local unsigned g_nerr = 0;
void fm_set_error(const std::string &msg) { g_err = msg; g_nerr++; }
unsigned fm_error_count() { return g_nerr; }
extern "C" const char *fm_last_error(void) { return g_err.c_str(); }
extern "C" int fm_set_sell_split(int W) { return fmk::sell_split(W); }

// Element-buffer assembly (element matrices to a buffer, then a gather into the SELL slots) for the levels whose
// buffers fit the budget; 0 keeps every level on the coloured scatter.
static int g_asm_gather = 1;
static constexpr size_t kGatherBudget = (size_t)12 << 30;  // element buffers + gather index of one system
extern "C" int fm_set_assembly_gather(int on) {
  if (on != 0 && on != 1) return fm_einval(__func__);
  const int old = g_asm_gather;
  g_asm_gather = on;
  return old;
}
// narrowest column class of SELL copies built from now on (FEMUS_SELL_P2=0 starts it at 8, FEMUS_SELL_C8=0 at 16)
static int g_sell_minbits = -1;
static int sell_minbits() {
  if (g_sell_minbits < 0) {
    const char *c8env = getenv("FEMUS_SELL_C8");
    const char *p2env = getenv("FEMUS_SELL_P2");
    g_sell_minbits = (c8env && atoi(c8env) == 0) ? 16 : ((p2env && atoi(p2env) == 0) ? 8 : 2);
  }
  return g_sell_minbits;
}
extern "C" int fm_set_sell_offsets(int bits) {
  const int old = sell_minbits();
  if (bits < 0) return old;
  if (bits != 2 && bits != 8 && bits != 16 && bits != 32) return fm_einval(__func__);
  g_sell_minbits = bits;
  return old;
}

template <typename T> static int dalloc(T **p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) { fm_set_error(std::string("hipMalloc: ") + hipGetErrorString(e)); return FM_ENOMEM; }
  return 0;
}
template <typename T> static void dfree(T *&p) { if (p) { (void)hipFree(p); p = nullptr; } }

// ===========================================================================
// context
// ===========================================================================
// frees what a context holds (any subset of it: also a context whose creation failed half-way)
static void ctx_release(fm_ctx *c) {
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  dfree(c->scratch);
  if (c->h_scratch) (void)hipHostFree(c->h_scratch);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int fm_noalloc_attach(int device);  // noalloc.hip
static int ctx_setup(fm_ctx *c) {
  FM_HIP(hipSetDevice(c->device));
  FM_CHECK(fm_noalloc_attach(c->device));
  FM_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  FM_HIP(hipEventCreate(&c->ev0));
  FM_HIP(hipEventCreate(&c->ev1));
  FM_CHECK(dalloc(&c->scratch, 4096));
  FM_HIP(hipHostMalloc((void **)&c->h_scratch, 4096 * sizeof(double), hipHostMallocDefault));
  return FM_OK;
}

extern "C" int fm_ctx_create(int device, fm_ctx **out) {
  if (!out) return fm_einval(__func__);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { fm_set_error("no HIP device"); return FM_EHIP; }
  if (device < 0 || device >= ndev) return fm_einval(__func__);
  fm_ctx *c = new fm_ctx();
  c->device = device;
  const int rc = ctx_setup(c);
  if (rc != FM_OK) { ctx_release(c); return rc; }
  *out = c;
  return FM_OK;
}

extern "C" int fm_ctx_destroy(fm_ctx *c) {
  if (!c) return fm_einval(__func__);
  if (c->nsys > 0) { fm_set_error("fm_ctx_destroy: systems still alive on this context"); return FM_ESTATE; }
  ctx_release(c);
  return FM_OK;
}

extern "C" int fm_sync(fm_ctx *c) {
  if (!c) return fm_einval(__func__);
  FM_HIP(hipStreamSynchronize(c->stream));
  return FM_OK;
}

// ===========================================================================
// mesh
// ===========================================================================
extern "C" int fm_mesh_create_box(int dim, const int nx0[3], const double lo[3], const double hi[3], int nlev,
                                  fm_mesh **out) {
  if (!out || !nx0 || !lo || !hi) return fm_einval(__func__);
  fm_mesh *m = new fm_mesh();
  int rc = fm_build_box(dim, nx0, lo, hi, nlev, m);
  if (rc) { delete m; fm_set_error("fm_mesh_create_box: invalid sizes"); return rc; }
  *out = m;
  return FM_OK;
}
extern "C" int fm_mesh_create(int dim, int nlev, int n_nodes, const double *xyz, const int *off_el, const int *el_map,
                              const int *el_neighbor, const int *nodes_q2, const int *nodes_q1, const int *parent,
                              const int *child, fm_mesh **out) {
  if (!out) return fm_einval(__func__);
  fm_mesh *m = new fm_mesh();
  std::string err;
  if (fm_mesh_from_arrays(dim, nlev, n_nodes, xyz, off_el, el_map, el_neighbor, nodes_q2, nodes_q1, parent, child, m,
                          err)) {
    delete m;
    fm_set_error(err);
    return FM_EINVAL;
  }
  *out = m;
  return FM_OK;
}
extern "C" int fm_mesh_destroy(fm_mesh *m) { delete m; return FM_OK; }
extern "C" int fm_mesh_info(const fm_mesh *m, int *info) {
  if (!m || !info) return fm_einval(__func__);
  info[0] = m->dim; info[1] = m->nlev; info[2] = m->n2; info[3] = m->n1; info[4] = m->nsides; info[5] = m->Ntop;
  return FM_OK;
}
extern "C" int fm_mesh_level(const fm_mesh *m, int L, int *n_el, int *n_q2, int *n_q1) {
  if (!m || L < 0 || L >= m->nlev) return fm_einval(__func__);
  if (n_el) *n_el = m->nel[L];
  if (n_q2) *n_q2 = m->nq2[L];
  if (n_q1) *n_q1 = m->nq1[L];
  return FM_OK;
}
extern "C" int fm_mesh_coords(const fm_mesh *m, double *xyz) {
  if (!m || !xyz) return fm_einval(__func__);
  std::memcpy(xyz, m->xyz.data(), m->xyz.size() * sizeof(double));
  return FM_OK;
}
extern "C" int fm_mesh_conn(const fm_mesh *m, int L, int *conn, int *neigh, int *parent) {
  if (!m || L < 0 || L >= m->nlev) return fm_einval(__func__);
  if (conn) std::memcpy(conn, m->conn[L].data(), m->conn[L].size() * sizeof(int));
  if (neigh) std::memcpy(neigh, m->neigh[L].data(), m->neigh[L].size() * sizeof(int));
  if (parent) std::memcpy(parent, m->parent[L].data(), m->parent[L].size() * sizeof(int));
  return FM_OK;
}

// ===========================================================================
// preprocessor output (the datasets of gencase's mesh and operator files)
// ===========================================================================
static bool gencase_kind(int k) { return k == 0 || k == 1 || k == 2; }
extern "C" int fm_gencase_matrix_block(const fm_mesh *m, int L, int rkind, int ckind, int *len, int *pos,
                                       int64_t *nnz) {
  if (!m || L < 0 || L >= m->nlev || !gencase_kind(rkind) || !gencase_kind(ckind) || !len || !nnz)
    return fm_einval(__func__);
  std::vector<int> ln, ps;
  *nnz = fm_gencase_block(*m, L, rkind, ckind, ln, pos ? &ps : nullptr);
  std::memcpy(len, ln.data(), ln.size() * sizeof(int));
  if (pos) std::memcpy(pos, ps.data(), ps.size() * sizeof(int));
  return FM_OK;
}
extern "C" int fm_gencase_transfer_block(const fm_mesh *m, int L, int kind, int rest, int *len, int *pos,
                                         double *val, int64_t *nnz) {
  if (!m || !gencase_kind(kind) || !len || !nnz || (pos == nullptr) != (val == nullptr)) return fm_einval(__func__);
  if (rest ? (L < 0 || L + 1 >= m->nlev) : (L < 1 || L >= m->nlev)) {
    fm_set_error(std::string(__func__) + (rest ? ": restriction needs 0 <= L < nlev-1" : ": prolongation needs 1 <= L < nlev"));
    return FM_EINVAL;
  }
  std::vector<std::vector<std::pair<int, double>>> blk;
  if (kind == 0) fm_gencase_transfer_kk(*m, L, rest != 0, blk);
  else fm_transfer_block(*m, L, kind, rest != 0, blk);
  int64_t q = 0;
  len[0] = 0;
  for (size_t r = 0; r < blk.size(); r++) {
    for (auto &pr : blk[r]) {
      if (pos) { pos[q] = pr.first; val[q] = pr.second; }
      q++;
    }
    len[r + 1] = (int)q;
  }
  *nnz = q;
  return FM_OK;
}
extern "C" int fm_gencase_boundary(const fm_mesh *m, int L, int *nodes, int *n_bd) {
  if (!m || L < 0 || L >= m->nlev || !n_bd) return fm_einval(__func__);
  std::vector<int> nd;
  fm_boundary_mesh(*m, L, nd);
  *n_bd = (int)(nd.size() / m->nb2);
  if (nodes) std::memcpy(nodes, nd.data(), nd.size() * sizeof(int));
  return FM_OK;
}

// ===========================================================================
// FE tables (closed-form Q2/Q1 Lagrange at 3-point Gauss-Legendre points)
// ===========================================================================
static double l2f(int c, double x) { return c < 0 ? .5 * x * (x - 1.) : (c > 0 ? .5 * x * (x + 1.) : 1. - x * x); }
static double dl2f(int c, double x) { return c < 0 ? x - .5 : (c > 0 ? x + .5 : -2. * x); }
static double l1f(int c, double x) { return c < 0 ? .5 * (1. - x) : .5 * (1. + x); }
static int ref_c(int dim, int k, int d) {
  if (dim == 1) return k == 0 ? -1 : (k == 1 ? 1 : 0);
  return dim == 2 ? fm_quad9_ref[k][d] : fm_hex27_ref[k][d];
}
static void gauss_pt(int dim, int qp, double *xi, double *w) {
  const double a = -std::sqrt(3. / 5.), x1[3] = {a, 0., -a}, w1[3] = {5. / 9., 8. / 9., 5. / 9.};
  int idx[3] = {0, 0, 0};
  if (dim == 1) idx[0] = qp;
  else if (dim == 2) { idx[0] = qp / 3; idx[1] = qp % 3; }
  else { idx[0] = qp / 9; idx[1] = (qp / 3) % 3; idx[2] = qp % 3; }
  *w = 1.;
  for (int d = 0; d < dim; d++) { xi[d] = x1[idx[d]]; *w *= w1[idx[d]]; }
}
static void build_tables(int dim, FeTab &t) {
  std::memset(&t, 0, sizeof(t));
  int n2 = dim == 2 ? 9 : 27, n1 = dim == 2 ? 4 : 8, ng = n2;
  for (int qp = 0; qp < ng; qp++) {
    double xi[3], w;
    gauss_pt(dim, qp, xi, &w);
    t.w[qp] = w;
    for (int s = 0; s < n2; s++) {
      double v = 1.;
      for (int d = 0; d < dim; d++) v *= l2f(ref_c(dim, s, d), xi[d]);
      t.phi2[s * ng + qp] = v;
      for (int dir = 0; dir < dim; dir++) {
        double g = 1.;
        for (int d = 0; d < dim; d++) g *= d == dir ? dl2f(ref_c(dim, s, d), xi[d]) : l2f(ref_c(dim, s, d), xi[d]);
        t.dphi2[(s * ng + qp) + dir * n2 * ng] = g;
      }
    }
    for (int s = 0; s < n1; s++) {
      double v = 1.;
      for (int d = 0; d < dim; d++) v *= l1f(ref_c(dim, s, d), xi[d]);
      t.phi1[s * ng + qp] = v;
    }
  }
  int bd = dim - 1, nb = dim == 2 ? 3 : 9, ngb = nb;
  for (int qp = 0; qp < ngb; qp++) {
    double xi[3], w;
    gauss_pt(bd, qp, xi, &w);
    t.bw[qp] = w;
    for (int s = 0; s < nb; s++) {
      double v = 1.;
      for (int d = 0; d < bd; d++) v *= l2f(ref_c(bd, s, d), xi[d]);
      t.bphi2[s * ngb + qp] = v;
      for (int dir = 0; dir < bd; dir++) {
        double g = 1.;
        for (int d = 0; d < bd; d++) g *= d == dir ? dl2f(ref_c(bd, s, d), xi[d]) : l2f(ref_c(bd, s, d), xi[d]);
        t.bdphi2[(s * ngb + qp) + dir * nb * ngb] = g;
      }
    }
  }
}

// ===========================================================================
// system
// ===========================================================================
extern "C" int fm_system_create(fm_ctx *ctx, const fm_mesh *m, int nq2, int nq1, fm_system **out) {
  if (!ctx || !m || !out || nq2 < 0 || nq1 < 0 || nq2 + nq1 == 0 || nq2 + nq1 > 8) return fm_einval(__func__);
  if (nq1 > 0 && nq2 == 0) return fm_einval(__func__);
  fm_system *s = new fm_system();
  s->ctx = ctx; s->mesh = m; s->nq2 = nq2; s->nq1 = nq1;
  s->nloc = nq2 * m->n2 + nq1 * m->n1;
  s->node_dof.resize(m->nlev);
  s->Dim.resize(m->nlev);
  for (int L = 0; L < m->nlev; L++) fm_init_dof(*m, L, nq2, nq1, s->node_dof[L], s->Dim[L]);
  int64_t nt = s->Dim[m->nlev - 1];
  s->bc0.assign(nt, 1); s->bc1.assign(nt, 1);
  s->lev.resize(m->nlev);
  s->u_pat.resize(m->nlev); s->u_rst.resize(m->nlev); s->u_prl.resize(m->nlev);
  ctx->nsys++;
  *out = s;
  return FM_OK;
}

// A caller's CSR (reference numbering) checked and stored with sorted columns; vals may be NULL (pattern only)
static int user_csr(const char *fn, int64_t nr, int64_t nc, const int64_t *rp, const int *ci, const double *v,
                    UserCsr &u) {
  if (!rp || !ci || rp[0] != 0) return fm_einval(fn);
  for (int64_t i = 0; i < nr; i++)
    if (rp[i + 1] < rp[i]) return fm_einval(fn);
  const int64_t nnz = rp[nr];
  u.rp.assign(rp, rp + nr + 1);
  u.ci.resize(nnz);
  u.v.assign(v ? nnz : 0, 0.);
  std::vector<std::pair<int, double>> row;
  for (int64_t i = 0; i < nr; i++) {
    row.clear();
    for (int64_t q = rp[i]; q < rp[i + 1]; q++) {
      if (ci[q] < 0 || ci[q] >= nc) { fm_set_error(std::string(fn) + ": column out of range in row " + std::to_string(i)); return FM_EINVAL; }
      row.emplace_back(ci[q], v ? v[q] : 0.);
    }
    std::sort(row.begin(), row.end(), [](const std::pair<int, double> &a, const std::pair<int, double> &b) { return a.first < b.first; });
    for (size_t k = 0; k < row.size(); k++) {
      if (k > 0 && row[k].first == row[k - 1].first) { fm_set_error(std::string(fn) + ": repeated column in row " + std::to_string(i)); return FM_EINVAL; }
      u.ci[rp[i] + k] = row[k].first;
      if (v) u.v[rp[i] + k] = row[k].second;
    }
  }
  u.nc = nc;
  u.set = true;
  return 0;
}

// MGSolDA::ReadMatrix (src/MGSolverDA.C:1785-2004): the pattern of A on level L from the caller (rows and columns in
// the reference DoF numbering of level L).  It must hold every coupling of the level's elements (checked at
// fm_system_init); extra entries stay zero.  NULL rowptr drops it (the pattern is built from the mesh again).
extern "C" int fm_system_set_pattern(fm_system *s, int L, const int64_t *rowptr, const int *colind) {
  if (!s || L < 0 || L >= (int)s->Dim.size()) return fm_einval(__func__);
  if (!rowptr) { s->u_pat[L] = UserCsr(); s->initialized = false; return FM_OK; }
  FM_CHECK(user_csr(__func__, s->Dim[L], s->Dim[L], rowptr, colind, nullptr, s->u_pat[L]));
  s->initialized = false;
  return FM_OK;
}

// MGSolDA::ReadRest / ReadProl (src/MGSolverDA.C:2236-2444, 2011-2230): which = 1: Rst[L] (coarse L <- fine L+1;
// fm_system_init multiplies its rows by the Dirichlet mask bc0 % 2 of the row's top-level DoF, as ReadRest does),
// which = 2: Prl[L] (fine L <- coarse L-1); reference numbering of both levels.  NULL rowptr drops it.
extern "C" int fm_system_set_transfer(fm_system *s, int which, int L, const int64_t *rowptr, const int *colind,
                                      const double *val) {
  if (!s || L < 0 || L >= (int)s->Dim.size() || (which != 1 && which != 2)) return fm_einval(__func__);
  const int top = (int)s->Dim.size() - 1;
  if ((which == 1 && L >= top) || (which == 2 && L == 0)) return fm_einval(__func__);
  UserCsr &u = which == 1 ? s->u_rst[L] : s->u_prl[L];
  if (!rowptr) { u = UserCsr(); s->initialized = false; return FM_OK; }
  if (!val) return fm_einval(__func__);
  FM_CHECK(user_csr(__func__, s->Dim[L], s->Dim[which == 1 ? L + 1 : L - 1], rowptr, colind, val, u));
  s->initialized = false;
  return FM_OK;
}

static void free_sell(DevSELL &q) {
  if (q.borrowed) {  // only the values are this matrix's own
    dfree(q.v); dfree(q.vf);
    q = DevSELL();
    return;
  }
  dfree(q.sl_off); dfree(q.sl_c32); dfree(q.sl_w); dfree(q.perm); dfree(q.c16); dfree(q.base); dfree(q.c32);
  dfree(q.sl_c8); dfree(q.c8); dfree(q.sl_c16); dfree(q.sl_c2); dfree(q.c2);
  dfree(q.pos); dfree(q.rbase); dfree(q.rslot); dfree(q.v); dfree(q.vf);
  q.nsl = 0; q.nslot = 0; q.n32 = 0; q.n8 = 0; q.n16 = 0; q.n2 = 0;
}

void fm_free_csr(DevCSR &c) {
  dfree(c.rp); dfree(c.ci); dfree(c.v); dfree(c.vf);
  free_sell(c.sell);
}
static void free_csr(DevCSR &c) { fm_free_csr(c); }

// SELL-64 copy of a CSR pattern in internal DoF order: internal row i holds CSR
// row ro[i], CSR column j becomes internal column cr[j]; sigma window 4096 rows
// (rows sorted by decreasing length inside each window), entries of a row
// sorted by internal column, columns per slice in the narrowest class it allows
// (2-bit lane-predicted, 8- or 16-bit offsets from a per-position base, absolute
// 32-bit columns).  Optional row groups (boundaries in
// internal rows) start their own slices, so a product can be restricted to
// the slices of chosen groups.
int fm_build_sell(hipStream_t st, DevCSR &c, const int *ro, const int *cr, const std::vector<int64_t> *groups) {
  // W: rows are length-sorted inside windows of W rows (FEMUS_SELL_SIGMA, default 4096; 64 = inside a slice);
  // column offsets no narrower than fm_set_sell_offsets allows
  const char *sg = getenv("FEMUS_SELL_SIGMA");
  const int C = 64, W = sg && atoi(sg) >= 64 ? atoi(sg) : 4096;
  const int64_t n = c.nr;
  SellHost h;
  std::string err;
  const int hrc = sell_build_host(n, c.h_rp, c.h_ci, ro, cr, groups, W, sell_minbits(), h, err);
  if (hrc) { fm_set_error(err); return hrc == -1 ? FM_EINVAL : FM_ESTATE; }
  c.sell.h_gslice = h.gslice;
  const int64_t nsl = h.nsl, nslot = h.nslot, n8 = h.n8, n16 = h.n16, n32 = h.n32, n2 = h.n2;
  const std::vector<int64_t> &off = h.off, &o8 = h.o8, &o16 = h.o16, &o32 = h.o32, &rbase = h.rbase, &rslot = h.rslot;
  const std::vector<int> &wd = h.wd, &perm = h.perm, &base = h.base, &c32 = h.c32;
  const std::vector<unsigned short> &c16 = h.c16, &pos = h.pos;
  const std::vector<unsigned char> &c8 = h.c8;
  DevSELL &q = c.sell;
  free_sell(q);
  q.nsl = nsl; q.nslot = nslot; q.n32 = n32; q.n8 = n8; q.n16 = n16; q.n2 = n2;
  FM_CHECK(dalloc(&q.sl_c2, nsl)); FM_CHECK(dalloc(&q.c2, h.c2.size()));
  FM_HIP(hipMemcpyAsync(q.sl_c2, h.o2.data(), nsl * sizeof(int64_t), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(q.c2, h.c2.data(), h.c2.size() * sizeof(unsigned), hipMemcpyHostToDevice, st));
  FM_CHECK(dalloc(&q.sl_off, nsl + 1)); FM_CHECK(dalloc(&q.sl_c32, nsl)); FM_CHECK(dalloc(&q.sl_w, nsl));
  FM_CHECK(dalloc(&q.sl_c8, nsl)); FM_CHECK(dalloc(&q.c8, c8.size())); FM_CHECK(dalloc(&q.sl_c16, nsl));
  FM_CHECK(dalloc(&q.perm, nsl * C)); FM_CHECK(dalloc(&q.c16, c16.size())); FM_CHECK(dalloc(&q.base, nslot / C));
  FM_CHECK(dalloc(&q.c32, c32.size())); FM_CHECK(dalloc(&q.pos, nslot)); FM_CHECK(dalloc(&q.v, nslot));
  FM_CHECK(dalloc(&q.rbase, n)); FM_CHECK(dalloc(&q.rslot, n));
  FM_CHECK(dalloc(&q.vf, nslot));
  FM_HIP(hipMemcpyAsync(q.sl_off, off.data(), (nsl + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(q.sl_c32, o32.data(), nsl * sizeof(int64_t), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(q.sl_c8, o8.data(), nsl * sizeof(int64_t), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(q.c8, c8.data(), c8.size(), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(q.sl_w, wd.data(), nsl * sizeof(int), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(q.perm, perm.data(), nsl * C * sizeof(int), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(q.c16, c16.data(), c16.size() * sizeof(unsigned short), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(q.sl_c16, o16.data(), nsl * sizeof(int64_t), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(q.base, base.data(), (nslot / C) * sizeof(int), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(q.c32, c32.data(), c32.size() * sizeof(int), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(q.pos, pos.data(), nslot * sizeof(unsigned short), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(q.rbase, rbase.data(), n * sizeof(int64_t), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(q.rslot, rslot.data(), n * sizeof(int64_t), hipMemcpyHostToDevice, st));
  FM_HIP(hipStreamSynchronize(st));
  return fm_sync_sell(st, c);
}

static SellMap sell_map(const DevSELL &q) {
  SellMap m;
  m.nsl = q.nsl; m.sl_off = q.sl_off; m.sl_w = q.sl_w; m.perm = q.perm; m.rbase = q.rbase; m.pos = q.pos;
  return m;
}

int fm_csr_from_sell(hipStream_t st, DevCSR &c) {
  if (c.sell.nsl == 0) return 0;
  if (!c.v) {  // first request: the CSR value arrays are allocated here
    FM_CHECK(dalloc(&c.v, c.nnz));
    FM_CHECK(dalloc(&c.vf, c.nnz));
  }
  if (!c.ci) {  // column indices of an assembled matrix live on the host after set-up; uploaded on request
    if ((int64_t)c.h_ci.size() != c.nnz) { fm_set_error("CSR columns missing"); return FM_ESTATE; }
    FM_CHECK(dalloc(&c.ci, c.nnz));
    FM_HIP(hipMemcpyAsync(c.ci, c.h_ci.data(), c.nnz * sizeof(int), hipMemcpyHostToDevice, st));
  }
  FM_CHECK(fmk::sell2csr(st, sell_map(c.sell), c.sell.v, c.v));
  FM_CHECK(fmk::cast_vals(st, c.nnz, c.v, c.vf));
  return 0;
}

// SELL slots of the diagonals (dslot, one per internal row) and, with eldof given, the element map of the level
// (elpos / elslot): CSR entries are translated to SELL positions through a temporary per-entry table
static int sell_slots(hipStream_t st, DevCSR &A, const int *ord, int nel, int nloc, const int *eldof,
                      const int *eldofi, unsigned short *elpos, int64_t *elslot, int64_t *dslot) {
  unsigned short *kof = nullptr;
  FM_CHECK(dalloc(&kof, A.nnz));
  int rc = fmk::sell_kof(st, sell_map(A.sell), kof);
  if (!rc && eldof) rc = fmk::elpos(st, nel, nloc, eldof, A.rp, A.ci, kof, elpos);
  if (!rc && eldof) rc = fmk::elslot(st, (int64_t)nel * nloc, eldofi, A.sell.rslot, elslot);
  if (!rc) rc = fmk::diag_slot(st, A.nr, A.rp, A.ci, ord, A.sell.rslot, kof, dslot);
  if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = FM_EHIP;
  dfree(kof);
  return rc;
}

// refresh SELL values (FP64 and FP32) from the CSR values
int fm_sync_sell(hipStream_t st, DevCSR &c) {
  if (c.sell.nsl == 0) return 0;
  if (!c.v) {  // no CSR values yet: the SELL values start at zero
    FM_CHECK(fmk::fill<double>(st, c.sell.nslot, c.sell.v, 0.0));
    return fmk::fill<float>(st, c.sell.nslot, c.sell.vf, 0.f);
  }
  return fmk::csr2sell(st, sell_map(c.sell), c.v, c.sell.v, c.sell.vf);
}

void fm_drop_graphs(fm_system *s) {
  for (auto &g : s->graphs) (void)hipGraphExecDestroy(g.exec);
  s->graphs.clear();
}

extern "C" int fm_system_destroy(fm_system *s) {
  if (!s) return fm_einval(__func__);
  fm_drop_graphs(s);
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  fm_krylov_free(s);
  for (auto &l : s->lev) fm_free_level(l);
  for (auto &l : s->alev) fm_free_level(l);
  dfree(s->ebuf); dfree(s->febuf);
  dfree(s->d_bc0); dfree(s->d_bc1); dfree(s->d_dof_top); dfree(s->xyz); dfree(s->x_old); dfree(s->vel); dfree(s->xbc);
  dfree(s->tnode);
  dfree(s->d_tab);
  s->ctx->nsys--;
  delete s;
  return FM_OK;
}

extern "C" int64_t fm_system_ndof(const fm_system *s, int L) {
  if (!s || L < 0 || L >= (int)s->Dim.size()) return -1;
  return s->Dim[L];
}
extern "C" int fm_system_node_dof(const fm_system *s, int L, int *nd) {
  if (!s || !nd || L < 0 || L >= (int)s->Dim.size()) return fm_einval(__func__);
  std::memcpy(nd, s->node_dof[L].data(), s->node_dof[L].size() * sizeof(int));
  return FM_OK;
}

// MGSolDA::GenBc_loop (src/MGSolverDA.C:786-862): interior nodes keep the
// default (bc_intern_read: all 1); nodes of top-level boundary faces take
// bc_read(xp).  Q1 variables are set on face vertices only.
extern "C" int fm_system_genbc(fm_system *s, fm_bc_fn fn, void *user) {
  if (!s || !fn) return fm_einval(__func__);
  const fm_mesh &m = *s->mesh;
  int top = m.nlev - 1, N = m.Ntop, dim = m.dim, nv = s->nq2 + s->nq1;
  int nb1 = dim == 2 ? 2 : 4;
  const std::vector<int> &dof = s->node_dof[top];
  std::fill(s->bc0.begin(), s->bc0.end(), 1);
  std::fill(s->bc1.begin(), s->bc1.end(), 1);
  int neum[8], flag[8];
  for (int e = 0; e < m.nel[top]; e++)
    for (int sd = 0; sd < m.nsides; sd++) {
      if (m.neigh[top][(size_t)e * m.nsides + sd] != -1) continue;
      for (int i = 0; i < m.nb2; i++) {
        int ln = dim == 2 ? fm_face2d[sd][i] : fm_face3d[sd][i];
        int k = m.conn[top][(size_t)e * m.n2 + ln];
        double xp[3] = {0, 0, 0};
        for (int d = 0; d < dim; d++) xp[d] = m.xyz[k + (size_t)d * N];
        for (int iv = 0; iv < nv; iv++) { neum[iv] = 1; flag[iv] = 1; }
        fn(xp, neum, flag, user);
        for (int iv = 0; iv < nv; iv++) {
          if (iv >= s->nq2 && i >= nb1) continue;
          int kd = dof[k + (size_t)iv * N];
          s->bc0[kd] = neum[iv];
          s->bc1[kd] = flag[iv];
        }
      }
    }
  s->initialized = false;
  return FM_OK;
}

extern "C" int fm_system_get_bc(const fm_system *s, int *bc0, int *bc1) {
  if (!s) return fm_einval(__func__);
  if (bc0) std::memcpy(bc0, s->bc0.data(), s->bc0.size() * sizeof(int));
  if (bc1) std::memcpy(bc1, s->bc1.data(), s->bc1.size() * sizeof(int));
  return FM_OK;
}

static int upload_csr(hipStream_t st, DevCSR &c, int64_t nr, int64_t nc, std::vector<int64_t> &rp, std::vector<int> &ci,
                      std::vector<double> *v, const std::vector<int> &ro, const std::vector<int> &cr,
                      const std::vector<int64_t> *groups = nullptr) {
  free_csr(c);
  c.nr = nr; c.nc = nc; c.nnz = rp[nr];
  FM_CHECK(dalloc(&c.rp, nr + 1));
  FM_CHECK(dalloc(&c.ci, c.nnz));
  FM_HIP(hipMemcpyAsync(c.rp, rp.data(), (nr + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(c.ci, ci.data(), c.nnz * sizeof(int), hipMemcpyHostToDevice, st));
  if (v) {  // fixed values (transfers); an assembled matrix gets its CSR values only on request (fm_csr_from_sell)
    FM_CHECK(dalloc(&c.v, c.nnz));
    FM_CHECK(dalloc(&c.vf, c.nnz));
    FM_HIP(hipMemcpyAsync(c.v, v->data(), c.nnz * sizeof(double), hipMemcpyHostToDevice, st));
    FM_CHECK(fmk::cast_vals(st, c.nnz, c.v, c.vf));
  }
  FM_HIP(hipStreamSynchronize(st));
  c.h_rp.swap(rp);
  c.h_ci.swap(ci);
  if (v) c.h_v.swap(*v);
  return fm_build_sell(st, c, ro.data(), cr.data(), groups);
}

template <typename T> static inline CsrView<T> view(const DevCSR &c) { return fm_view<T>(c); }

int fm_alloc_level_vectors(Level &l, bool top) {
  dfree(l.x); dfree(l.b); dfree(l.r); dfree(l.t); dfree(l.d); dfree(l.d2); dfree(l.dinv); dfree(l.res);
  dfree(l.xf); dfree(l.bf); dfree(l.rf); dfree(l.tf); dfree(l.df); dfree(l.d2f); dfree(l.dinvf); dfree(l.nrm);
  dfree(l.sb); dfree(l.sx);
  FM_CHECK(dalloc(&l.x, l.n)); FM_CHECK(dalloc(&l.b, l.n)); FM_CHECK(dalloc(&l.r, l.n)); FM_CHECK(dalloc(&l.t, l.n));
  FM_CHECK(dalloc(&l.d, l.n)); FM_CHECK(dalloc(&l.d2, l.n)); FM_CHECK(dalloc(&l.dinv, l.n)); FM_CHECK(dalloc(&l.res, l.n));
  FM_CHECK(dalloc(&l.xf, l.n)); FM_CHECK(dalloc(&l.bf, l.n)); FM_CHECK(dalloc(&l.rf, l.n)); FM_CHECK(dalloc(&l.tf, l.n));
  FM_CHECK(dalloc(&l.df, l.n)); FM_CHECK(dalloc(&l.d2f, l.n)); FM_CHECK(dalloc(&l.dinvf, l.n));
  FM_CHECK(dalloc(&l.nrm, 8));
  if (top) { FM_CHECK(dalloc(&l.sb, l.n)); FM_CHECK(dalloc(&l.sx, l.n)); }
  return 0;
}

void fm_free_level(Level &l) {
  free_csr(l.A); free_csr(l.R); free_csr(l.P);
  dfree(l.x); dfree(l.b); dfree(l.r); dfree(l.t); dfree(l.d); dfree(l.d2); dfree(l.dinv);
  dfree(l.xf); dfree(l.bf); dfree(l.rf); dfree(l.tf); dfree(l.df); dfree(l.d2f); dfree(l.dinvf);
  dfree(l.conn); dfree(l.eldof); dfree(l.eldofi); dfree(l.elpos); dfree(l.elslot); dfree(l.bmask); dfree(l.color_elems);
  dfree(l.coef);
  dfree(l.ord); dfree(l.rank);
  dfree(l.Ainv); dfree(l.Ainvf); dfree(l.vinv); dfree(l.vinvf); dfree(l.vwt); dfree(l.vwtf); dfree(l.nrm); dfree(l.res);
  dfree(l.sb); dfree(l.sx); dfree(l.vone); dfree(l.cslist); dfree(l.dslot);
  fm_free_egather(l.eg);
}

// Per colour: the SELL slices of the row groups whose signature contains the
// colour (the rows that colour's elements touch), concatenated in l.cslist.
static int colour_slices(hipStream_t st, Level &l) {
  dfree(l.cslist);
  l.cs_off.assign(1, 0);
  const std::vector<int64_t> &gs = l.A.sell.h_gslice;
  if (l.h_gmask.empty() || gs.size() != l.h_gmask.size() + 1) return 0;
  int all = 0;
  for (int g : l.h_gmask) all |= g;
  if (all == 0) return 0;
  std::vector<int> list;
  for (int c = 0; c < 8; c++) {
    if (all & (1 << c))
      for (size_t g = 0; g < l.h_gmask.size(); g++)
        if (l.h_gmask[g] & (1 << c))
          for (int64_t sl = gs[g]; sl < gs[g + 1]; sl++) list.push_back((int)sl);
    l.cs_off.push_back((int64_t)list.size());
  }
  FM_CHECK(dalloc(&l.cslist, list.size()));
  FM_HIP(hipMemcpyAsync(l.cslist, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice, st));
  FM_HIP(hipStreamSynchronize(st));
  return 0;
}

// every coupling of the level's elements must be an entry of the caller's pattern (the assembly writes through
// the element -> entry map and would drop what the pattern lacks)
static int check_user_pattern(const fm_system *s, int L) {
  const fm_mesh &m = *s->mesh;
  const UserCsr &u = s->u_pat[L];
  std::vector<int> ed(s->nloc);
  for (int e = 0; e < m.nel[L]; e++) {
    fm_elem_dofs(m, L, e, s->nq2, s->nq1, s->node_dof[L], ed.data());
    for (int a = 0; a < s->nloc; a++)
      for (int b = 0; b < s->nloc; b++) {
        const int *c0 = u.ci.data() + u.rp[ed[a]], *c1 = u.ci.data() + u.rp[ed[a] + 1];
        if (!std::binary_search(c0, c1, ed[b])) {
          fm_set_error("fm_system_init: the caller's pattern of level " + std::to_string(L) + " lacks entry (" +
                       std::to_string(ed[a]) + ", " + std::to_string(ed[b]) + ") of element " + std::to_string(e));
          return FM_EINVAL;
        }
      }
  }
  return 0;
}

// ReadRest's row mask (src/MGSolverDA.C:2422-2424): rows of a coarse DoF whose top-level DoF is Dirichlet vanish
static void mask_user_rest(const fm_system *s, int L, const std::vector<int64_t> &rp, std::vector<double> &v) {
  const fm_mesh &m = *s->mesh;
  const int N = m.Ntop, top = m.nlev - 1, nv = s->nq2 + s->nq1;
  for (int iv = 0; iv < nv; iv++)
    for (int k = 0; k < N; k++) {
      const int row = s->node_dof[L][k + (size_t)iv * N];
      if (row < 0) continue;
      const int td = s->node_dof[top][k + (size_t)iv * N];
      if (td >= 0 && s->bc0[td] % 2 == 0)
        for (int64_t q = rp[row]; q < rp[row + 1]; q++) v[q] = 0.;
    }
}

extern "C" int fm_system_init(fm_system *s) {
  if (!s) return fm_einval(__func__);
  fm_ctx *c = s->ctx;
  FM_HIP(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  const fm_mesh &m = *s->mesh;
  int nlev = m.nlev, top = nlev - 1, N = m.Ntop, nv = s->nq2 + s->nq1, nloc = s->nloc;
  // top-level maps, coordinates and BC on the device
  dfree(s->d_bc0); dfree(s->d_bc1); dfree(s->d_dof_top); dfree(s->xyz); dfree(s->x_old);
  int64_t nt = s->Dim[top];
  FM_CHECK(dalloc(&s->d_bc0, nt)); FM_CHECK(dalloc(&s->d_bc1, nt));
  FM_CHECK(dalloc(&s->d_dof_top, (size_t)nv * N));
  FM_CHECK(dalloc(&s->xyz, (size_t)m.dim * N));
  FM_CHECK(dalloc(&s->x_old, nt));
  FM_HIP(hipMemcpyAsync(s->d_bc0, s->bc0.data(), nt * sizeof(int), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(s->d_bc1, s->bc1.data(), nt * sizeof(int), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(s->d_dof_top, s->node_dof[top].data(), (size_t)nv * N * sizeof(int), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(s->xyz, m.xyz.data(), m.xyz.size() * sizeof(double), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemsetAsync(s->x_old, 0, nt * sizeof(double), st));
  if (!s->d_tab) {
    FeTab t;
    build_tables(m.dim, t);
    FM_CHECK(dalloc(&s->d_tab, 1));
    FM_HIP(hipMemcpy(s->d_tab, &t, sizeof(FeTab), hipMemcpyHostToDevice));
  }
  for (int L = 0; L < nlev; L++) {
    Level &l = s->lev[L];
    l.n = s->Dim[L];
    l.nel = m.nel[L];
    // colour signature of every node (coupled systems only: the coloured Vanka sweep uses it)
    std::vector<unsigned char> nmask;
    if (s->nq1 > 0) fm_node_colours(m, L, nmask);
    std::string err;
    if (fm_level_order(m, L, nv, s->node_dof[L], l.n, l.h_ord, l.h_rank, s->nq1 > 0 ? &nmask : nullptr, l.h_goff,
                       l.h_gmask, err)) {
      fm_set_error(err);
      return FM_ESTATE;
    }
    dfree(l.ord); dfree(l.rank);
    FM_CHECK(dalloc(&l.ord, l.n)); FM_CHECK(dalloc(&l.rank, l.n));
    FM_HIP(hipMemcpyAsync(l.ord, l.h_ord.data(), l.n * sizeof(int), hipMemcpyHostToDevice, st));
    FM_HIP(hipMemcpyAsync(l.rank, l.h_rank.data(), l.n * sizeof(int), hipMemcpyHostToDevice, st));
  }
  for (int L = 0; L < nlev; L++) {
    Level &l = s->lev[L];
    // pattern: the caller's (ReadMatrix) or the element couplings of the mesh
    std::vector<int64_t> rp;
    std::vector<int> ci;
    if (s->u_pat[L].set) {
      FM_CHECK(check_user_pattern(s, L));
      rp = s->u_pat[L].rp;
      ci = s->u_pat[L].ci;
    } else {
      fm_pattern(m, L, s->nq2, s->nq1, s->node_dof[L], rp, ci);
    }
    FM_CHECK(upload_csr(st, l.A, l.n, l.n, rp, ci, nullptr, l.h_ord, l.h_rank, &l.h_goff));
    FM_CHECK(colour_slices(st, l));
    // transfers: the caller's (ReadRest with the Dirichlet row mask, ReadProl) or built from the refinement
    if (L < top) {
      std::vector<int64_t> trp; std::vector<int> tci; std::vector<double> tv;
      if (s->u_rst[L].set) {
        trp = s->u_rst[L].rp; tci = s->u_rst[L].ci; tv = s->u_rst[L].v;
        mask_user_rest(s, L, trp, tv);
      } else {
        fm_transfer(m, L, s->nq2, s->nq1, true, s->node_dof[L], s->node_dof[L + 1], &s->node_dof[top], &s->bc0, trp, tci, tv);
      }
      FM_CHECK(upload_csr(st, l.R, l.n, s->Dim[L + 1], trp, tci, &tv, l.h_ord, s->lev[L + 1].h_rank));
    }
    if (L > 0) {
      std::vector<int64_t> trp; std::vector<int> tci; std::vector<double> tv;
      if (s->u_prl[L].set) {
        trp = s->u_prl[L].rp; tci = s->u_prl[L].ci; tv = s->u_prl[L].v;
      } else {
        fm_transfer(m, L, s->nq2, s->nq1, false, s->node_dof[L], s->node_dof[L - 1], nullptr, nullptr, trp, tci, tv);
      }
      FM_CHECK(upload_csr(st, l.P, l.n, s->Dim[L - 1], trp, tci, &tv, l.h_ord, s->lev[L - 1].h_rank));
    }
    // vectors
    FM_CHECK(fm_alloc_level_vectors(l, L == top));
    FM_HIP(hipMemsetAsync(l.x, 0, l.n * sizeof(double), st));
    FM_HIP(hipMemsetAsync(l.b, 0, l.n * sizeof(double), st));
    FM_HIP(hipMemsetAsync(l.res, 0, l.n * sizeof(double), st));
    // elements: connectivity, DoFs, colours, boundary masks, CSR positions
    std::vector<int> eld((size_t)l.nel * nloc);
    std::vector<unsigned char> bm(l.nel, 0);
    std::vector<int> col(l.nel);
    for (int e = 0; e < l.nel; e++) {
      fm_elem_dofs(m, L, e, s->nq2, s->nq1, s->node_dof[L], &eld[(size_t)e * nloc]);
      for (int sd = 0; sd < m.nsides; sd++) if (m.neigh[L][(size_t)e * m.nsides + sd] == -1) bm[e] |= (unsigned char)(1 << sd);
      col[e] = fm_elem_colour(m, L, e);
    }
    l.ncolor = fm_ncolours(m, L);
    l.color_off.assign(l.ncolor + 1, 0);
    for (int e = 0; e < l.nel; e++) l.color_off[col[e] + 1]++;
    for (int k = 0; k < l.ncolor; k++) l.color_off[k + 1] += l.color_off[k];
    std::vector<int> ce(l.nel), fillp(l.color_off.begin(), l.color_off.end() - 1);
    for (int e = 0; e < l.nel; e++) ce[fillp[col[e]]++] = e;
    std::vector<int> eldi(eld.size());
    for (size_t q = 0; q < eld.size(); q++) eldi[q] = eld[q] >= 0 ? l.h_rank[eld[q]] : -1;
    dfree(l.conn); dfree(l.eldof); dfree(l.eldofi); dfree(l.elpos); dfree(l.elslot); dfree(l.bmask);
    dfree(l.color_elems);
    fm_free_egather(l.eg);  // built from the slot map and the SELL layout made here
    FM_CHECK(dalloc(&l.conn, (size_t)l.nel * m.n2));
    FM_CHECK(dalloc(&l.eldof, (size_t)l.nel * nloc));
    FM_CHECK(dalloc(&l.eldofi, (size_t)l.nel * nloc));
    FM_CHECK(dalloc(&l.elpos, (size_t)l.nel * nloc * nloc + FM_ELMAP_PAD));
    FM_CHECK(dalloc(&l.elslot, (size_t)l.nel * nloc + FM_ELMAP_PAD));
    FM_HIP(hipMemsetAsync(l.elpos + (size_t)l.nel * nloc * nloc, 0xFF, FM_ELMAP_PAD * sizeof(unsigned short), st));
    FM_HIP(hipMemsetAsync(l.elslot + (size_t)l.nel * nloc, 0xFF, FM_ELMAP_PAD * sizeof(int64_t), st));
    FM_CHECK(dalloc(&l.bmask, l.nel));
    FM_CHECK(dalloc(&l.color_elems, l.nel));
    FM_HIP(hipMemcpyAsync(l.conn, m.conn[L].data(), m.conn[L].size() * sizeof(int), hipMemcpyHostToDevice, st));
    FM_HIP(hipMemcpyAsync(l.eldof, eld.data(), eld.size() * sizeof(int), hipMemcpyHostToDevice, st));
    FM_HIP(hipMemcpyAsync(l.eldofi, eldi.data(), eldi.size() * sizeof(int), hipMemcpyHostToDevice, st));
    FM_HIP(hipMemcpyAsync(l.bmask, bm.data(), bm.size(), hipMemcpyHostToDevice, st));
    FM_HIP(hipMemcpyAsync(l.color_elems, ce.data(), ce.size() * sizeof(int), hipMemcpyHostToDevice, st));
    dfree(l.dslot);
    FM_CHECK(dalloc(&l.dslot, l.n));
    FM_CHECK(sell_slots(st, l.A, l.ord, l.nel, nloc, l.eldof, l.eldofi, l.elpos, l.elslot, l.dslot));
    dfree(l.A.ci);  // the solver runs on the SELL copy; CSR requests re-upload the columns (fm_csr_from_sell)
    FM_HIP(hipStreamSynchronize(st));
  }
  fm_drop_graphs(s);
  for (auto &l : s->alev) fm_free_level(l);
  s->alev.clear();
  s->a_assembled.clear();
  s->adj_ready = false;
  s->initialized = true;
  s->mg_ready = false;
  return FM_OK;
}

extern "C" int64_t fm_system_csr_nnz(const fm_system *s, int which, int L) {
  if (!s || L < 0 || L >= (int)s->lev.size() || which < 0 || which > 2) return -1;
  const Level &l = s->lev[L];
  const DevCSR &c = which == 0 ? l.A : (which == 1 ? l.R : l.P);
  return c.nnz;
}

extern "C" int fm_system_sell_info(const fm_system *s, int which, int L, int64_t *info) {
  if (!s || !info || L < 0 || L >= (int)s->lev.size() || !s->initialized || which < 0 || which > 2) return fm_einval(__func__);
  const Level &l = s->lev[L];
  const DevCSR &c = which == 0 ? l.A : (which == 1 ? l.R : l.P);
  info[0] = c.sell.nsl; info[1] = c.sell.nslot; info[2] = c.sell.n8; info[3] = c.sell.n16; info[4] = c.sell.n32;
  info[5] = c.nnz; info[6] = c.sell.n2;
  return FM_OK;
}

extern "C" int fm_system_csr_get(const fm_system *s, int which, int L, int64_t *rowptr, int *colind, double *val) {
  if (!s || L < 0 || L >= (int)s->lev.size() || !s->initialized || which < 0 || which > 2) return fm_einval(__func__);
  const Level &l = s->lev[L];
  const DevCSR &c = which == 0 ? l.A : (which == 1 ? l.R : l.P);
  if (c.nr == 0) return fm_einval(__func__);
  FM_HIP(hipSetDevice(s->ctx->device));
  if (which == 0 && val) FM_CHECK(fm_csr_from_sell(s->ctx->stream, const_cast<DevCSR &>(c)));
  FM_HIP(hipStreamSynchronize(s->ctx->stream));
  if (rowptr) FM_HIP(hipMemcpy(rowptr, c.rp, (c.nr + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
  if (colind) {
    if (c.ci) FM_HIP(hipMemcpy(colind, c.ci, c.nnz * sizeof(int), hipMemcpyDeviceToHost));
    else std::copy(c.h_ci.begin(), c.h_ci.end(), colind);
  }
  if (val) FM_HIP(hipMemcpy(val, c.v, c.nnz * sizeof(double), hipMemcpyDeviceToHost));
  return FM_OK;
}

static int check_level(const fm_system *s, int L) {
  if (!s || !s->initialized || L < 0 || L >= (int)s->lev.size()) { fm_set_error("system not initialised or bad level"); return FM_ESTATE; }
  return 0;
}

extern "C" int fm_system_set_xold(fm_system *s, const double *x) {
  FM_CHECK(check_level(s, 0));
  if (!x) return fm_einval(__func__);
  FM_HIP(hipMemcpyAsync(s->x_old, x, s->Dim.back() * sizeof(double), hipMemcpyHostToDevice, s->ctx->stream));
  FM_HIP(hipStreamSynchronize(s->ctx->stream));
  return FM_OK;
}
// Level vectors live in internal DoF order; the ABI exchanges them in the
// reference numbering: out[dof] = v[rank[dof]], v[i] = in[ord[i]].
static int vec_to_host(const fm_system *s, int L, const double *v, double *host) {
  const Level &l = s->lev[L];
  hipStream_t st = s->ctx->stream;
  FM_CHECK(fmk::gather(st, l.n, l.rank, v, l.t));
  FM_HIP(hipMemcpyAsync(host, l.t, l.n * sizeof(double), hipMemcpyDeviceToHost, st));
  FM_HIP(hipStreamSynchronize(st));
  return 0;
}
static int vec_from_host(fm_system *s, int L, const double *host, double *v) {
  Level &l = s->lev[L];
  hipStream_t st = s->ctx->stream;
  FM_HIP(hipMemcpyAsync(l.t, host, l.n * sizeof(double), hipMemcpyHostToDevice, st));
  FM_CHECK(fmk::gather(st, l.n, l.ord, l.t, v));
  FM_HIP(hipStreamSynchronize(st));
  return 0;
}
extern "C" int fm_system_get_x(const fm_system *s, int L, double *x) {
  FM_CHECK(check_level(s, L));
  if (!x) return fm_einval(__func__);
  return vec_to_host(s, L, s->lev[L].x, x);
}
extern "C" int fm_system_set_x(fm_system *s, int L, const double *x) {
  FM_CHECK(check_level(s, L));
  if (!x) return fm_einval(__func__);
  return vec_from_host(s, L, x, s->lev[L].x);
}
extern "C" int fm_system_get_rhs(const fm_system *s, int L, double *b) {
  FM_CHECK(check_level(s, L));
  if (!b) return fm_einval(__func__);
  return vec_to_host(s, L, s->lev[L].b, b);
}
extern "C" int fm_system_set_rhs(fm_system *s, int L, const double *b) {
  FM_CHECK(check_level(s, L));
  if (!b) return fm_einval(__func__);
  return vec_from_host(s, L, b, s->lev[L].b);
}
// Device-pointer variants: the caller's vector (reference numbering) lives in device memory of the context's
// device and is gathered to / from the level's internal order on the device, with no host copy
static int check_dev_ptr(const char *fn, const void *p) {
  if (!p) return fm_einval(fn);
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) != hipSuccess || (at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged)) {
    (void)hipGetLastError();
    fm_set_error(std::string(fn) + ": not a device pointer");
    return FM_EINVAL;
  }
  return 0;
}
static int vec_to_dev(const fm_system *s, int L, const double *v, double *dev) {
  FM_HIP(hipSetDevice(s->ctx->device));
  FM_CHECK(fmk::gather(s->ctx->stream, s->lev[L].n, s->lev[L].rank, v, dev));
  FM_HIP(hipStreamSynchronize(s->ctx->stream));
  return 0;
}
static int vec_from_dev(fm_system *s, int L, const double *dev, double *v) {
  FM_HIP(hipSetDevice(s->ctx->device));
  FM_CHECK(fmk::gather(s->ctx->stream, s->lev[L].n, s->lev[L].ord, dev, v));
  FM_HIP(hipStreamSynchronize(s->ctx->stream));
  return 0;
}
extern "C" int fm_system_get_x_dev(const fm_system *s, int L, double *x_dev) {
  FM_CHECK(check_level(s, L));
  FM_CHECK(check_dev_ptr(__func__, x_dev));
  return vec_to_dev(s, L, s->lev[L].x, x_dev);
}
extern "C" int fm_system_set_x_dev(fm_system *s, int L, const double *x_dev) {
  FM_CHECK(check_level(s, L));
  FM_CHECK(check_dev_ptr(__func__, x_dev));
  return vec_from_dev(s, L, x_dev, s->lev[L].x);
}
extern "C" int fm_system_get_rhs_dev(const fm_system *s, int L, double *b_dev) {
  FM_CHECK(check_level(s, L));
  FM_CHECK(check_dev_ptr(__func__, b_dev));
  return vec_to_dev(s, L, s->lev[L].b, b_dev);
}
extern "C" int fm_system_set_rhs_dev(fm_system *s, int L, const double *b_dev) {
  FM_CHECK(check_level(s, L));
  FM_CHECK(check_dev_ptr(__func__, b_dev));
  return vec_from_dev(s, L, b_dev, s->lev[L].b);
}
extern "C" int fm_get_res(const fm_system *s, int L, double *r) {
  FM_CHECK(check_level(s, L));
  if (!r) return fm_einval(__func__);
  return vec_to_host(s, L, s->lev[L].res, r);
}

extern "C" int fm_system_set_elem_coef(fm_system *s, int L, const double *coef) {
  FM_CHECK(check_level(s, L));
  Level &l = s->lev[L];
  if (!coef) { dfree(l.coef); return FM_OK; }
  if (!l.coef) FM_CHECK(dalloc(&l.coef, l.nel));
  FM_HIP(hipMemcpyAsync(l.coef, coef, l.nel * sizeof(double), hipMemcpyHostToDevice, s->ctx->stream));
  FM_HIP(hipStreamSynchronize(s->ctx->stream));
  return FM_OK;
}
// Dirichlet data of the boundary-face penalty RHS (top-level DoFs, reference
// numbering).  The reference takes it from x_old (which also carries the
// advection field); a separate vector lets a parameter act on the boundary
// data alone.  nullptr restores x_old.
extern "C" int fm_system_set_bcdata(fm_system *s, const double *g) {
  FM_CHECK(check_level(s, 0));
  if (!g) { dfree(s->xbc); return FM_OK; }
  if (!s->xbc) FM_CHECK(dalloc(&s->xbc, s->Dim.back()));
  FM_HIP(hipMemcpyAsync(s->xbc, g, s->Dim.back() * sizeof(double), hipMemcpyHostToDevice, s->ctx->stream));
  FM_HIP(hipStreamSynchronize(s->ctx->stream));
  return FM_OK;
}

// Boundary-face penalty RHS of the top level for Dirichlet data g (reference
// numbering in and out), matrix untouched: the linear map g -> b_bd(g).
static int bc_rhs(fm_system *s, const void *params, const double *g, double *rhs) {
  FM_CHECK(check_level(s, 0));
  if (!params || !g || !rhs) return fm_einval(__func__);
  const fm_mesh &m = *s->mesh;
  int top = (int)s->lev.size() - 1;
  Level &l = s->lev[top];
  hipStream_t st = s->ctx->stream;
  FM_HIP(hipMemcpyAsync(l.t, g, l.n * sizeof(double), hipMemcpyHostToDevice, st));
  FM_CHECK(fmk::fill<double>(st, l.n, l.d, 0.0));
  const bool heat = s->nq2 == 1 && s->nq1 == 0;
  if (!heat && !(s->nq2 == m.dim && s->nq1 == 1)) return fm_einval(__func__);
  for (int k = 0; k < l.ncolor; k++) {
    int n = l.color_off[k + 1] - l.color_off[k];
    const int *el = l.color_elems + l.color_off[k];
    if (heat) {
      const fm_heat_params *p = (const fm_heat_params *)params;
      if (p->dt == 0.) return fm_einval(__func__);
      HeatArgs a;
      a.alpha0 = p->alpha0; a.rhocp = p->rhocp; a.dt = p->dt; a.tc = p->time_coeff; a.source = p->source;
      a.upwind = p->upwind;
      for (int d = 0; d < 3; d++) a.vel_const[d] = p->vel_const[d];
      FM_CHECK(fmk::heat_assemble(st, m.dim, n, el, l.conn, l.eldof, l.elpos, l.elslot, l.bmask, s->xyz, m.Ntop, s->d_dof_top,
                                  s->d_bc0, s->d_bc1, s->x_old, l.t, l.coef, s->vel, s->d_tab, a, 1, 0, 1, nullptr, l.d));
    } else {
      const fm_ns_params *p = (const fm_ns_params *)params;
      if (p->dt == 0.) return fm_einval(__func__);
      NsArgs a;
      a.IRe = p->IRe; a.dt = p->dt; a.tc = p->time_coeff; a.upwind = p->upwind; a.IFr = p->IFr;
      for (int d = 0; d < 3; d++) { a.dirg[d] = p->dirg[d]; a.src[d] = p->src[d]; a.buoy[d] = 0.; }
      a.Tn = nullptr;
      FM_CHECK(fmk::ns_assemble(st, m.dim, n, el, l.conn, l.eldof, l.elpos, l.elslot, l.bmask, s->xyz, m.Ntop, s->d_dof_top,
                                s->d_bc0, s->d_bc1, s->x_old, l.t, l.coef, s->d_tab, a, 1, 0, 1, nullptr, l.d));
    }
  }
  FM_HIP(hipMemcpyAsync(rhs, l.d, l.n * sizeof(double), hipMemcpyDeviceToHost, st));
  FM_HIP(hipStreamSynchronize(st));
  return FM_OK;
}

extern "C" int fm_heat_bc_rhs(fm_system *s, const fm_heat_params *p, const double *g, double *rhs) {
  if (!s || s->nq2 != 1 || s->nq1 != 0) return fm_einval(__func__);
  return bc_rhs(s, p, g, rhs);
}
extern "C" int fm_ns_bc_rhs(fm_system *s, const fm_ns_params *p, const double *g, double *rhs) {
  if (!s || !s->mesh || s->nq2 != s->mesh->dim || s->nq1 != 1) return fm_einval(__func__);
  return bc_rhs(s, p, g, rhs);
}

// Boussinesq coupling: body force coef * T in the NS momentum RHS, T a nodal
// temperature field (top-level nodes); NULL T removes it.
extern "C" int fm_ns_set_buoyancy(fm_system *s, const double *T_node, const double coef[3]) {
  FM_CHECK(check_level(s, 0));
  if (!s->mesh || s->nq2 != s->mesh->dim || s->nq1 != 1) return fm_einval(__func__);
  if (!T_node) { dfree(s->tnode); for (int d = 0; d < 3; d++) s->buoy[d] = 0.; return FM_OK; }
  if (!coef) return fm_einval(__func__);
  const size_t N = (size_t)s->mesh->Ntop;
  if (!s->tnode) FM_CHECK(dalloc(&s->tnode, N));
  FM_HIP(hipMemcpyAsync(s->tnode, T_node, N * sizeof(double), hipMemcpyHostToDevice, s->ctx->stream));
  FM_HIP(hipStreamSynchronize(s->ctx->stream));
  for (int d = 0; d < 3; d++) s->buoy[d] = coef[d];
  return FM_OK;
}

extern "C" int fm_system_set_velocity(fm_system *s, const double *vel) {
  FM_CHECK(check_level(s, 0));
  if (!vel) { dfree(s->vel); return FM_OK; }
  size_t n = (size_t)s->mesh->dim * s->mesh->Ntop;
  if (!s->vel) FM_CHECK(dalloc(&s->vel, n));
  FM_HIP(hipMemcpyAsync(s->vel, vel, n * sizeof(double), hipMemcpyHostToDevice, s->ctx->stream));
  FM_HIP(hipStreamSynchronize(s->ctx->stream));
  return FM_OK;
}

// Coupling fields from another system's top-level solution, gathered on the device (both systems on the same
// context and mesh): which 0 = the temperature of a heat system as dst's buoyancy field (coef as in
// fm_ns_set_buoyancy), 1 = the velocity of an NS system as dst's advection field.
extern "C" int fm_system_couple(fm_system *dst, const fm_system *src, int which, const double coef[3]) {
  FM_CHECK(check_level(dst, 0));
  FM_CHECK(check_level(src, 0));
  if (dst->ctx != src->ctx || dst->mesh != src->mesh || which < 0 || which > 1) return fm_einval(__func__);
  const fm_mesh &m = *dst->mesh;
  const size_t N = (size_t)m.Ntop;
  const int top = m.nlev - 1;
  const Level &ls = src->lev[top];
  hipStream_t st = dst->ctx->stream;
  FM_HIP(hipSetDevice(dst->ctx->device));
  if (which == 0) {
    if (!coef || dst->nq2 != m.dim || dst->nq1 != 1 || src->nq2 < 1) return fm_einval(__func__);
    if (!dst->tnode) FM_CHECK(dalloc(&dst->tnode, N));
    FM_CHECK(fmk::node_gather(st, (int64_t)N, 1, src->d_dof_top, ls.rank, ls.x, dst->tnode));
    for (int d = 0; d < 3; d++) dst->buoy[d] = coef[d];
  } else {
    if (src->nq2 < m.dim) return fm_einval(__func__);
    if (!dst->vel) FM_CHECK(dalloc(&dst->vel, (size_t)m.dim * N));
    FM_CHECK(fmk::node_gather(st, (int64_t)N, m.dim, src->d_dof_top, ls.rank, ls.x, dst->vel));
  }
  FM_HIP(hipStreamSynchronize(st));
  return FM_OK;
}

// Time-step advance on the device: x_old = x (top level, reference numbering); *rel_change (if non-null)
// = |x - x_old| / |x| with the x_old before the copy.
extern "C" int fm_system_advance(fm_system *s, double *rel_change) {
  FM_CHECK(check_level(s, 0));
  const int top = (int)s->lev.size() - 1;
  Level &l = s->lev[top];
  fm_ctx *c = s->ctx;
  hipStream_t st = c->stream;
  FM_HIP(hipSetDevice(c->device));
  FM_CHECK(fmk::gather(st, l.n, l.rank, l.x, l.t));  // x in the reference numbering
  if (rel_change) {
    FM_HIP(hipMemcpyAsync(l.d, l.t, l.n * sizeof(double), hipMemcpyDeviceToDevice, st));
    FM_CHECK(fmk::axpby<double>(st, l.n, 1.0, l.d, -1.0, s->x_old));
    FM_CHECK(fmk::dot_dev<double>(st, l.n, l.d, l.d, c->scratch, c->scratch + 2048, 1));
    FM_CHECK(fmk::dot_dev<double>(st, l.n, l.t, l.t, c->scratch, c->scratch + 2049, 1));
    FM_HIP(hipMemcpyAsync(c->h_scratch, c->scratch + 2048, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  FM_HIP(hipMemcpyAsync(s->x_old, l.t, l.n * sizeof(double), hipMemcpyDeviceToDevice, st));
  FM_HIP(hipStreamSynchronize(st));
  if (rel_change) *rel_change = c->h_scratch[0] / std::max(c->h_scratch[1], 1e-300);
  return FM_OK;
}

void fm_free_egather(EGather &g) {
  for (void *p : {(void *)g.lseg, (void *)g.bbase, (void *)g.seg, (void *)g.cnt, (void *)g.cst, (void *)g.cidx})
    if (p) (void)hipFree(p);
  g = EGather();
}

template <typename T>
static int eg_upload(hipStream_t st, T **d, const std::vector<T> &h) {
  if (hipMalloc((void **)d, std::max<size_t>(h.size(), 1) * sizeof(T)) != hipSuccess) return -1;
  if (!h.empty() && hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st) != hipSuccess) return -1;
  return 0;
}

int fm_build_egather(hipStream_t st, Level &l, int nl, size_t max_lds, EGather &g, std::string &err) {
  const DevSELL &S = l.A.sell;
  const int64_t nrow_el = (int64_t)l.nel * nl;
  std::vector<unsigned short> pos((size_t)nrow_el * nl);
  std::vector<int64_t> slot0(nrow_el), off(S.nsl + 1);
  std::vector<int> wd(S.nsl);
  if (hipMemcpyAsync(pos.data(), l.elpos, pos.size() * sizeof(unsigned short), hipMemcpyDeviceToHost, st) ||
      hipMemcpyAsync(slot0.data(), l.elslot, slot0.size() * sizeof(int64_t), hipMemcpyDeviceToHost, st) ||
      hipMemcpyAsync(off.data(), S.sl_off, off.size() * sizeof(int64_t), hipMemcpyDeviceToHost, st) ||
      hipMemcpyAsync(wd.data(), S.sl_w, wd.size() * sizeof(int), hipMemcpyDeviceToHost, st) || hipStreamSynchronize(st)) {
    err = "element-buffer gather: download of the slot map failed";
    return -1;
  }
  if (off.back() != S.nslot) { err = "element-buffer gather: slice offsets do not cover the slots"; return -1; }
  EGatherPlan p;
  if (egather_plan(l.nel, nl, pos, slot0, off, wd, max_lds, p, err)) return -1;
  g = EGather();
  g.lg = p.lg; g.nl = nl; g.maxseg = p.maxseg; g.maxctr = p.maxctr; g.maxit = p.maxit;
  g.nblk = p.nblk; g.nseg = p.nseg; g.nctr = p.nctr;
  // bbase holds nblk + 1 offsets: a block reads its own and the next one (the end of its stream)
  if (eg_upload(st, &g.lseg, p.lseg) || eg_upload(st, &g.bbase, p.bbase) || eg_upload(st, &g.seg, p.seg) ||
      eg_upload(st, &g.cnt, p.cnt) || eg_upload(st, &g.cst, p.cst) || eg_upload(st, &g.cidx, p.cidx) ||
      hipStreamSynchronize(st) != hipSuccess) {
    fm_free_egather(g);
    err = "element-buffer gather: upload failed";
    return -1;
  }
  g.bytes = p.lseg.size() * 8 + p.bbase.size() * 8 + p.seg.size() * 4 + p.cnt.size() + p.cst.size() * 2 + p.cidx.size() * 2;
  g.built = true;
  return 0;
}

// Whether level L assembles through the element buffer: builds the gather at the first call, sizes the
// system's element buffers; false (the coloured scatter) when switched off, over budget or not buildable.
static int egather_ready(fm_system *s, int L, bool &use) {
  use = false;
  Level &l = s->lev[L];
  if (!g_asm_gather || l.eg.unusable || l.nel <= 0) return 0;
  const int nl = s->nloc;
  const size_t ne = (size_t)l.nel * nl * nl, nf = (size_t)l.nel * nl;
  if (!l.eg.built) {
    size_t held = std::max(s->ebuf_n, ne) * 8 + std::max(s->febuf_n, nf) * 8;
    for (const Level &o : s->lev) held += o.eg.bytes;
    // index: per slot count + start, per contribution (at most one per element entry) an LDS position
    const size_t index = (size_t)l.A.sell.nslot * 3 + ne * 2 + nf * 12;
    if (held + index > kGatherBudget) { l.eg.unusable = true; return 0; }
    std::string err;
    // staging per block (FEMUS_EGATHER_LDS_KB): 24 KB for scalar Q2 rows (six or more blocks resident per CU;
    // 56 KB took the 64^3 heat gather from 3.49 to 3.79 ms), 56 KB for Q2-Q1 rows, whose element rows fill 24 KB
    // with two lanes (96x24x24 channel: 9.41 -> 8.60 ms, four or eight lanes per block)
    const char *lk = getenv("FEMUS_EGATHER_LDS_KB");
    const size_t lds = (size_t)(lk && atoi(lk) >= 8 && atoi(lk) <= 64 ? atoi(lk) : (nl > 32 ? 56 : 24)) * 1024;
    if (fm_build_egather(s->ctx->stream, l, nl, lds, l.eg, err)) {
      (void)hipGetLastError();
      l.eg.unusable = true;
      return 0;
    }
  }
  if (s->ebuf_n < ne) {
    dfree(s->ebuf);
    s->ebuf_n = 0;
    FM_CHECK(dalloc(&s->ebuf, ne));
    s->ebuf_n = ne;
  }
  if (s->febuf_n < nf) {
    dfree(s->febuf);
    s->febuf_n = 0;
    FM_CHECK(dalloc(&s->febuf, nf));
    s->febuf_n = nf;
  }
  use = true;
  return 0;
}

// sum the element buffers of level L into the values (FP64 + FP32) of A (or of A^T: same layout) and the rhs
static int run_egather(fm_system *s, int L, DevSELL &target, double *rhs) {
  Level &l = s->lev[L];
  const EGather &g = l.eg;
  const DevSELL &S = l.A.sell;
  size_t oc, os, on;
  const size_t lds = egather_lds(g.maxseg, g.nl, g.maxctr, g.maxit, &oc, &os, &on);
  return fmk::egather(s->ctx->stream, g.nblk, g.lg, g.nl, (int)oc, (int)os, (int)on, lds, S.sl_off, S.sl_w, S.perm,
                      g.lseg, g.seg, g.bbase, g.cnt, g.cst, g.cidx, s->ebuf, s->febuf, target.v, target.vf, rhs);
}

// refresh FP32 values and diagonal inverses after an assembly (the gather writes the FP32 values itself)
static int after_assembly(fm_system *s, int L, bool gathered = false) {
  Level &l = s->lev[L];
  hipStream_t st = s->ctx->stream;
  if (!gathered) FM_CHECK(fmk::cast_vals(st, l.A.sell.nslot, l.A.sell.v, l.A.sell.vf));
  FM_CHECK(fmk::dinv_slot(st, l.n, l.dslot, l.A.sell.v, l.dinv, l.dinvf));
  l.lam = 0.;
  s->mg_ready = false;
  return 0;
}

extern "C" int fm_heat_genmatrhs(fm_system *s, int L, int mode, const fm_heat_params *p) {
  FM_CHECK(check_level(s, L));
  if (!p || s->nq2 != 1 || s->nq1 != 0 || p->dt == 0. || p->alpha0 <= 0.) return fm_einval(__func__);
  FM_HIP(hipSetDevice(s->ctx->device));
  Level &l = s->lev[L];
  hipStream_t st = s->ctx->stream;
  const fm_mesh &m = *s->mesh;
  HeatArgs a;
  a.alpha0 = p->alpha0; a.rhocp = p->rhocp; a.dt = p->dt; a.tc = p->time_coeff; a.source = p->source; a.upwind = p->upwind;
  for (int d = 0; d < 3; d++) a.vel_const[d] = p->vel_const[d];
  bool gather = false;
  FM_CHECK(egather_ready(s, L, gather));
  if (gather) {
    FM_CHECK(fmk::heat_assemble_e(st, m.dim, l.nel, l.conn, l.bmask, s->xyz, m.Ntop, s->d_dof_top, s->d_bc0, s->d_bc1,
                                  s->x_old, s->xbc, l.coef, s->vel, s->d_tab, a, mode, 0, s->ebuf, s->febuf));
    FM_CHECK(run_egather(s, L, l.A.sell, mode == 1 ? l.b : nullptr));
  } else {
    FM_HIP(hipMemsetAsync(l.A.sell.v, 0, l.A.sell.nslot * sizeof(double), st));
    if (mode == 1) FM_HIP(hipMemsetAsync(l.b, 0, l.n * sizeof(double), st));
    for (int k = 0; k < l.ncolor; k++) {
      int n = l.color_off[k + 1] - l.color_off[k];
      FM_CHECK(fmk::heat_assemble(st, m.dim, n, l.color_elems + l.color_off[k], l.conn, l.eldofi, l.elpos, l.elslot, l.bmask,
                                  s->xyz, m.Ntop, s->d_dof_top, s->d_bc0, s->d_bc1, s->x_old, s->xbc, l.coef, s->vel,
                                  s->d_tab, a, mode, 0, 0, l.A.sell.v, l.b));
    }
  }
  FM_CHECK(after_assembly(s, L, gather));
  FM_HIP(hipStreamSynchronize(st));
  return FM_OK;
}

// A caller's own values of A[L] (SparseMatrixM::zero and set(i, j, v) over the whole pattern,
// contrib/matrix/sparse_matrixM.h:113-116, before LinearSolverM::solve, contrib/matrix/linear_solverM.h:150-154):
// val holds the nnz values in the CSR order of the level's pattern as fm_system_csr_get (which 0) reports it
// (reference numbering, columns sorted).  They replace the FP64 and FP32 SELL values and the Jacobi diagonal;
// the multigrid set-up is rebuilt at the next solve.
static int set_matrix(fm_system *s, int L, const double *val, bool dev) {
  Level &l = s->lev[L];
  hipStream_t st = s->ctx->stream;
  FM_HIP(hipSetDevice(s->ctx->device));
  if (l.A.nr == 0 || l.A.sell.nsl == 0) { fm_set_error("set_matrix: level has no matrix"); return FM_ESTATE; }
  const double *src = val;
  if (!dev) {  // staged in the CSR value array, which holds CSR values only on request anyway
    if (!l.A.v) {
      FM_CHECK(dalloc(&l.A.v, l.A.nnz));
      FM_CHECK(dalloc(&l.A.vf, l.A.nnz));
    }
    FM_HIP(hipMemcpyAsync(l.A.v, val, l.A.nnz * sizeof(double), hipMemcpyHostToDevice, st));
    src = l.A.v;
  }
  FM_CHECK(fmk::csr2sell(st, sell_map(l.A.sell), src, l.A.sell.v, l.A.sell.vf));
  FM_CHECK(after_assembly(s, L, true));
  FM_HIP(hipStreamSynchronize(st));
  return FM_OK;
}
extern "C" int fm_system_set_matrix(fm_system *s, int L, const double *val) {
  FM_CHECK(check_level(s, L));
  if (!val) return fm_einval(__func__);
  return set_matrix(s, L, val, false);
}
extern "C" int fm_system_set_matrix_dev(fm_system *s, int L, const double *val_dev) {
  FM_CHECK(check_level(s, L));
  FM_CHECK(check_dev_ptr(__func__, val_dev));
  return set_matrix(s, L, val_dev, true);
}

extern "C" int fm_ns_genmatrhs(fm_system *s, int L, int mode, const fm_ns_params *p) {
  FM_CHECK(check_level(s, L));
  const fm_mesh &m = *s->mesh;
  if (!p || s->nq2 != m.dim || s->nq1 != 1 || p->dt == 0. || p->IRe <= 0.) return fm_einval(__func__);
  FM_HIP(hipSetDevice(s->ctx->device));
  Level &l = s->lev[L];
  hipStream_t st = s->ctx->stream;
  NsArgs a;
  a.IRe = p->IRe; a.dt = p->dt; a.tc = p->time_coeff; a.upwind = p->upwind; a.IFr = p->IFr;
  for (int d = 0; d < 3; d++) { a.dirg[d] = p->dirg[d]; a.src[d] = p->src[d]; a.buoy[d] = s->buoy[d]; }
  a.Tn = s->tnode;
  bool gather = false;
  FM_CHECK(egather_ready(s, L, gather));
  if (gather) {
    FM_CHECK(fmk::ns_assemble_e(st, m.dim, l.nel, l.conn, l.bmask, s->xyz, m.Ntop, s->d_dof_top, s->d_bc0, s->d_bc1,
                                s->x_old, s->xbc, l.coef, s->d_tab, a, mode, 0, s->ebuf, s->febuf));
    FM_CHECK(run_egather(s, L, l.A.sell, mode == 1 ? l.b : nullptr));
  } else {
    FM_HIP(hipMemsetAsync(l.A.sell.v, 0, l.A.sell.nslot * sizeof(double), st));
    if (mode == 1) FM_HIP(hipMemsetAsync(l.b, 0, l.n * sizeof(double), st));
    for (int k = 0; k < l.ncolor; k++) {
      int n = l.color_off[k + 1] - l.color_off[k];
      FM_CHECK(fmk::ns_assemble(st, m.dim, n, l.color_elems + l.color_off[k], l.conn, l.eldofi, l.elpos, l.elslot, l.bmask,
                                s->xyz, m.Ntop, s->d_dof_top, s->d_bc0, s->d_bc1, s->x_old, s->xbc, l.coef, s->d_tab, a,
                                mode, 0, 0, l.A.sell.v, l.b));
    }
  }
  FM_CHECK(after_assembly(s, L, gather));
  FM_HIP(hipStreamSynchronize(st));
  return FM_OK;
}

// ===========================================================================
// level operators
// ===========================================================================
extern "C" int fm_spmv(fm_system *s, int L, const double *xh, double *yh) {
  FM_CHECK(check_level(s, L));
  if (!xh || !yh) return fm_einval(__func__);
  Level &l = s->lev[L];
  hipStream_t st = s->ctx->stream;
  FM_HIP(hipMemcpyAsync(l.t, xh, l.n * sizeof(double), hipMemcpyHostToDevice, st));
  FM_CHECK(fm_csr_from_sell(st, l.A));
  CsrView<double> A = view<double>(l.A);
  A.nsl = 0;  // CSR copy: reference numbering on both sides
  FM_CHECK(fmk::spmv(st, A, l.t, l.d));
  FM_HIP(hipMemcpyAsync(yh, l.d, l.n * sizeof(double), hipMemcpyDeviceToHost, st));
  FM_HIP(hipStreamSynchronize(st));
  return FM_OK;
}
// y = A^T x with the transposed matrix of the adjoint assembly, vectors in the reference DoF order
extern "C" int fm_adjoint_spmv(fm_system *s, int L, const double *xh, double *yh) {
  FM_CHECK(check_level(s, L));
  if (!xh || !yh) return fm_einval(__func__);
  if (L >= (int)s->alev.size() || L >= (int)s->a_assembled.size() || !s->a_assembled[L]) {
    fm_set_error("fm_adjoint_spmv: level not assembled transposed");
    return FM_ESTATE;
  }
  Level &l = s->lev[L];
  hipStream_t st = s->ctx->stream;
  FM_HIP(hipSetDevice(s->ctx->device));
  FM_HIP(hipMemcpyAsync(l.t, xh, l.n * sizeof(double), hipMemcpyHostToDevice, st));
  // the SELL copy of A^T in the level's internal order: x gathered in (d[i] = x[ord[i]]), the product, the
  // result gathered back to the reference numbering (y[j] = d2[rank[j]]); no CSR values are formed
  FM_CHECK(fmk::gather(st, l.n, l.ord, l.t, l.d));
  FM_CHECK(fmk::spmv(st, view<double>(s->alev[L].A), l.d, l.d2));
  FM_CHECK(fmk::gather(st, l.n, l.rank, l.d2, l.t));
  FM_HIP(hipMemcpyAsync(yh, l.t, l.n * sizeof(double), hipMemcpyDeviceToHost, st));
  FM_HIP(hipStreamSynchronize(st));
  return FM_OK;
}
extern "C" int fm_residual(fm_system *s, int L) {
  FM_CHECK(check_level(s, L));
  Level &l = s->lev[L];
  FM_CHECK(fmk::resid(s->ctx->stream, view<double>(l.A), l.b, l.x, l.res));
  FM_HIP(hipStreamSynchronize(s->ctx->stream));
  return FM_OK;
}
extern "C" int fm_restrict(fm_system *s, int L) {
  FM_CHECK(check_level(s, L));
  if (L < 1) return fm_einval(__func__);
  Level &lc = s->lev[L - 1], &lf = s->lev[L];
  FM_CHECK(fmk::spmv(s->ctx->stream, view<double>(lc.R), lf.res, lc.b));
  FM_HIP(hipStreamSynchronize(s->ctx->stream));
  return FM_OK;
}
extern "C" int fm_prolong(fm_system *s, int L) {
  FM_CHECK(check_level(s, L));
  if (L < 1) return fm_einval(__func__);
  Level &lc = s->lev[L - 1], &lf = s->lev[L];
  FM_CHECK(fmk::spmv(s->ctx->stream, view<double>(lf.P), lc.x, lf.res));
  FM_HIP(hipStreamSynchronize(s->ctx->stream));
  return FM_OK;
}

extern "C" int fm_bench_spmv(fm_system *s, int L, int prec, int reps, double *sec) {
  FM_CHECK(check_level(s, L));
  if (reps < 1 || !sec || prec < 0 || prec > 5) return fm_einval(__func__);
  Level &l = s->lev[L];
  hipStream_t st = s->ctx->stream;
  // prec 0 / 1: SELL-64 FP64 / FP32 y = A x (the solver's format); 2 / 3: CSR FP64 / FP32 y = A x;
  // 4 / 5: the SELL Chebyshev smoothing step FP64 / FP32 (r -= A d, d' = c1 d + c2 D^-1 r, x += d'), the
  // product the V-cycle spends its time in.  The level's own work vectors are overwritten.
  CsrView<double> v64 = view<double>(l.A);
  CsrView<float> v32 = view<float>(l.A);
  if (prec == 2 || prec == 3) {
    FM_CHECK(fm_csr_from_sell(st, l.A));
    v64.nsl = 0; v32.nsl = 0;
  }
  const bool f32 = (prec & 1) != 0, cheb = prec >= 4;
  FM_CHECK(fmk::fill<double>(st, l.n, l.t, 1.0));
  FM_CHECK(fmk::fill<float>(st, l.n, l.tf, 1.0f));
  if (cheb) {
    // a damped momentum-Jacobi recurrence (c1 = 0.3, c2 = 0.25 with the Jacobi diagonal): bounded iterates
    FM_CHECK(fmk::dinv_slot(st, l.n, l.dslot, l.A.sell.v, l.dinv, l.dinvf));
    FM_CHECK(fmk::fill<double>(st, l.n, l.r, 1.0)); FM_CHECK(fmk::fill<double>(st, l.n, l.x, 0.0));
    FM_CHECK(fmk::fill<double>(st, l.n, l.d, 0.0)); FM_CHECK(fmk::fill<double>(st, l.n, l.d2, 0.0));
    FM_CHECK(fmk::fill<float>(st, l.n, l.rf, 1.0f)); FM_CHECK(fmk::fill<float>(st, l.n, l.xf, 0.0f));
    FM_CHECK(fmk::fill<float>(st, l.n, l.df, 0.0f)); FM_CHECK(fmk::fill<float>(st, l.n, l.d2f, 0.0f));
  }
  int k2 = 0;
  auto one = [&]() -> int {
    if (cheb) {
      const int k = k2++;
      if (!f32) return fmk::cheb_step(st, v64, l.r, l.x, k & 1 ? l.d2 : l.d, k & 1 ? l.d : l.d2, l.dinv, 0.3, 0.25);
      return fmk::cheb_step(st, v32, l.rf, l.xf, k & 1 ? l.d2f : l.df, k & 1 ? l.df : l.d2f, l.dinvf, 0.3f, 0.25f);
    }
    if (!f32) return fmk::spmv(st, v64, l.t, l.d);
    return fmk::spmv(st, v32, l.tf, l.df);
  };
  FM_CHECK(one());
  FM_HIP(hipEventRecord(s->ctx->ev0, st));
  for (int k = 0; k < reps; k++) FM_CHECK(one());
  FM_HIP(hipEventRecord(s->ctx->ev1, st));
  FM_HIP(hipEventSynchronize(s->ctx->ev1));
  float ms = 0.f;
  FM_HIP(hipEventElapsedTime(&ms, s->ctx->ev0, s->ctx->ev1));
  *sec = ms * 1e-3 / reps;
  return FM_OK;
}

// ===========================================================================
// discrete adjoint: transposed hierarchy, transposed assembly, functionals
// ===========================================================================
static void transpose_host(int64_t nr, int64_t nc, const std::vector<int64_t> &rp, const std::vector<int> &ci,
                           const std::vector<double> &v, std::vector<int64_t> &rpT, std::vector<int> &ciT,
                           std::vector<double> &vT) {
  rpT.assign(nc + 1, 0);
  for (int64_t q = 0; q < rp[nr]; q++) rpT[ci[q] + 1]++;
  for (int64_t j = 0; j < nc; j++) rpT[j + 1] += rpT[j];
  ciT.resize(rp[nr]);
  vT.resize(rp[nr]);
  std::vector<int64_t> f(rpT.begin(), rpT.end() - 1);
  for (int64_t i = 0; i < nr; i++)
    for (int64_t q = rp[i]; q < rp[i + 1]; q++) { int64_t t = f[ci[q]]++; ciT[t] = (int)i; vT[t] = v[q]; }
}

// A matrix with the pattern of fwd (the transposed operator of a structurally symmetric pattern): its own row
// pointers and SELL values, fwd's SELL layout (column offsets, slot maps, row order) borrowed, the CSR columns
// kept on the host until a caller asks for the CSR form.
static int share_pattern(hipStream_t st, DevCSR &c, const DevCSR &fwd) {
  free_csr(c);
  c.nr = fwd.nr; c.nc = fwd.nc; c.nnz = fwd.nnz;
  FM_CHECK(dalloc(&c.rp, c.nr + 1));
  FM_HIP(hipMemcpyAsync(c.rp, fwd.rp, (c.nr + 1) * sizeof(int64_t), hipMemcpyDeviceToDevice, st));
  c.h_rp = fwd.h_rp;
  c.h_ci = fwd.h_ci;
  c.sell = fwd.sell;
  c.sell.borrowed = true;
  c.sell.v = nullptr;
  c.sell.vf = nullptr;
  FM_CHECK(dalloc(&c.sell.v, c.sell.nslot));
  FM_CHECK(dalloc(&c.sell.vf, c.sell.nslot));
  return fm_sync_sell(st, c);  // no CSR values: zeros
}

static int ensure_adjoint(fm_system *s) {
  if (!s->alev.empty()) return 0;
  hipStream_t st = s->ctx->stream;
  int nlev = (int)s->lev.size(), top = nlev - 1;
  s->alev.resize(nlev);
  s->a_assembled.assign(nlev, 0);
  for (int L = 0; L < nlev; L++) {
    Level &a = s->alev[L], &l = s->lev[L];
    a.n = l.n;
    a.nel = l.nel;
    a.h_ord = l.h_ord;
    a.h_rank = l.h_rank;
    FM_CHECK(dalloc(&a.ord, a.n)); FM_CHECK(dalloc(&a.rank, a.n));
    FM_HIP(hipMemcpyAsync(a.ord, l.h_ord.data(), a.n * sizeof(int), hipMemcpyHostToDevice, st));
    FM_HIP(hipMemcpyAsync(a.rank, l.h_rank.data(), a.n * sizeof(int), hipMemcpyHostToDevice, st));
    a.h_goff = l.h_goff;
    a.h_gmask = l.h_gmask;
    FM_CHECK(share_pattern(st, a.A, l.A));  // A^T has A's pattern: same SELL layout, own values
    FM_CHECK(colour_slices(st, a));
    FM_CHECK(dalloc(&a.dslot, a.n));
    FM_HIP(hipMemcpyAsync(a.dslot, l.dslot, a.n * sizeof(int64_t), hipMemcpyDeviceToDevice, st));
    if (L < top) {  // restriction of the adjoint = (prolongation of level L+1)^T
      const DevCSR &P = s->lev[L + 1].P;
      std::vector<int64_t> rpT; std::vector<int> ciT; std::vector<double> vT;
      transpose_host(P.nr, P.nc, P.h_rp, P.h_ci, P.h_v, rpT, ciT, vT);
      FM_CHECK(upload_csr(st, a.R, P.nc, P.nr, rpT, ciT, &vT, l.h_ord, s->lev[L + 1].h_rank));
    }
    if (L > 0) {  // prolongation of the adjoint = (restriction of level L-1)^T
      const DevCSR &R = s->lev[L - 1].R;
      std::vector<int64_t> rpT; std::vector<int> ciT; std::vector<double> vT;
      transpose_host(R.nr, R.nc, R.h_rp, R.h_ci, R.h_v, rpT, ciT, vT);
      FM_CHECK(upload_csr(st, a.P, R.nc, R.nr, rpT, ciT, &vT, l.h_ord, s->lev[L - 1].h_rank));
    }
    FM_CHECK(fm_alloc_level_vectors(a, L == top));
    FM_HIP(hipMemsetAsync(a.x, 0, a.n * sizeof(double), st));
  }
  FM_HIP(hipStreamSynchronize(st));
  return 0;
}

static int after_assembly_T(fm_system *s, int L, bool gathered = false) {
  Level &a = s->alev[L];
  hipStream_t st = s->ctx->stream;
  if (!gathered) FM_CHECK(fmk::cast_vals(st, a.A.sell.nslot, a.A.sell.v, a.A.sell.vf));
  FM_CHECK(fmk::dinv_slot(st, a.n, a.dslot, a.A.sell.v, a.dinv, a.dinvf));
  a.lam = 0.;
  s->a_assembled[L] = 1;
  s->adj_ready = false;
  FM_HIP(hipStreamSynchronize(st));
  return 0;
}

extern "C" int fm_heat_genmat_transpose(fm_system *s, int L, const fm_heat_params *p) {
  FM_CHECK(check_level(s, L));
  if (!p || s->nq2 != 1 || s->nq1 != 0 || p->dt == 0. || p->alpha0 <= 0.) return fm_einval(__func__);
  FM_HIP(hipSetDevice(s->ctx->device));
  FM_CHECK(ensure_adjoint(s));
  Level &l = s->lev[L], &a = s->alev[L];
  hipStream_t st = s->ctx->stream;
  const fm_mesh &m = *s->mesh;
  HeatArgs h;
  h.alpha0 = p->alpha0; h.rhocp = p->rhocp; h.dt = p->dt; h.tc = p->time_coeff; h.source = p->source; h.upwind = p->upwind;
  for (int d = 0; d < 3; d++) h.vel_const[d] = p->vel_const[d];
  bool gather = false;
  FM_CHECK(egather_ready(s, L, gather));
  if (gather) {
    FM_CHECK(fmk::heat_assemble_e(st, m.dim, l.nel, l.conn, l.bmask, s->xyz, m.Ntop, s->d_dof_top, s->d_bc0, s->d_bc1,
                                  s->x_old, s->xbc, l.coef, s->vel, s->d_tab, h, 0, 1, s->ebuf, s->febuf));
    FM_CHECK(run_egather(s, L, a.A.sell, nullptr));
  } else {
    FM_HIP(hipMemsetAsync(a.A.sell.v, 0, a.A.sell.nslot * sizeof(double), st));
    for (int k = 0; k < l.ncolor; k++) {
      int n = l.color_off[k + 1] - l.color_off[k];
      FM_CHECK(fmk::heat_assemble(st, m.dim, n, l.color_elems + l.color_off[k], l.conn, l.eldof, l.elpos, l.elslot, l.bmask,
                                  s->xyz, m.Ntop, s->d_dof_top, s->d_bc0, s->d_bc1, s->x_old, s->xbc, l.coef, s->vel,
                                  s->d_tab, h, 0, 1, 0, a.A.sell.v, nullptr));
    }
  }
  return after_assembly_T(s, L, gather);
}

extern "C" int fm_ns_genmat_transpose(fm_system *s, int L, const fm_ns_params *p) {
  FM_CHECK(check_level(s, L));
  const fm_mesh &m = *s->mesh;
  if (!p || s->nq2 != m.dim || s->nq1 != 1 || p->dt == 0. || p->IRe <= 0.) return fm_einval(__func__);
  FM_HIP(hipSetDevice(s->ctx->device));
  FM_CHECK(ensure_adjoint(s));
  Level &l = s->lev[L], &a = s->alev[L];
  hipStream_t st = s->ctx->stream;
  NsArgs h;
  h.IRe = p->IRe; h.dt = p->dt; h.tc = p->time_coeff; h.upwind = p->upwind; h.IFr = p->IFr;
  for (int d = 0; d < 3; d++) { h.dirg[d] = p->dirg[d]; h.src[d] = p->src[d]; h.buoy[d] = 0.; }
  h.Tn = nullptr;
  bool gather = false;
  FM_CHECK(egather_ready(s, L, gather));
  if (gather) {
    FM_CHECK(fmk::ns_assemble_e(st, m.dim, l.nel, l.conn, l.bmask, s->xyz, m.Ntop, s->d_dof_top, s->d_bc0, s->d_bc1,
                                s->x_old, s->xbc, l.coef, s->d_tab, h, 0, 1, s->ebuf, s->febuf));
    FM_CHECK(run_egather(s, L, a.A.sell, nullptr));
  } else {
    FM_HIP(hipMemsetAsync(a.A.sell.v, 0, a.A.sell.nslot * sizeof(double), st));
    for (int k = 0; k < l.ncolor; k++) {
      int n = l.color_off[k + 1] - l.color_off[k];
      FM_CHECK(fmk::ns_assemble(st, m.dim, n, l.color_elems + l.color_off[k], l.conn, l.eldof, l.elpos, l.elslot, l.bmask,
                                s->xyz, m.Ntop, s->d_dof_top, s->d_bc0, s->d_bc1, s->x_old, s->xbc, l.coef, s->d_tab, h,
                                0, 1, 0, a.A.sell.v, nullptr));
    }
  }
  return after_assembly_T(s, L, gather);
}

// Newton Jacobian of the steady residual R(u) = A(u) u - b at u = x_old (the Oseen matrix of GenMatRhs plus the
// derivatives of the advection term and of the upwind coefficients with respect to the advecting field); trans = 1
// scatters its transpose into the adjoint matrices.
extern "C" int fm_ns_genjac(fm_system *s, int L, int trans, const fm_ns_params *p) {
  FM_CHECK(check_level(s, L));
  const fm_mesh &m = *s->mesh;
  if (!p || s->nq2 != m.dim || s->nq1 != 1 || p->dt == 0. || p->IRe <= 0. || (trans != 0 && trans != 1)) return fm_einval(__func__);
  FM_HIP(hipSetDevice(s->ctx->device));
  if (trans) FM_CHECK(ensure_adjoint(s));
  Level &l = s->lev[L];
  double *val = trans ? s->alev[L].A.sell.v : l.A.sell.v;
  const int64_t nslot = trans ? s->alev[L].A.sell.nslot : l.A.sell.nslot;
  hipStream_t st = s->ctx->stream;
  NsArgs a;
  a.IRe = p->IRe; a.dt = p->dt; a.tc = p->time_coeff; a.upwind = p->upwind; a.IFr = p->IFr;
  for (int d = 0; d < 3; d++) { a.dirg[d] = p->dirg[d]; a.src[d] = p->src[d]; a.buoy[d] = 0.; }
  a.Tn = nullptr;
  a.newton = 1;
  bool gather = false;
  FM_CHECK(egather_ready(s, L, gather));
  if (gather) {
    FM_CHECK(fmk::ns_assemble_e(st, m.dim, l.nel, l.conn, l.bmask, s->xyz, m.Ntop, s->d_dof_top, s->d_bc0, s->d_bc1,
                                s->x_old, s->xbc, l.coef, s->d_tab, a, 0, trans, s->ebuf, s->febuf));
    FM_CHECK(run_egather(s, L, trans ? s->alev[L].A.sell : l.A.sell, nullptr));
  } else {
    FM_HIP(hipMemsetAsync(val, 0, nslot * sizeof(double), st));
    for (int k = 0; k < l.ncolor; k++) {
      int n = l.color_off[k + 1] - l.color_off[k];
      FM_CHECK(fmk::ns_assemble(st, m.dim, n, l.color_elems + l.color_off[k], l.conn, l.eldofi, l.elpos, l.elslot, l.bmask,
                                s->xyz, m.Ntop, s->d_dof_top, s->d_bc0, s->d_bc1, s->x_old, s->xbc, l.coef, s->d_tab, a, 0,
                                trans, 0, val, nullptr));
    }
  }
  if (trans) return after_assembly_T(s, L, gather);
  FM_CHECK(after_assembly(s, L, gather));
  FM_HIP(hipStreamSynchronize(st));
  return FM_OK;
}

extern "C" int fm_get_adjoint_csr(const fm_system *s, int L, double *val) {
  if (!s || L < 0 || L >= (int)s->alev.size() || !val) return fm_einval(__func__);
  DevCSR &c = const_cast<DevCSR &>(s->alev[L].A);
  FM_CHECK(fm_csr_from_sell(s->ctx->stream, c));
  FM_HIP(hipStreamSynchronize(s->ctx->stream));
  FM_HIP(hipMemcpy(val, c.v, c.nnz * sizeof(double), hipMemcpyDeviceToHost));
  return FM_OK;
}

// J = alpha0 * sum_e k_e T_e^T Klap_e T_e;  dJ/dT = 2 alpha0 sum_e k_e Klap_e T_e;  dJ/dk_e = alpha0 T_e^T Klap_e T_e
extern "C" int fm_heat_dissipation(fm_system *s, const fm_heat_params *p, const double *T_host, double *J,
                                   double *dJdT_host, double *dJdk_host) {
  FM_CHECK(check_level(s, 0));
  if (!p || !T_host || s->nq2 != 1 || s->nq1 != 0) return fm_einval(__func__);
  int top = (int)s->lev.size() - 1;
  Level &l = s->lev[top];
  hipStream_t st = s->ctx->stream;
  const fm_mesh &m = *s->mesh;
  FM_HIP(hipMemcpyAsync(l.t, T_host, l.n * sizeof(double), hipMemcpyHostToDevice, st));
  FM_CHECK(fmk::fill<double>(st, l.n, l.d, 0.0));
  double *diss = nullptr;
  FM_CHECK(dalloc(&diss, l.nel));
  int rc = 0;
  for (int k = 0; k < l.ncolor && !rc; k++) {
    int n = l.color_off[k + 1] - l.color_off[k];
    rc = fmk::heat_lap(st, m.dim, n, l.color_elems + l.color_off[k], l.conn, l.eldof, s->xyz, m.Ntop, s->d_dof_top,
                       s->d_bc0, l.t, nullptr, l.coef, s->d_tab, diss, nullptr, l.d, 2.0 * p->alpha0);
  }
  std::vector<double> h(l.nel), hk(l.nel, 1.0);
  if (!rc) {
    FM_HIP(hipMemcpyAsync(h.data(), diss, l.nel * sizeof(double), hipMemcpyDeviceToHost, st));
    if (l.coef) FM_HIP(hipMemcpyAsync(hk.data(), l.coef, l.nel * sizeof(double), hipMemcpyDeviceToHost, st));
    if (dJdT_host) FM_HIP(hipMemcpyAsync(dJdT_host, l.d, l.n * sizeof(double), hipMemcpyDeviceToHost, st));
    FM_HIP(hipStreamSynchronize(st));
  }
  dfree(diss);
  if (rc) return FM_EHIP;
  double acc = 0.;
  for (int e = 0; e < l.nel; e++) {
    acc += hk[e] * p->alpha0 * h[e];
    if (dJdk_host) dJdk_host[e] = p->alpha0 * h[e];
  }
  if (J) *J = acc;
  return FM_OK;
}

// adjoint part of dJ/dk_e: -lambda^T (dA/dk_e) T = -alpha0 sum_i bc_vol_i lambda_i (Klap_e T_e)_i
extern "C" int fm_heat_grad_coef(fm_system *s, const fm_heat_params *p, const double *lam_host, const double *T_host,
                                 double *grad_host) {
  FM_CHECK(check_level(s, 0));
  if (!p || !lam_host || !T_host || !grad_host || s->nq2 != 1 || s->nq1 != 0) return fm_einval(__func__);
  int top = (int)s->lev.size() - 1;
  Level &l = s->lev[top];
  hipStream_t st = s->ctx->stream;
  const fm_mesh &m = *s->mesh;
  FM_HIP(hipMemcpyAsync(l.t, T_host, l.n * sizeof(double), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(l.d2, lam_host, l.n * sizeof(double), hipMemcpyHostToDevice, st));
  double *adj = nullptr;
  FM_CHECK(dalloc(&adj, l.nel));
  int rc = fmk::heat_lap(st, m.dim, l.nel, l.color_elems, l.conn, l.eldof, s->xyz, m.Ntop, s->d_dof_top, s->d_bc0,
                         l.t, l.d2, l.coef, s->d_tab, nullptr, adj, nullptr, 0.0);
  std::vector<double> h(l.nel);
  if (!rc) {
    FM_HIP(hipMemcpyAsync(h.data(), adj, l.nel * sizeof(double), hipMemcpyDeviceToHost, st));
    FM_HIP(hipStreamSynchronize(st));
  }
  dfree(adj);
  if (rc) return FM_EHIP;
  for (int e = 0; e < l.nel; e++) grad_host[e] = -p->alpha0 * h[e];
  return FM_OK;
}

// STREAM copy of `bytes` (rounded down to 16 B) between two device buffers;
// *sec = average kernel time over reps (HIP events on the context stream).
extern "C" int fm_bench_stream(fm_ctx *c, int64_t bytes, int reps, double *sec) {
  if (!c || bytes < 16 || reps < 1 || !sec) return fm_einval(__func__);
  FM_HIP(hipSetDevice(c->device));
  int64_t n16 = bytes / 16;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n16 * 16) != hipSuccess) return FM_ENOMEM;
  if (hipMalloc(&b, n16 * 16) != hipSuccess) { (void)hipFree(a); return FM_ENOMEM; }
  int rc = 0;
  if (hipMemsetAsync(a, 0, n16 * 16, c->stream) != hipSuccess) rc = FM_EHIP;
  if (!rc) rc = fmk::copy16(c->stream, n16, a, b);
  if (!rc && hipEventRecord(c->ev0, c->stream) != hipSuccess) rc = FM_EHIP;
  for (int k = 0; k < reps && !rc; k++) rc = fmk::copy16(c->stream, n16, a, b);
  if (!rc && hipEventRecord(c->ev1, c->stream) != hipSuccess) rc = FM_EHIP;
  if (!rc && hipEventSynchronize(c->ev1) != hipSuccess) rc = FM_EHIP;
  float ms = 0.f;
  if (!rc && hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess) rc = FM_EHIP;
  (void)hipFree(a);
  (void)hipFree(b);
  if (rc) return rc < 0 ? rc : FM_EHIP;
  *sec = ms * 1e-3 / reps;
  return FM_OK;
}

// NS viscous dissipation J = sum_e nu_e int_e |grad u|^2 (top level), dJ/dU (DoF order; pressure rows 0),
// partials dJ/dnu_e.
extern "C" int fm_ns_dissipation(fm_system *s, const fm_ns_params *p, const double *U_host, double *J,
                                 double *dJdU_host, double *dJdnu_host) {
  FM_CHECK(check_level(s, 0));
  if (!p || !U_host || s->nq2 != s->mesh->dim || s->nq1 != 1) return fm_einval(__func__);
  int top = (int)s->lev.size() - 1;
  Level &l = s->lev[top];
  hipStream_t st = s->ctx->stream;
  const fm_mesh &m = *s->mesh;
  FM_HIP(hipMemcpyAsync(l.t, U_host, l.n * sizeof(double), hipMemcpyHostToDevice, st));
  FM_CHECK(fmk::fill<double>(st, l.n, l.d, 0.0));
  double *diss = nullptr;
  FM_CHECK(dalloc(&diss, l.nel));
  int rc = 0;
  for (int k = 0; k < l.ncolor && !rc; k++) {
    int n = l.color_off[k + 1] - l.color_off[k];
    rc = fmk::ns_visc(st, m.dim, n, l.color_elems + l.color_off[k], l.conn, l.eldof, s->xyz, m.Ntop, s->d_dof_top,
                      s->d_bc0, l.t, nullptr, l.coef, p->IRe, s->d_tab, diss, nullptr, l.d, 2.0);
  }
  std::vector<double> h(l.nel), hk(l.nel, p->IRe);
  if (!rc) {
    FM_HIP(hipMemcpyAsync(h.data(), diss, l.nel * sizeof(double), hipMemcpyDeviceToHost, st));
    if (l.coef) FM_HIP(hipMemcpyAsync(hk.data(), l.coef, l.nel * sizeof(double), hipMemcpyDeviceToHost, st));
    if (dJdU_host) FM_HIP(hipMemcpyAsync(dJdU_host, l.d, l.n * sizeof(double), hipMemcpyDeviceToHost, st));
    FM_HIP(hipStreamSynchronize(st));
  }
  dfree(diss);
  if (rc) return FM_EHIP;
  double acc = 0.;
  for (int e = 0; e < l.nel; e++) {
    acc += hk[e] * h[e];
    if (dJdnu_host) dJdnu_host[e] = h[e];
  }
  if (J) *J = acc;
  return FM_OK;
}

// adjoint term of dJ/dnu_e for the Oseen step: -lambda^T (dA/dnu_e) U
extern "C" int fm_ns_grad_coef(fm_system *s, const fm_ns_params *p, const double *lam_host, const double *U_host,
                               double *grad_host) {
  FM_CHECK(check_level(s, 0));
  if (!p || !lam_host || !U_host || !grad_host || s->nq2 != s->mesh->dim || s->nq1 != 1) return fm_einval(__func__);
  int top = (int)s->lev.size() - 1;
  Level &l = s->lev[top];
  hipStream_t st = s->ctx->stream;
  const fm_mesh &m = *s->mesh;
  FM_HIP(hipMemcpyAsync(l.t, U_host, l.n * sizeof(double), hipMemcpyHostToDevice, st));
  FM_HIP(hipMemcpyAsync(l.d2, lam_host, l.n * sizeof(double), hipMemcpyHostToDevice, st));
  double *adj = nullptr;
  FM_CHECK(dalloc(&adj, l.nel));
  int rc = fmk::ns_visc(st, m.dim, l.nel, l.color_elems, l.conn, l.eldof, s->xyz, m.Ntop, s->d_dof_top, s->d_bc0,
                        l.t, l.d2, l.coef, p->IRe, s->d_tab, nullptr, adj, nullptr, 0.0);
  std::vector<double> h(l.nel);
  if (!rc) {
    FM_HIP(hipMemcpyAsync(h.data(), adj, l.nel * sizeof(double), hipMemcpyDeviceToHost, st));
    FM_HIP(hipStreamSynchronize(st));
  }
  dfree(adj);
  if (rc) return FM_EHIP;
  for (int e = 0; e < l.nel; e++) grad_host[e] = -h[e];
  return FM_OK;
}

// Pressure drop between the faces x = lo[0] (inlet) and x = hi[0] (outlet) of
// the top level: J = mean_inlet(p) - mean_outlet(p), face means of the Q1
// pressure (bilinear on each face: weight area/4 per vertex in 3D, length/2
// per vertex in 2D on the box's axis-aligned faces).  J is linear in U;
// dJdU (reference numbering, zero on velocity DoFs) is returned with J.
extern "C" int fm_ns_pressure_drop(fm_system *s, const double *U, double *J, double *dJdU) {
  FM_CHECK(check_level(s, 0));
  const fm_mesh &m = *s->mesh;
  if (!U || s->nq2 != m.dim || s->nq1 != 1) return fm_einval(__func__);
  const int top = m.nlev - 1, N = m.Ntop, dim = m.dim, nvert = dim == 2 ? 2 : 4;
  const std::vector<int> &dof = s->node_dof[top];
  const int64_t nt = s->Dim[top];
  std::vector<double> w(nt, 0.);
  double area[2] = {0., 0.};
  const double tol = 1e-9 * (m.hi[0] - m.lo[0]);
  for (int e = 0; e < m.nel[top]; e++)
    for (int sd = 0; sd < m.nsides; sd++) {
      if (m.neigh[top][(size_t)e * m.nsides + sd] != -1) continue;
      int nodes[4];
      bool in = true, out = true;
      for (int i = 0; i < nvert; i++) {
        int ln = dim == 2 ? fm_face2d[sd][i] : fm_face3d[sd][i];
        nodes[i] = m.conn[top][(size_t)e * m.n2 + ln];
        double x = m.xyz[nodes[i]];
        in = in && std::fabs(x - m.lo[0]) < tol;
        out = out && std::fabs(x - m.hi[0]) < tol;
      }
      if (!in && !out) continue;
      double a = 1.;
      for (int d = 1; d < dim; d++) {
        double lo = 1e300, hi = -1e300;
        for (int i = 0; i < nvert; i++) { double x = m.xyz[nodes[i] + (size_t)d * N]; lo = std::min(lo, x); hi = std::max(hi, x); }
        a *= hi - lo;
      }
      const int side = in ? 0 : 1;
      area[side] += a;
      for (int i = 0; i < nvert; i++) {
        int pd = dof[nodes[i] + (size_t)dim * N];
        if (pd < 0) { fm_set_error("pressure drop: face vertex without a pressure DoF"); return FM_ESTATE; }
        w[pd] += (side == 0 ? 1. : -1.) * a / nvert;
      }
    }
  if (area[0] <= 0. || area[1] <= 0.) { fm_set_error("pressure drop: no inlet or outlet face"); return FM_EINVAL; }
  double acc = 0.;
  for (int64_t i = 0; i < nt; i++) {
    if (w[i] != 0.) w[i] /= (w[i] > 0. ? area[0] : area[1]);
    acc += w[i] * U[i];
  }
  if (J) *J = acc;
  if (dJdU) std::memcpy(dJdU, w.data(), nt * sizeof(double));
  return FM_OK;
}

// Nusselt number of a differentially heated box along `axis` (top level):
// J = sum_e int_e (u_a T / kappa - k_e dT/dx_a) with the nodal advection field
// of fm_system_set_velocity, kappa = alpha0; dJ/dT (reference numbering) and
// the partials dJ/dk_e = -int_e dT/dx_a.
// Coupling products of the Boussinesq system on the heat system's top level (node-space outputs; see k_coupling):
// mode 0: dNu/du_axis (needs T), mode 1: (dR_T/du)^T lam (T, lam; out has dim * Ntop entries), mode 2: M m (m nodal).
extern "C" int fm_heat_coupling(fm_system *s, const fm_heat_params *p, int mode, double scale, const double *T_host,
                                const double *lam_host, const double *m_host, double *out_host) {
  FM_CHECK(check_level(s, 0));
  if (!p || s->nq2 != 1 || s->nq1 != 0 || mode < 0 || mode > 2 || !out_host || !(p->alpha0 > 0.)) return fm_einval(__func__);
  if (mode == 1 && p->upwind != 0. && !s->vel) { fm_set_error("fm_heat_coupling: upwind term needs the velocity"); return FM_ESTATE; }
  if ((mode != 2 && !T_host) || (mode == 1 && !lam_host) || (mode == 2 && !m_host)) return fm_einval(__func__);
  const int top = (int)s->lev.size() - 1;
  Level &l = s->lev[top];
  hipStream_t st = s->ctx->stream;
  const fm_mesh &m = *s->mesh;
  FM_HIP(hipSetDevice(s->ctx->device));
  const int64_t nout = (int64_t)(mode == 1 ? m.dim : 1) * m.Ntop;
  double *dm = nullptr, *dout = nullptr;
  FM_CHECK(dalloc(&dout, nout));
  if (mode == 2 && dalloc(&dm, m.Ntop)) { dfree(dout); return FM_ENOMEM; }
  int rc = 0;
  if (mode != 2 && hipMemcpyAsync(l.t, T_host, l.n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess) rc = FM_EHIP;
  if (!rc && mode == 1 && hipMemcpyAsync(l.d, lam_host, l.n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess) rc = FM_EHIP;
  if (!rc && mode == 2 && hipMemcpyAsync(dm, m_host, m.Ntop * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess) rc = FM_EHIP;
  if (!rc) rc = fmk::fill<double>(st, nout, dout, 0.0);
  for (int k = 0; k < l.ncolor && !rc; k++) {
    int n = l.color_off[k + 1] - l.color_off[k];
    rc = fmk::coupling(st, m.dim, n, l.color_elems + l.color_off[k], l.conn, s->xyz, m.Ntop, s->d_dof_top, s->d_bc0, l.t,
                       l.d, dm, scale, mode, s->vel, p->upwind, p->rhocp, p->alpha0, s->d_tab, dout);
  }
  if (!rc && (hipMemcpyAsync(out_host, dout, nout * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
              hipStreamSynchronize(st) != hipSuccess)) rc = FM_EHIP;
  dfree(dout);
  dfree(dm);
  return rc ? (rc < 0 ? rc : FM_EHIP) : FM_OK;
}

extern "C" int fm_heat_nusselt(fm_system *s, const fm_heat_params *p, int axis, const double *T_host, double *J,
                               double *dJdT_host, double *dJdk_host) {
  FM_CHECK(check_level(s, 0));
  if (!p || !T_host || s->nq2 != 1 || s->nq1 != 0 || axis < 0 || axis >= s->mesh->dim || !(p->alpha0 > 0.)) return fm_einval(__func__);
  int top = (int)s->lev.size() - 1;
  Level &l = s->lev[top];
  hipStream_t st = s->ctx->stream;
  const fm_mesh &m = *s->mesh;
  FM_HIP(hipMemcpyAsync(l.t, T_host, l.n * sizeof(double), hipMemcpyHostToDevice, st));
  FM_CHECK(fmk::fill<double>(st, l.n, l.d, 0.0));
  double *jel = nullptr, *dk = nullptr;
  FM_CHECK(dalloc(&jel, l.nel));
  if (dalloc(&dk, l.nel)) { dfree(jel); return FM_ENOMEM; }
  int rc = 0;
  for (int k = 0; k < l.ncolor && !rc; k++) {
    int n = l.color_off[k + 1] - l.color_off[k];
    rc = fmk::heat_nusselt(st, m.dim, n, l.color_elems + l.color_off[k], l.conn, l.eldof, s->xyz, m.Ntop, s->d_dof_top,
                           s->vel, l.t, l.coef, p->alpha0, axis, s->d_tab, jel, dk, l.d);
  }
  std::vector<double> h(l.nel);
  if (!rc) {
    if (hipMemcpyAsync(h.data(), jel, l.nel * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess) rc = FM_EHIP;
    if (!rc && dJdk_host && hipMemcpyAsync(dJdk_host, dk, l.nel * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess) rc = FM_EHIP;
    if (!rc && dJdT_host && hipMemcpyAsync(dJdT_host, l.d, l.n * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess) rc = FM_EHIP;
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = FM_EHIP;
  }
  dfree(jel);
  dfree(dk);
  if (rc) return rc < 0 ? rc : FM_EHIP;
  double acc = 0.;
  for (int e = 0; e < l.nel; e++) acc += h[e];
  if (J) *J = acc;
  return FM_OK;
}
