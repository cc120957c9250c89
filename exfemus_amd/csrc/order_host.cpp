// order_host.cpp -- internal DoF order of a level and the node colour signatures it groups by
// (host only; the library uploads the order, tests/cpp/order_check.cpp verifies it).
#include <algorithm>
#include <cmath>
#include "fm_mesh.h"

void fm_node_colours(const fm_mesh &m, int L, std::vector<unsigned char> &nmask) {
  nmask.assign(m.Ntop, 0);
  for (int e = 0; e < m.nel[L]; e++) {
    const int cl = fm_elem_colour(m, L, e);
    for (int a = 0; a < m.n2; a++) nmask[m.conn[L][(size_t)e * m.n2 + a]] |= (unsigned char)(1 << cl);
  }
}

// Internal DoF order of level L: the level's nodes in lexicographic order of
// their grid coordinates (the axis with the most nodes slowest, so a row's
// neighbours lie within ~2 node planes of the smallest cross-section), all
// variables of a node adjacent (Q2 variables, then the Q1 variable where the
// node carries one).  This keeps the columns of a SELL slice inside few
// narrow windows, so the gather of x hits cache and the column offsets fit
// 16 bits.
//
// With nmask (per node: bit c set when an element of colour c contains the
// node) the domain is cut into slabs of SLAB node planes along the slowest
// axis and, inside each slab, the nodes are grouped by that colour signature,
// lexicographic inside a group; goff / gmask return the groups' DoF ranges
// and signatures.  The rows a colour's elements touch are then whole groups,
// so a coloured smoother can form the residual on those rows only, and the
// slabs keep the gathered part of x local (grouping over the whole domain
// made the full product 40% slower on the 96x24x24 channel).  Without nmask
// the groups are the slabs' parity classes (vertex / edge / face / interior
// node of the Q2 lattice).
// Node keys of a level without a lattice (a caller's mesh): a Cuthill-McKee numbering of the node graph (nodes
// joined when they share an element), started from a node of least degree and visiting neighbours in order of
// increasing degree; slab = breadth-first level / SLAB, so the groups play the role of the lattice's slabs of
// node planes.
static void graph_keys(const fm_mesh &m, int L, std::vector<int64_t> &key, std::vector<int> &slab, int SLAB) {
  const int nn = m.nq2[L], ne = m.nel[L], n2 = m.n2;
  std::vector<int64_t> off(nn + 1, 0);
  for (int e = 0; e < ne; e++)
    for (int a = 0; a < n2; a++) off[m.conn[L][(size_t)e * n2 + a] + 1]++;
  for (int k = 0; k < nn; k++) off[k + 1] += off[k];
  std::vector<int> els(off[nn]);
  {
    std::vector<int64_t> f(off.begin(), off.end() - 1);
    for (int e = 0; e < ne; e++)
      for (int a = 0; a < n2; a++) els[f[m.conn[L][(size_t)e * n2 + a]]++] = e;
  }
  auto neighbours = [&](int k, std::vector<int> &out) {
    out.clear();
    for (int64_t q = off[k]; q < off[k + 1]; q++)
      for (int a = 0; a < n2; a++) out.push_back(m.conn[L][(size_t)els[q] * n2 + a]);
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
  };
  std::vector<int> deg(nn), nb;
  for (int k = 0; k < nn; k++) { neighbours(k, nb); deg[k] = (int)nb.size(); }
  std::vector<int> num(nn, -1), lvl(nn, 0);
  int cnt = 0;
  for (int start = 0; start < nn; start++) {  // every connected component
    if (num[start] >= 0) continue;
    int s0 = start;
    for (int k = 0; k < nn; k++) if (num[k] < 0 && deg[k] < deg[s0]) s0 = k;
    std::vector<int> queue{s0};
    num[s0] = cnt++;
    for (size_t h = 0; h < queue.size(); h++) {
      const int k = queue[h];
      neighbours(k, nb);
      std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] < deg[b]; });
      for (int j : nb)
        if (num[j] < 0) { num[j] = cnt++; lvl[j] = lvl[k] + 1; queue.push_back(j); }
    }
  }
  for (int k = 0; k < nn; k++) { key[k] = num[k]; slab[k] = lvl[k] / SLAB; }
}

int fm_level_order(const fm_mesh &m, int L, int nv, const std::vector<int> &dofL, int64_t ndof,
                   std::vector<int> &ord, std::vector<int> &rank, const std::vector<unsigned char> *nmask,
                   std::vector<int64_t> &goff, std::vector<int> &gmask, std::string &err) {
  const int N = m.Ntop, dim = m.dim;
  std::vector<int> nodes;
  nodes.reserve(m.nq2[L]);
  for (int k = 0; k < N; k++) if (dofL[k] >= 0) nodes.push_back(k);
  if (!m.lattice) {
    // no lattice: Cuthill-McKee keys, slabs of breadth-first levels, groups by colour signature (coupled systems)
    std::vector<int64_t> key(N, 0);
    std::vector<int> slab(N, 0);
    graph_keys(m, L, key, slab, 4);
    auto sig = [&](int k) -> int { return nmask ? (int)(*nmask)[k] : 0; };
    auto grp = [&](int k) -> int64_t { return (int64_t)slab[k] * 256 + sig(k); };
    std::stable_sort(nodes.begin(), nodes.end(), [&](int a, int b) {
      return grp(a) != grp(b) ? grp(a) < grp(b) : key[a] < key[b];
    });
    ord.clear();
    goff.assign(1, 0);
    gmask.clear();
    for (size_t t = 0; t < nodes.size(); t++) {
      const int k = nodes[t];
      if (t > 0 && grp(k) != grp(nodes[t - 1])) { goff.push_back((int64_t)ord.size()); gmask.push_back(sig(nodes[t - 1])); }
      for (int iv = 0; iv < nv; iv++) {
        int d = dofL[k + (size_t)iv * N];
        if (d >= 0) ord.push_back(d);
      }
    }
    goff.push_back((int64_t)ord.size());
    gmask.push_back(nodes.empty() ? 0 : sig(nodes.back()));
    if ((int64_t)ord.size() != ndof) { err = "level order does not cover the DoFs"; return FM_ESTATE; }
    rank.assign(ndof, -1);
    for (int64_t i = 0; i < ndof; i++) {
      if (ord[i] < 0 || ord[i] >= ndof || rank[ord[i]] >= 0) { err = "level order is not a permutation"; return FM_ESTATE; }
      rank[ord[i]] = (int)i;
    }
    return 0;
  }
  int axes[3] = {0, 1, 2};
  int64_t cnt[3];
  double h[3];
  for (int d = 0; d < dim; d++) {
    cnt[d] = (int64_t)m.nx0[d] << (m.nlev - 1);
    h[d] = (m.hi[d] - m.lo[d]) / (2.0 * (double)cnt[d]);
  }
  std::stable_sort(axes, axes + dim, [&](int a, int b) { return cnt[a] > cnt[b] || (cnt[a] == cnt[b] && a > b); });
  const int SLAB = 8;
  std::vector<int64_t> key(N, 0);
  std::vector<int> slab(N, 0), par(N, 0);
  for (int k : nodes) {
    int64_t kk = 0;
    int pc = 0;
    for (int t = 0; t < dim; t++) {
      int d = axes[t];
      int64_t q = std::llround((m.xyz[k + (size_t)d * N] - m.lo[d]) / h[d]);
      kk = kk * (2 * cnt[d] + 1) + q;
      pc = pc * 2 + (int)(q & 1);
      if (t == 0) slab[k] = (int)(q / SLAB);
    }
    key[k] = kk;
    par[k] = pc;
  }
  // group = (slab, signature): the colour signature for coupled systems, the node's parity class otherwise
  auto sig = [&](int k) -> int { return nmask ? (int)(*nmask)[k] : par[k]; };
  auto grp = [&](int k) -> int64_t { return (int64_t)slab[k] * 256 + sig(k); };
  std::stable_sort(nodes.begin(), nodes.end(), [&](int a, int b) {
    return grp(a) != grp(b) ? grp(a) < grp(b) : key[a] < key[b];
  });
  ord.clear();
  ord.reserve(ndof);
  goff.assign(1, 0);
  gmask.clear();
  for (size_t t = 0; t < nodes.size(); t++) {
    const int k = nodes[t];
    if (t > 0 && grp(k) != grp(nodes[t - 1])) { goff.push_back((int64_t)ord.size()); gmask.push_back(sig(nodes[t - 1])); }
    for (int iv = 0; iv < nv; iv++) {
      int d = dofL[k + (size_t)iv * N];
      if (d >= 0) ord.push_back(d);
    }
  }
  goff.push_back((int64_t)ord.size());
  gmask.push_back(nodes.empty() ? 0 : sig(nodes.back()));
  if ((int64_t)ord.size() != ndof) { err = "level order does not cover the DoFs"; return FM_ESTATE; }
  rank.assign(ndof, -1);
  for (int64_t i = 0; i < ndof; i++) {
    if (ord[i] < 0 || ord[i] >= ndof || rank[ord[i]] >= 0) { err = "level order is not a permutation"; return FM_ESTATE; }
    rank[ord[i]] = (int)i;
  }
  return 0;
}
