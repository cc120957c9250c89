// Host check of the internal DoF order the library builds for every level (fm_level_order and
// fm_node_colours, exfemus_amd/csrc/order_host.cpp) on box meshes from the library's own mesh builder
// (exfemus_amd/csrc/gencase.cpp).  Everything the order promises is recomputed here from the node
// coordinates and the element geometry alone:
//   - ord is a permutation of the level's DoFs and rank its inverse;
//   - the DoFs of a node are adjacent, in variable order;
//   - the groups (goff) tile [0, ndof); all nodes of a group lie in one slab of 8 node planes along the axis
//     with the most cells and carry the group's signature (gmask); slabs never decrease from one group to
//     the next, and inside a group the nodes follow the lexicographic order of their lattice indices;
//   - element colours are the parities of the elements' cell indices, and two elements of one colour share
//     no node (the coloured Vanka sweep updates a colour's patches without conflicts);
//   - coupled systems (nq1 > 0): the signature of a node is the set of colours of the elements containing
//     it, and for every colour c the rows of the groups whose signature holds c are exactly the DoFs of the
//     elements of colour c (the coloured smoother forms the residual on those groups only);
//   - otherwise the signature is the node's parity class on the Q2 lattice.
// Prints one JSON line.
//   usage: order_check dim nx ny nz nlev nq2 nq1
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include "../../exfemus_amd/csrc/fm_mesh.h"

static int fail(const std::string &m) {
  std::printf("{\"ok\": false, \"error\": \"%s\"}\n", m.c_str());
  return 1;
}

int main(int argc, char **argv) {
  if (argc < 8) return fail("usage: dim nx ny nz nlev nq2 nq1");
  const int dim = atoi(argv[1]), nlev = atoi(argv[5]), nq2 = atoi(argv[6]), nq1 = atoi(argv[7]);
  const int nx0[3] = {atoi(argv[2]), atoi(argv[3]), dim == 3 ? atoi(argv[4]) : 1};
  const double lo[3] = {-1.0, 0.0, 0.25};
  const double hi[3] = {lo[0] + 1.5 * nx0[0], lo[1] + 0.7 * nx0[1], lo[2] + 1.1 * nx0[2]};
  fm_mesh m;
  if (fm_build_box(dim, nx0, lo, hi, nlev, &m) != 0) return fail("fm_build_box failed");
  const int N = m.Ntop, top = nlev - 1, nv = nq2 + nq1, nloc = nq2 * m.n2 + nq1 * m.n1;
  // lattice of the top-level nodes: spacing h, index q per axis; the slowest axis has the most cells (ties: the
  // higher axis), the fastest the fewest
  int64_t cells[3];
  double h[3];
  for (int d = 0; d < dim; d++) {
    cells[d] = (int64_t)nx0[d] << top;
    h[d] = (hi[d] - lo[d]) / (2.0 * (double)cells[d]);
  }
  std::vector<int> axes;
  for (int d = dim - 1; d >= 0; d--) axes.push_back(d);
  std::stable_sort(axes.begin(), axes.end(), [&](int a, int b) { return cells[a] > cells[b]; });
  std::vector<int64_t> q((size_t)N * 3, 0);
  for (int k = 0; k < N; k++)
    for (int d = 0; d < dim; d++) {
      const double t = (m.xyz[k + (size_t)d * N] - lo[d]) / h[d];
      q[3 * (size_t)k + d] = std::llround(t);
      if (std::fabs(t - (double)q[3 * (size_t)k + d]) > 1e-6) return fail("node off the Q2 lattice");
    }
  int64_t total_groups = 0, max_slab = 0;
  for (int L = 0; L < nlev; L++) {
    std::vector<int> dof;
    int64_t ndof = 0;
    fm_init_dof(m, L, nq2, nq1, dof, ndof);
    std::vector<unsigned char> nmask;
    if (nq1 > 0) fm_node_colours(m, L, nmask);
    std::vector<int> ord, rank, gmask;
    std::vector<int64_t> goff;
    std::string err;
    if (fm_level_order(m, L, nv, dof, ndof, ord, rank, nq1 > 0 ? &nmask : nullptr, goff, gmask, err) != 0)
      return fail("level " + std::to_string(L) + ": " + err);
    const std::string lv = "level " + std::to_string(L) + ": ";
    // permutation
    if ((int64_t)ord.size() != ndof || (int64_t)rank.size() != ndof) return fail(lv + "order size");
    std::vector<char> seen(ndof, 0);
    for (int64_t i = 0; i < ndof; i++) {
      if (ord[i] < 0 || ord[i] >= ndof || seen[ord[i]]) return fail(lv + "ord is not a permutation");
      seen[ord[i]] = 1;
      if (rank[ord[i]] != i) return fail(lv + "rank is not the inverse of ord");
    }
    // natural DoF -> (node, variable)
    std::vector<int> dnode(ndof, -1), dvar(ndof, -1);
    for (int iv = 0; iv < nv; iv++)
      for (int k = 0; k < N; k++) {
        const int d = dof[k + (size_t)iv * N];
        if (d >= 0) { dnode[d] = k; dvar[d] = iv; }
      }
    // the DoFs of a node adjacent, in variable order: walk the internal order node by node
    std::vector<int> inode;  // internal node sequence
    for (int64_t i = 0; i < ndof;) {
      const int k = dnode[ord[i]];
      int last = -1;
      inode.push_back(k);
      while (i < ndof && dnode[ord[i]] == k) {
        if (dvar[ord[i]] <= last) return fail(lv + "variables of a node out of order");
        last = dvar[ord[i]];
        i++;
      }
      if (last != nv - 1 && last != nq2 - 1) return fail(lv + "a node's DoFs are split");
    }
    {
      std::vector<int> u(inode);
      std::sort(u.begin(), u.end());
      if (std::adjacent_find(u.begin(), u.end()) != u.end()) return fail(lv + "a node's DoFs are not adjacent");
      if ((int)u.size() != m.nq2[L]) return fail(lv + "node count");
    }
    // element colours from geometry: cell index of the element's centre along each axis
    const int64_t span = (int64_t)2 << (top - L);  // element width in lattice steps
    std::vector<int> ecol(m.nel[L]);
    std::vector<int> cover((size_t)N * 8, 0);
    for (int e = 0; e < m.nel[L]; e++) {
      int col = 0;
      for (int d = 0; d < dim; d++) {
        int64_t s = 0;
        for (int a = 0; a < m.n2; a++) s += q[3 * (size_t)m.conn[L][(size_t)e * m.n2 + a] + d];
        if (s % m.n2) return fail(lv + "element centre off the lattice");
        const int64_t c = (s / m.n2 - span / 2) / span;
        col |= (int)(c & 1) << d;
      }
      if (col != fm_elem_colour(m, L, e)) return fail(lv + "element colour is not the parity of its cell index");
      ecol[e] = col;
      for (int a = 0; a < m.n2; a++) {
        int &cnt = cover[8 * (size_t)m.conn[L][(size_t)e * m.n2 + a] + col];
        if (++cnt > 1) return fail(lv + "two elements of one colour share a node");
      }
    }
    // signature of each node: colour set (coupled) or parity class on the lattice
    auto signature = [&](int k) {
      int sg = 0;
      if (nq1 > 0) {
        for (int c = 0; c < 8; c++) if (cover[8 * (size_t)k + c]) sg |= 1 << c;
      } else {
        for (int d : axes) sg = sg * 2 + (int)(q[3 * (size_t)k + d] & 1);
      }
      return sg;
    };
    if (nq1 > 0)
      for (int k = 0; k < m.nq2[L]; k++)
        if (nmask[k] != signature(k)) return fail(lv + "node colour signature");
    // groups
    const size_t ng = gmask.size();
    if (goff.size() != ng + 1 || goff.front() != 0 || goff.back() != ndof) return fail(lv + "groups do not tile the DoFs");
    std::vector<int> grp_of(ndof);
    int64_t prev_slab = -1;
    for (size_t g = 0; g < ng; g++) {
      if (goff[g + 1] <= goff[g]) return fail(lv + "empty group");
      const int64_t slab = q[3 * (size_t)dnode[ord[goff[g]]] + axes[0]] / 8;
      if (slab < prev_slab) return fail(lv + "slabs decrease");
      if (g > 0 && slab == prev_slab && gmask[g] == gmask[g - 1]) return fail(lv + "two adjacent groups with one key");
      prev_slab = slab;
      max_slab = std::max(max_slab, slab);
      int pk = -1;
      for (int64_t i = goff[g]; i < goff[g + 1]; i++) {
        grp_of[i] = (int)g;
        const int k = dnode[ord[i]];
        if (q[3 * (size_t)k + axes[0]] / 8 != slab) return fail(lv + "group spans two slabs");
        if (signature(k) != gmask[g]) return fail(lv + "node signature differs from its group's");
        if (pk >= 0 && k != pk) {
          bool less = false, eq = true;
          for (int d : axes) {
            const int64_t a = q[3 * (size_t)pk + d], b = q[3 * (size_t)k + d];
            if (a != b) { less = a < b; eq = false; break; }
          }
          if (eq || !less) return fail(lv + "nodes of a group not in lexicographic order");
        }
        pk = k;
      }
    }
    total_groups += (int64_t)ng;
    // coupled systems: rows of the groups holding colour c == DoFs of the elements of colour c
    if (nq1 > 0) {
      std::vector<int> eld(nloc);
      for (int c = 0; c < (1 << dim); c++) {
        std::vector<char> touched(ndof, 0), listed(ndof, 0);
        for (int e = 0; e < m.nel[L]; e++) {
          if (ecol[e] != c) continue;
          fm_elem_dofs(m, L, e, nq2, nq1, dof, eld.data());
          for (int a = 0; a < nloc; a++) if (eld[a] >= 0) touched[rank[eld[a]]] = 1;
        }
        for (size_t g = 0; g < ng; g++)
          if (gmask[g] & (1 << c))
            for (int64_t i = goff[g]; i < goff[g + 1]; i++) listed[i] = 1;
        if (touched != listed) return fail(lv + "colour " + std::to_string(c) + ": group rows differ from its elements' DoFs");
      }
    }
    if (L == top)
      std::printf("{\"ok\": true, \"ndof\": %lld, \"groups\": %zu, \"slabs\": %lld, \"levels\": %d, "
                  "\"groups_all_levels\": %lld}\n",
                  (long long)ndof, ng, (long long)max_slab + 1, nlev, (long long)total_groups);
  }
  return 0;
}
