// fm_mesh.h -- host-side mesh hierarchy, DoF maps, patterns and the internal DoF order (no device code:
// the library uploads what these build, the host checkers under tests/cpp verify it).
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include "../../include/femus_hip.h"

// Host-side mesh hierarchy (gencase result; src/MGMesh.C layout).
struct fm_mesh {
  int dim = 0, nlev = 0, n2 = 0, n1 = 0, nsides = 0, nb2 = 0, nch = 0, Ntop = 0;
  int nx0[3] = {1, 1, 1};
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  std::vector<int> nq2, nq1, nel;
  std::vector<double> xyz;                    // xyz[node + d*Ntop]
  std::vector<std::vector<int>> conn, neigh, parent, egrid;  // egrid: element grid index (3 ints)
  // lattice: built by fm_build_box (egrid, nx0, lo, hi describe the nodes); false for a hierarchy handed in by a
  // caller (fm_mesh_from_arrays), whose element colours and child maps are stored instead
  bool lattice = true;
  std::vector<std::vector<unsigned char>> ecol;  // per level: element colour (caller meshes)
  std::vector<std::vector<int>> kids;            // per level L >= 1: kids[L][p * nch + c] = child c of parent p
  std::vector<std::vector<unsigned char>> chid;  // per level L >= 1: index of each element among its parent's children
};

// gencase.cpp
int fm_build_box(int dim, const int nx0[3], const double lo[3], const double hi[3], int nlev, fm_mesh *m);
void fm_init_dof(const fm_mesh &m, int L, int nq2, int nq1, std::vector<int> &node_dof, int64_t &ndof);
void fm_pattern(const fm_mesh &m, int L, int nq2, int nq1, const std::vector<int> &dofL,
                std::vector<int64_t> &rp, std::vector<int> &ci);
void fm_transfer(const fm_mesh &m, int L, int nq2, int nq1, bool rest, const std::vector<int> &dof_row,
                 const std::vector<int> &dof_col, const std::vector<int> *dof_top, const std::vector<int> *bc0,
                 std::vector<int64_t> &rp, std::vector<int> &ci, std::vector<double> &v);
void fm_elem_dofs(const fm_mesh &m, int L, int e, int nq2, int nq1, const std::vector<int> &dof, int *out);
// node-level operators and the boundary mesh in gencase's order (the preprocessor's output files)
void fm_transfer_block(const fm_mesh &m, int L, int kind, bool rest,
                       std::vector<std::vector<std::pair<int, double>>> &blk);
int64_t fm_gencase_block(const fm_mesh &m, int L, int rkind, int ckind, std::vector<int> &len, std::vector<int> *pos);
void fm_gencase_transfer_kk(const fm_mesh &m, int L, bool rest, std::vector<std::vector<std::pair<int, double>>> &blk);
void fm_boundary_mesh(const fm_mesh &m, int L, std::vector<int> &nodes);
extern const int fm_hex27_ref[27][3];
extern const int fm_quad9_ref[9][2];
extern const int fm_face2d[4][3];
extern const int fm_face3d[6][9];

// Colour of element e of level L for the coloured Vanka sweep: the parities of its grid indices
// (2^dim colours).  Two elements of one colour are at least two elements apart along some axis, so they
// share no node.
inline int fm_elem_colour(const fm_mesh &m, int L, int e) {
  if (!m.lattice) return m.ecol[L][e];
  const int *g = &m.egrid[L][3 * (size_t)e];
  return (g[0] & 1) + 2 * (g[1] & 1) + (m.dim == 3 ? 4 * (g[2] & 1) : 0);
}
// number of element colours of level L
inline int fm_ncolours(const fm_mesh &m, int L) {
  if (m.lattice) return 1 << m.dim;
  int c = 0;
  for (unsigned char v : m.ecol[L]) c = v + 1 > c ? v + 1 : c;
  return c;
}

// A caller's hierarchy in MGMesh's layout (src/MGMesh.C:493-899): nlev levels, Ntop nodes whose coordinates are
// xyz[node + d Ntop]; level L holds the elements off_el[L] .. off_el[L+1]-1 of el_map (n2 top-level node ids per
// element in libMesh Quad9 / Hex27 local order) and el_neigh (nsides level-local neighbours, -1 on the boundary);
// nq2[L] / nq1[L] Q2 / Q1 nodes of level L (the first nq2[L] / nq1[L] top-level ids); parent / child the level
// L-1 parent of every element of L >= 1 and its index among the parent's 2^dim children (corner bits (c >> d) & 1
// along axis d of the parent's reference element), or both NULL: then they are inferred from the nested node sets.
// Validates, colours the elements (no two of one colour share a node) and builds the child maps.  Returns 0, or -1
// with err set.
int fm_mesh_from_arrays(int dim, int nlev, int Ntop, const double *xyz, const int *off_el, const int *el_map,
                        const int *el_neigh, const int *nq2, const int *nq1, const int *parent, const int *child,
                        fm_mesh *m, std::string &err);

// order_host.cpp
// Colour signature of every node of level L: bit c set when an element of colour c contains the node
// (indexed by top-level node number; nodes absent from level L stay 0).
void fm_node_colours(const fm_mesh &m, int L, std::vector<unsigned char> &nmask);
// Internal DoF order of level L (ord: internal -> natural DoF, rank: its inverse) and its row groups
// (goff: group boundaries in internal DoFs, gmask: each group's signature).  Returns 0, or -5 with err set
// when the DoF map is inconsistent.
int fm_level_order(const fm_mesh &m, int L, int nv, const std::vector<int> &dofL, int64_t ndof,
                   std::vector<int> &ord, std::vector<int> &rank, const std::vector<unsigned char> *nmask,
                   std::vector<int64_t> &goff, std::vector<int> &gmask, std::string &err);
