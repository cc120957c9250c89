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
extern const int fm_hex27_ref[27][3];
extern const int fm_quad9_ref[9][2];
extern const int fm_face2d[4][3];
extern const int fm_face3d[6][9];

// Colour of element e of level L for the coloured Vanka sweep: the parities of its grid indices
// (2^dim colours).  Two elements of one colour are at least two elements apart along some axis, so they
// share no node.
inline int fm_elem_colour(const fm_mesh &m, int L, int e) {
  const int *g = &m.egrid[L][3 * (size_t)e];
  return (g[0] & 1) + 2 * (g[1] & 1) + (m.dim == 3 ? 4 * (g[2] & 1) : 0);
}

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
