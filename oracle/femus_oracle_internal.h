/* femus_oracle_internal.h -- TEST INFRASTRUCTURE (oracle): private layout. */
#ifndef FEMUS_ORACLE_INTERNAL_H
#define FEMUS_ORACLE_INTERNAL_H
#include <stdint.h>

struct om_mesh {
  int dim, nlev, n2, n1, nsides, nb2, nch;
  int nx0[3], G[3];
  double lo[3], hi[3];
  int Ntop;
  int *nq2, *nq1;          /* Q2 / Q1 node counts per level (prefixes)      */
  double *xyz;             /* xyz[node + d*Ntop]                            */
  int *nel;                /* elements per level                            */
  int **corner;            /* integer lower corner on the top grid          */
  int **parent, **conn, **neigh;
  int *gnew;               /* top-grid point -> node index                  */
  int *node_lid, *node_lev, *node_levp;
};

extern const int om_quad9_ref[9][2];
extern const int om_hex27_ref[27][3];
extern const int om_surf2d[4][3];
extern const int om_surf3d[6][9];
int om_ref_coord(int dim, int k, int d);
int om_face_node(int dim, int s, int k);
void om_side_dir(int dim, int s, int *axis, int *sgn);
double om_shape(int kind, int dim, int k, const double *xi);
double om_dshape(int kind, int dim, int k, int dir, const double *xi);
void om_gauss_point(int dim, int qp, double *xi, double *w);
int om_elem_dofs(const struct om_mesh *m, int L, int e, int nq2, int nq1, const int *node_dof, int *dofs);
#endif
