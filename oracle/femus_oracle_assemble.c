/*
 * femus_oracle_assemble.c -- TEST INFRASTRUCTURE (oracle).  Element
 * Jacobian/RHS assembly of the heat (Q2) and coupled Navier-Stokes (Q2-Q1)
 * systems, restated from:
 *   src/MGFE.C:1066-1110 (Jac3D), 1143-1166 (Jac2D), 1357-1379 (JacSur2D),
 *              1424-1464 (JacSur3D), 1496-1519 (get_dphi_gl_g), 1603-1645 (normal_g)
 *   template_appl/nse/SRC/MGSolverT.C:93-525   (MGSolT::GenMatRhs)
 *   template_appl/nse/SRC/MGSolverNS.C:108-717 (MGSolNS::GenMatRhs, NS_EQUATIONS==1)
 *   src/MGSolverBase.C:336-353 (get_el_sol: fields read from top-level x_old)
 * Additions of this build (no reference counterpart): per-element k / nu
 * scaling, the steady switch time_coeff, a constant volume source, the
 * Boussinesq body force buoy * T from a nodal temperature field.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "femus_oracle.h"
#include "femus_oracle_internal.h"

typedef struct {
  int dim, n2, n1, ng, ngb, nb2;
  double phi2[27 * 27], dphi2[3 * 27 * 27], w[27];
  double phi1[8 * 27];
  double bphi2[9 * 9], bdphi2[2 * 9 * 9], bw[9]; /* face (dim-1) Q2 tables */
} fe_t;

static void fe_init(fe_t *f, int dim) {
  f->dim = dim;
  f->n2 = om_fe_tables(2, dim, f->phi2, f->dphi2, f->w, &f->ng);
  f->n1 = om_fe_tables(1, dim, f->phi1, NULL, NULL, NULL);
  f->nb2 = om_fe_tables(2, dim - 1, f->bphi2, f->bdphi2, f->bw, &f->ngb);
}

/* Volume Jacobian, MGFE::Jac2D / Jac3D (xx[node + d*n2]) */
static double jac(const fe_t *f, int qp, const double *xx, double *IJ) {
  int ns = f->n2, ng = f->ng, off = ns * ng;
  if (f->dim == 2) {
    double x_xi = 0., x_eta = 0., y_xi = 0., y_eta = 0.;
    for (int s = 0; s < ns; s++) {
      int sng = s * ng + qp;
      x_xi += xx[s] * f->dphi2[sng]; x_eta += xx[s] * f->dphi2[sng + off];
      y_xi += xx[s + ns] * f->dphi2[sng]; y_eta += xx[s + ns] * f->dphi2[sng + off];
    }
    double det = x_xi * y_eta - y_xi * x_eta, idet = 1. / det;
    IJ[0] = y_eta * idet; IJ[1] = -y_xi * idet; IJ[2] = -x_eta * idet; IJ[3] = x_xi * idet;
    return det;
  }
  double x_xi = 0., x_eta = 0., x_zeta = 0., y_xi = 0., y_eta = 0., y_zeta = 0., z_xi = 0., z_eta = 0., z_zeta = 0.;
  for (int s = 0; s < ns; s++) {
    int sng = s * ng + qp;
    double a = f->dphi2[sng], b = f->dphi2[sng + off], c = f->dphi2[sng + 2 * off];
    x_xi += xx[s] * a; x_eta += xx[s] * b; x_zeta += xx[s] * c;
    y_xi += xx[s + ns] * a; y_eta += xx[s + ns] * b; y_zeta += xx[s + ns] * c;
    z_xi += xx[s + 2 * ns] * a; z_eta += xx[s + 2 * ns] * b; z_zeta += xx[s + 2 * ns] * c;
  }
  double det = x_xi * (y_eta * z_zeta - y_zeta * z_eta) - x_eta * (y_xi * z_zeta - y_zeta * z_xi) +
               x_zeta * (y_xi * z_eta - y_eta * z_xi);
  double idet = 1. / det;
  IJ[0] = (y_eta * z_zeta - y_zeta * z_eta) * idet;
  IJ[3] = -(x_eta * z_zeta - x_zeta * z_eta) * idet;
  IJ[6] = (y_zeta * x_eta - x_zeta * y_eta) * idet;
  IJ[1] = -(y_xi * z_zeta - y_zeta * z_xi) * idet;
  IJ[4] = (x_xi * z_zeta - x_zeta * z_xi) * idet;
  IJ[7] = -(y_zeta * x_xi - y_xi * x_zeta) * idet;
  IJ[2] = (z_eta * y_xi - y_eta * z_xi) * idet;
  IJ[5] = -(x_xi * z_eta - x_eta * z_xi) * idet;
  IJ[8] = (x_xi * y_eta - y_xi * x_eta) * idet;
  return det;
}

/* global derivatives dphi[eln + idim*n2] (MGFE::get_dphi_gl_g) */
static void dphi_global(const fe_t *f, int qp, const double *IJ, double *dphi) {
  int ns = f->n2, ng = f->ng, off = ns * ng, dim = f->dim;
  for (int e = 0; e < ns; e++) {
    double g[3];
    for (int d = 0; d < dim; d++) g[d] = f->dphi2[e * ng + qp + d * off];
    for (int i = 0; i < dim; i++) {
      double sum = 0.;
      for (int j = 0; j < dim; j++) sum += IJ[j + i * dim] * g[j];
      dphi[e + i * ns] = sum;
    }
  }
}

/* surface Jacobian determinant, MGFE::JacSur2D / JacSur3D (xb[node + d*nb2]) */
static double jac_sur(const fe_t *f, int qp, const double *xb) {
  int nb = f->nb2, ng = f->ngb;
  if (f->dim == 2) {
    double dx = 0., dy = 0.;
    for (int s = 0; s < nb; s++) { double d = f->bdphi2[s * ng + qp]; dx += xb[s] * d; dy += xb[s + nb] * d; }
    return sqrt(dx * dx + dy * dy);
  }
  int off = nb * ng;
  double dxdxi = 0., dxdeta = 0., dydxi = 0., dydeta = 0., dzdxi = 0., dzdeta = 0.;
  for (int s = 0; s < nb; s++) {
    int sng = s * ng + qp;
    double a = f->bdphi2[sng], b = f->bdphi2[sng + off];
    dxdxi += xb[s] * a; dxdeta += xb[s] * b;
    dydxi += xb[s + nb] * a; dydeta += xb[s + nb] * b;
    dzdxi += xb[s + 2 * nb] * a; dzdeta += xb[s + 2 * nb] * b;
  }
  double c1 = dxdxi * dydeta - dxdeta * dydxi, c2 = dydxi * dzdeta - dydeta * dzdxi, c3 = dzdxi * dxdeta - dzdeta * dxdxi;
  return sqrt(c1 * c1 + c2 * c2 + c3 * c3);
}

/* MGFE::normal_g */
static void normal_g(int dim, int nb, const double *xb, double *n) {
  if (dim == 2) {
    double t0 = xb[1] - xb[0], t1 = xb[1 + nb] - xb[nb];
    n[0] = t1; n[1] = -t0;
  } else {
    double a[3], b[3];
    for (int i = 0; i < 3; i++) { a[i] = xb[1 + i * nb] - xb[i * nb]; b[i] = xb[3 + i * nb] - xb[i * nb]; }
    n[0] = a[1] * b[2] - a[2] * b[1]; n[1] = a[2] * b[0] - a[0] * b[2]; n[2] = a[0] * b[1] - a[1] * b[0];
  }
  double mm = 0.;
  for (int d = 0; d < dim; d++) mm += n[d] * n[d];
  mm = sqrt(mm);
  for (int d = 0; d < dim; d++) n[d] /= mm;
}

static void elem_geom(const om_mesh *m, int L, int e, double *xx, double *h_eff) {
  int dim = m->dim, n2 = m->n2, N = m->Ntop;
  for (int d = 0; d < dim; d++)
    for (int k = 0; k < n2; k++) xx[k + d * n2] = m->xyz[m->conn[L][e * n2 + k] + d * N];
  for (int d = 0; d < dim; d++) {  /* MGSolverT.C:193-198 */
    double h = fabs(xx[d * n2 + 2 + 4 * (dim - 2)] - xx[d * n2 + 0]);
    double h1 = fabs(xx[d * n2 + 3 + 4 * (dim - 2)] - xx[d * n2 + 1]);
    h_eff[d] = h < h1 ? h1 : h;
  }
}

/* ------------------------------------------------------------------------ */
int om_heat_element(const om_mesh *m, int L, int iel, int mode, const om_heat_params *p,
                    const int *node_dof_top, const int *bc0, const int *bc1,
                    const double *x_old_top, double k_e, const double *vel_top,
                    double *KeM, double *FeM, int *dofs, const int *node_dof_L) {
  static fe_t F2, F3; static int init = 0;
  if (!init) { fe_init(&F2, 2); fe_init(&F3, 3); init = 1; }
  const fe_t *f = m->dim == 2 ? &F2 : &F3;
  int dim = m->dim, n2 = m->n2, N = m->Ntop;
  double xx[81], h_eff[3], IJ[9], dphi[81], Told[27], vel_n[3][27];
  int bc_vol[27], bc_bd[27];
  const int *conn = m->conn[L] + iel * n2;
  elem_geom(m, L, iel, xx, h_eff);
  for (int k = 0; k < n2; k++) {
    int kd = node_dof_top[conn[k]];
    bc_vol[k] = bc0[kd]; bc_bd[k] = bc1[kd];
    Told[k] = x_old_top[kd];
    if (node_dof_L) dofs[k] = node_dof_L[conn[k]];
    for (int d = 0; d < dim; d++) vel_n[d][k] = vel_top ? vel_top[conn[k] + d * N] : p->vel_const[d];
  }
  memset(KeM, 0, sizeof(double) * n2 * n2);
  memset(FeM, 0, sizeof(double) * n2);
  const double rhocp = p->rhocp, dt = p->dt, tc = p->time_coeff;
  for (int qp = 0; qp < f->ng; qp++) {
    double det = jac(f, qp, xx, IJ);
    double JxW = det * f->w[qp];
    const double *phi = f->phi2;  /* phi[s*ng+qp] */
    dphi_global(f, qp, IJ, dphi);
    double alpha_eff = p->alpha0 * k_e;
    double T_g = 0., vel_g[3] = {0., 0., 0.};
    for (int s = 0; s < n2; s++) {
      double ph = phi[s * f->ng + qp];
      T_g += ph * Told[s];
      for (int d = 0; d < dim; d++) vel_g[d] += ph * vel_n[d][s];
    }
    double mod2 = 1.e-20;
    for (int d = 0; d < dim; d++) mod2 += vel_g[d] * vel_g[d];
    mod2 = sqrt(mod2);
    double f_up[3];
    for (int d = 0; d < dim; d++) {  /* MGSolverT.C:276-279 */
      double Pe = 0.5 * mod2 * h_eff[d] / p->alpha0;
      f_up[d] = p->upwind * rhocp * 0.5 * (1. / tanh(Pe) - 1. / Pe) * h_eff[d] / mod2;
    }
    for (int i = 0; i < n2; i++) {  /* MGSolverT.C:300-345 */
      double phii = phi[i * f->ng + qp];
      double dtxJxW = JxW * bc_vol[i];
      if (mode == 1) FeM[i] += dtxJxW * (tc * rhocp * T_g * phii / dt + p->source * phii);
      for (int j = 0; j < n2; j++) {
        double phij = phi[j * f->ng + qp];
        double Adv = 0., Lap = 0.;
        for (int d = 0; d < dim; d++) {
          double dj = dphi[j + d * n2];
          Adv += vel_g[d] * dj;
          Lap += (alpha_eff + f_up[d] * vel_g[d] * vel_g[d]) * dj * dphi[i + d * n2];
        }
        KeM[i * n2 + j] += dtxJxW * (tc * rhocp * phii * phij / dt + rhocp * Adv * phii + Lap);
      }
    }
  }
  /* boundary faces (MGSolverT.C:366-507) */
  for (int s = 0; s < m->nsides; s++) {
    if (m->neigh[L][iel * m->nsides + s] != -1) continue;
    int nb = f->nb2, sur[9];
    double xb[27];
    for (int k = 0; k < nb; k++) {
      sur[k] = om_face_node(dim, s, k);
      for (int d = 0; d < dim; d++) xb[k + d * nb] = xx[sur[k] + d * n2];
    }
    int cen = sur[nb - 1];
    if (bc_vol[cen] == 0) {
      int bc_s = bc_bd[cen];
      double detb = jac_sur(f, f->ngb - 1, xb);
      double Ipen = detb / dt;
      for (int k = 0; k < nb; k++) {
        int lv = sur[k];
        int bc_beta = bc_s % 2;
        if (mode == 1) FeM[lv] += bc_beta * Ipen * Told[lv];
        KeM[lv * n2 + lv] += Ipen;
      }
    } else if (bc_bd[cen] != 0) {
      double alpha_b = p->alpha0;
      for (int qp = 0; qp < f->ngb; qp++) {
        double detb = jac_sur(f, qp, xb);
        double JxW = detb * f->bw[qp];
        for (int i = 0; i < nb; i++) {
          double phii = f->bphi2[i * f->ngb + qp];
          int lei = sur[i];
          int bcs = bc_bd[lei], bcv = bc_vol[lei];
          double dtxJxW = JxW * bcv;
          int bc_alpha = (bcs & 2) >> 1;
          for (int j = 0; j < nb; j++)
            KeM[lei * n2 + sur[j]] += dtxJxW * bc_alpha * alpha_b * phii * f->bphi2[j * f->ngb + qp];
        }
      }
    }
  }
  return n2;
}

/* ------------------------------------------------------------------------ */
int om_ns_element(const om_mesh *m, int L, int iel, int mode, const om_ns_params *p,
                  const int *node_dof_top, const int *bc0, const int *bc1,
                  const double *x_old_top, double nu_e,
                  double *KeM, double *FeM, int *dofs, const int *node_dof_L) {
  static fe_t F2, F3; static int init = 0;
  if (!init) { fe_init(&F2, 2); fe_init(&F3, 3); init = 1; }
  const fe_t *f = m->dim == 2 ? &F2 : &F3;
  int dim = m->dim, n2 = m->n2, n1 = m->n1, N = m->Ntop, NF = n2;
  int nl = dim * n2 + n1;
  double xx[81], h_eff[3], IJ[9], dphi[81], uold[3][27], Tn[27];
  int bc_vol[4 * 27], bc_bd[4 * 27];
  const int *conn = m->conn[L] + iel * n2;
  elem_geom(m, L, iel, xx, h_eff);
  for (int k = 0; k < n2; k++) Tn[k] = p->Tn ? p->Tn[conn[k]] : 0.;
  for (int iv = 0; iv < dim; iv++)
    for (int k = 0; k < n2; k++) {
      int kd = node_dof_top[conn[k] + iv * N];
      bc_vol[k + iv * NF] = bc0[kd]; bc_bd[k + iv * NF] = bc1[kd];
      uold[iv][k] = x_old_top[kd];
      if (node_dof_L) dofs[k + iv * n2] = node_dof_L[conn[k] + iv * N];
    }
  for (int k = 0; k < n1; k++) {
    int kd = node_dof_top[conn[k] + dim * N];
    bc_vol[k + dim * NF] = bc0[kd]; bc_bd[k + dim * NF] = bc1[kd];
    if (node_dof_L) dofs[dim * n2 + k] = node_dof_L[conn[k] + dim * N];
  }
  memset(KeM, 0, sizeof(double) * nl * nl);
  memset(FeM, 0, sizeof(double) * nl);
  const double rho = 1., dt = p->dt, tc = p->time_coeff;
  for (int qp = 0; qp < f->ng; qp++) {  /* MGSolverNS.C:285-495 */
    double det = jac(f, qp, xx, IJ);
    double JxW = det * f->w[qp];
    const double *phi = f->phi2;
    dphi_global(f, qp, IJ, dphi);
    double vel_g[3] = {0., 0., 0.}, T_g = 0.;
    for (int s = 0; s < n2; s++) {
      double ph = phi[s * f->ng + qp];
      for (int d = 0; d < dim; d++) vel_g[d] += ph * uold[d][s];
      T_g += ph * Tn[s];
    }
    double mod2 = 1.e-20;
    for (int d = 0; d < dim; d++) mod2 += vel_g[d] * vel_g[d];
    mod2 = sqrt(mod2);
    double f_up[3];
    for (int d = 0; d < dim; d++) {
      double Pe = 0.5 * mod2 * h_eff[d] / p->IRe;
      f_up[d] = p->upwind * rho * 0.5 * (1. / tanh(Pe) - 1. / Pe) * h_eff[d] / mod2;
    }
    double IRe_eff = nu_e;
    for (int i = 0; i < n2; i++) {
      double phii = phi[i * f->ng + qp];
      for (int iv = 0; iv < dim; iv++) {
        int indx = i + iv * n2;
        double dtxJxW = JxW * bc_vol[i + iv * NF];
        if (mode == 1)
          FeM[indx] += dtxJxW * (tc * rho * vel_g[iv] * phii / dt + rho * p->IFr * p->dirg[iv] * phii + p->src[iv] * phii +
                                 p->buoy[iv] * T_g * phii);
        for (int j = 0; j < n2; j++) {
          double phij = phi[j * f->ng + qp];
          double Lap = 0., Adv = 0.;
          for (int k = 0; k < dim; k++) {
            double dj = dphi[j + k * n2];
            Adv += vel_g[k] * dj * phii;
            Lap += (IRe_eff + f_up[k] * vel_g[k] * vel_g[k]) * (dj * dphi[i + k * n2]);
          }
          KeM[indx * nl + j + iv * n2] += dtxJxW * rho *
              (tc * phij * phii / dt + Adv + Lap +
               (IRe_eff + f_up[iv] * vel_g[iv] * vel_g[iv]) * (dphi[j + iv * n2] * dphi[i + iv * n2]));
          int i1 = (iv + 1) % dim;
          KeM[indx * nl + j + i1 * n2] += dtxJxW * rho * (IRe_eff + f_up[iv] * vel_g[iv] * vel_g[i1]) *
                                          dphi[j + iv * n2] * dphi[i + i1 * n2];
          if (dim == 3) {
            int i2 = (iv + 2) % dim;
            KeM[indx * nl + j + i2 * n2] += dtxJxW * rho * (IRe_eff + f_up[iv] * vel_g[iv] * vel_g[i2]) *
                                            dphi[j + iv * n2] * dphi[i + i2 * n2];
          }
        }
        for (int j = 0; j < n1; j++)  /* B^T, MGSolverNS.C:438-452 */
          KeM[indx * nl + j + dim * n2] += dtxJxW * (-f->phi1[j * f->ng + qp] * dphi[i + iv * n2]);
      }
    }
    for (int i = 0; i < n1; i++) {  /* B, MGSolverNS.C:458-489 */
      int indx = i + n2 * dim;
      double psii = f->phi1[i * f->ng + qp];
      for (int j = 0; j < n2; j++)
        for (int jv = 0; jv < dim; jv++)
          KeM[indx * nl + j + jv * n2] += JxW * rho * (psii * dphi[j + jv * n2]);
    }
  }
  /* boundary (MGSolverNS.C:503-690) */
  for (int s = 0; s < m->nsides; s++) {
    if (m->neigh[L][iel * m->nsides + s] != -1) continue;
    int nb = f->nb2, sur[9], elb[9];
    double xb[27], nrm[3];
    for (int k = 0; k < nb; k++) {
      sur[k] = om_face_node(dim, s, k);
      elb[k] = conn[sur[k]];
      for (int d = 0; d < dim; d++) xb[k + d * nb] = xx[sur[k] + d * n2];
    }
    normal_g(dim, nb, xb, nrm);
    for (int iv = 0; iv < dim; iv++) {
      int cen = sur[nb - 1] + iv * NF;
      if (bc_vol[cen] == 0) {
        double detb = jac_sur(f, f->ngb - 1, xb);
        double Ipen = detb / dt;
        for (int i = 0; i < nb; i++) {
          int row = sur[i] + iv * n2;
          int bcs = bc_bd[sur[i] + iv * NF];
          int bc_rhs = (bcs & 4) >> 2, bc_tg = (bcs & 2) >> 1, bc_n = bcs % 2;
          double vo[3], vdotn = 0.;
          for (int k = 0; k < dim; k++) {
            vo[k] = x_old_top[node_dof_top[elb[i] + k * N]];
            vdotn += nrm[k] * vo[k];
          }
          if (mode == 1)
            FeM[row] += bc_rhs * Ipen * ((1 - bc_n) * (1 - bc_tg) * vo[iv] + bc_n * vdotn * nrm[iv] +
                                         bc_tg * (vo[iv] - vdotn * nrm[iv]));
          KeM[row * nl + row] += ((1 - bc_n) * (1 - bc_tg) + bc_tg) * Ipen + Ipen * nrm[iv] * nrm[iv] * (bc_n - bc_tg);
          for (int jv = iv + 1; jv < dim; jv++)
            KeM[row * nl + sur[i] + jv * n2] += Ipen * nrm[jv] * nrm[iv] * (bc_n - bc_tg);
        }
      } else if (bc_bd[cen] != 0) {
        const double alpha_t = 1.;
        for (int qp = 0; qp < f->ngb; qp++) {
          double detb = jac_sur(f, qp, xb);
          double JxW = detb * f->bw[qp];
          for (int i = 0; i < nb; i++) {
            double phii = f->bphi2[i * f->ngb + qp];
            int row = sur[i] + iv * n2;
            int bcs = bc_bd[sur[i] + iv * NF], bcv = bc_vol[sur[i] + iv * NF];
            double dtJxW = JxW * (bcv % 2);
            int bc_tg = (bcs & 2) >> 1, bc_n = bcs % 2;
            for (int j = 0; j < nb; j++) {
              double phij = f->bphi2[j * f->ngb + qp];
              KeM[row * nl + row] += dtJxW * bc_tg * alpha_t * phij * phii +
                                     dtJxW * alpha_t * nrm[iv] * nrm[iv] * phij * phii * (bc_n - bc_tg);
              for (int jv = iv + 1; jv < dim; jv++)
                KeM[row * nl + sur[j] + jv * n2] += dtJxW * alpha_t * nrm[jv] * nrm[iv] * phij * phii * (bc_n - bc_tg);
            }
          }
        }
      }
    }
  }
  return nl;
}

static int64_t find_pos(const int64_t *rp, const int *ci, int row, int col) {
  int64_t lo = rp[row], hi = rp[row + 1] - 1;
  while (lo <= hi) {
    int64_t mid = (lo + hi) >> 1;
    if (ci[mid] == col) return mid;
    if (ci[mid] < col) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

int om_heat_genmatrhs(const om_mesh *m, int L, int mode, const om_heat_params *p,
                      const int *node_dof_L, const int *node_dof_top,
                      const int *bc0, const int *bc1, const double *x_old_top,
                      const double *k_elem, const double *vel_top,
                      const int64_t *rowptr, const int *colind, double *val, double *rhs) {
  int n2 = m->n2, n = m->nq2[L];
  double KeM[27 * 27], FeM[27];
  int dofs[27];
  memset(val, 0, sizeof(double) * rowptr[n]);
  if (mode == 1) memset(rhs, 0, sizeof(double) * n);
  for (int e = 0; e < m->nel[L]; e++) {
    om_heat_element(m, L, e, mode, p, node_dof_top, bc0, bc1, x_old_top, k_elem ? k_elem[e] : 1.,
                    vel_top, KeM, FeM, dofs, node_dof_L);
    for (int a = 0; a < n2; a++) {
      for (int b = 0; b < n2; b++) {
        int64_t q = find_pos(rowptr, colind, dofs[a], dofs[b]);
        if (q < 0) return -1;
        val[q] += KeM[a * n2 + b];
      }
      if (mode == 1) rhs[dofs[a]] += FeM[a];
    }
  }
  return 0;
}

int om_ns_genmatrhs(const om_mesh *m, int L, int mode, const om_ns_params *p,
                    const int *node_dof_L, const int *node_dof_top,
                    const int *bc0, const int *bc1, const double *x_old_top,
                    const double *nu_elem,
                    const int64_t *rowptr, const int *colind, double *val, double *rhs) {
  int nl = m->dim * m->n2 + m->n1;
  int n = m->dim * m->nq2[L] + m->nq1[L];
  double *KeM = (double *)malloc(sizeof(double) * nl * nl), FeM[89];
  int dofs[89];
  memset(val, 0, sizeof(double) * rowptr[n]);
  if (mode == 1) memset(rhs, 0, sizeof(double) * n);
  for (int e = 0; e < m->nel[L]; e++) {
    om_ns_element(m, L, e, mode, p, node_dof_top, bc0, bc1, x_old_top, nu_elem ? nu_elem[e] : p->IRe,
                  KeM, FeM, dofs, node_dof_L);
    for (int a = 0; a < nl; a++) {
      for (int b = 0; b < nl; b++) {
        int64_t q = find_pos(rowptr, colind, dofs[a], dofs[b]);
        if (q < 0) { free(KeM); return -1; }
        val[q] += KeM[a * nl + b];
      }
      if (mode == 1) rhs[dofs[a]] += FeM[a];
    }
  }
  free(KeM);
  return 0;
}
