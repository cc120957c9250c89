// assemble.hip -- GenMatRhs on gfx950: heat (Q2) and coupled Navier-Stokes
// (Q2-Q1 Taylor-Hood) element matrices and their sum into the level matrix.
//
// Two routes share the element kernels.  The element-buffer route (the default)
// computes every element of a level in one launch and writes its matrix and
// right-hand side to a buffer (E, Fe); the boundary-face kernel adds the face
// terms to the same element blocks; k_egather then sums the buffer into the
// SELL slots, one block per group of rows staged in LDS, in element order.  The
// coloured route adds each element matrix straight into the SELL values through
// the element -> slot map, one launch per element colour (elements of one
// colour share no node, so no atomics are needed); it serves the levels whose
// element buffer would not fit.
//
// Heat volume: one wavefront per element, reference tables staged once per
// block (k_heat_el).  NS volume: one 256-thread workgroup per element (k_ns_vol):
// Jacobian, inverse, JxW and global gradients (MGFE::Jac3D/Jac2D,
// get_dphi_gl_g: src/MGFE.C:1066-1166, 1496-1519) in LDS, element-matrix tiles
// in registers.  Boundary faces (penalty Dirichlet, Robin/Neumann): one thread
// per element (and velocity component).
//
// Heat:  template_appl/nse/SRC/MGSolverT.C:228-343 (volume), :366-507 (faces)
// NS:    template_appl/nse/SRC/MGSolverNS.C:285-495 (volume), :503-690 (faces)
#include <hip/hip_runtime.h>
#include <cstdint>
#include "fm_kernels.h"
#include "fm_device.h"

__constant__ int c_face2d[4][3] = {{0,1,4},{1,2,5},{2,3,6},{3,0,7}};
__constant__ int c_face3d[6][9] = {
  {0,1,2,3,8,9,10,11,20},{0,1,5,4,8,13,16,12,21},{1,2,6,5,9,14,17,13,22},
  {2,3,7,6,10,15,18,14,23},{3,0,4,7,11,12,19,15,24},{4,5,6,7,16,17,18,19,25}};

template <int DIM> __device__ inline int face_node(int s, int k) { return DIM == 2 ? c_face2d[s][k] : c_face3d[s][k]; }

// Jacobian at Gauss point qp: returns det, fills IJ (MGFE::Jac2D / Jac3D)
template <int DIM, int N2, int NG>
__device__ inline double jac_qp(int qp, const double *xx, const double *dref, double *IJ) {
  const int off = N2 * NG;
  if (DIM == 2) {
    double x_xi = 0., x_eta = 0., y_xi = 0., y_eta = 0.;
#pragma unroll 3
    for (int s = 0; s < N2; s++) {
      int sng = s * NG + qp;
      double a = dref[sng], b = dref[sng + off];
      x_xi += xx[s] * a; x_eta += xx[s] * b; y_xi += xx[s + N2] * a; y_eta += xx[s + N2] * b;
    }
    double det = x_xi * y_eta - y_xi * x_eta, idet = 1. / det;
    IJ[0] = y_eta * idet; IJ[1] = -y_xi * idet; IJ[2] = -x_eta * idet; IJ[3] = x_xi * idet;
    return det;
  } else {
    double x_xi = 0., x_eta = 0., x_zeta = 0., y_xi = 0., y_eta = 0., y_zeta = 0., z_xi = 0., z_eta = 0., z_zeta = 0.;
#pragma unroll 3
    for (int s = 0; s < N2; s++) {
      int sng = s * NG + qp;
      double a = dref[sng], b = dref[sng + off], c = dref[sng + 2 * off];
      x_xi += xx[s] * a; x_eta += xx[s] * b; x_zeta += xx[s] * c;
      y_xi += xx[s + N2] * a; y_eta += xx[s + N2] * b; y_zeta += xx[s + N2] * c;
      z_xi += xx[s + 2 * N2] * a; z_eta += xx[s + 2 * N2] * b; z_zeta += xx[s + 2 * N2] * c;
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
}

// surface Jacobian determinant on a face (MGFE::JacSur2D / JacSur3D); xb[k + d*NB]
template <int DIM, int NB, int NGB>
__device__ inline double jac_sur(int qp, const double *xb, const FeTab *tab) {
  if (DIM == 2) {
    double dx = 0., dy = 0.;
    for (int s = 0; s < NB; s++) { double d = tab->bdphi2[s * NGB + qp]; dx += xb[s] * d; dy += xb[s + NB] * d; }
    return sqrt(dx * dx + dy * dy);
  } else {
    const int off = NB * NGB;
    double dxdxi = 0., dxdeta = 0., dydxi = 0., dydeta = 0., dzdxi = 0., dzdeta = 0.;
    for (int s = 0; s < NB; s++) {
      int sng = s * NGB + qp;
      double a = tab->bdphi2[sng], b = tab->bdphi2[sng + off];
      dxdxi += xb[s] * a; dxdeta += xb[s] * b; dydxi += xb[s + NB] * a; dydeta += xb[s + NB] * b;
      dzdxi += xb[s + 2 * NB] * a; dzdeta += xb[s + 2 * NB] * b;
    }
    double c1 = dxdxi * dydeta - dxdeta * dydxi, c2 = dydxi * dzdeta - dydeta * dzdxi, c3 = dzdxi * dxdeta - dzdeta * dxdxi;
    return sqrt(c1 * c1 + c2 * c2 + c3 * c3);
  }
}

// SELL slot of entry (a, b) of element e's matrix: position elpos[e][a][b] within the row of local DoF a, whose
// position 0 is slot elslot[e][a]; -1 where the pattern has no such entry (0xFFFF) or the DoF does not exist
__device__ __forceinline__ int64_t el_slot(const int64_t *__restrict__ elslot, const unsigned short *__restrict__ elpos,
                                           int64_t e, int nl, int a, int b) {
  const unsigned k = elpos[(e * nl + a) * nl + b];
  const int64_t r = elslot[e * nl + a];
  return (k == 0xFFFFu || r < 0) ? -1 : r + 64 * (int64_t)k;
}

// ===========================================================================
// Heat volume kernel, one wavefront per element
//
// A 256-thread block stages the reference tables once for its four waves (dphi/dxi at the Gauss points,
// R[q][k][s], and phi[q][s]); afterwards every wave takes its own elements and never waits for the others.
// Per element:
//  1. the DIM x DIM Jacobian entries at the NG Gauss points (MGFE::Jac3D / Jac2D, src/MGFE.C:1066-1166) as
//     NG DIM^2 dot products over the element's nodes, spread over the 64 lanes;
//  2. per Gauss point (one lane each): the inverse Jacobian M_dk = dxi_k/dx_d, JxW, the interpolated velocity
//     and T_old, the upwind coefficients, and from them the weights of MGSolverT.C:228-330 folded into the
//     reference frame: the symmetric C_q,kl = sum_d JxW (alpha0 k_e + f_up,d v_d^2) M_dk M_dl and the
//     advection vector vt_q,k = rhocp sum_d v_d M_dk;
//  3. the element matrix K_ij = sum_q [ sum_kl C_q,kl R_qki R_qlj + a_qi b_qj ] with a_qi = JxW phi_i and
//     b_qj = tcdt phi_j + sum_k vt_q,k R_qkj (the mass / advection part only when present), each lane owning a
//     TI x TJ tile (rows i + 9u, columns TJ t + v: 63 tiles of 3 x 4 of the 27 x 27 hex27 matrix, 27 tiles of
//     1 x 3 of the 9 x 9 quad9 one);
//  4. Dirichlet rows dropped (the bc_vol row mask); the tile either added to A through the element -> SELL
//     slot map (coloured launches: the current values are read before step 1, since elements of one colour
//     share no slot) or written to the element-matrix buffer E[e][row][col] (TOE; row / column swapped for
//     the transposed matrix), and the element right-hand side likewise.
// ===========================================================================
__device__ __forceinline__ void wave_lds_sync() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); __builtin_amdgcn_wave_barrier(); }

template <int DIM, bool ADV, bool TOE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_heat_el(int nel, const int *__restrict__ elems, const int *__restrict__ conn,
                                                 const int *__restrict__ eldof, const unsigned short *__restrict__ elpos,
                                                 const int64_t *__restrict__ elslot, const double *__restrict__ xyz, int Ntop,
                                                 const int *__restrict__ dof_top, const int *__restrict__ bc0,
                                                 const double *__restrict__ x_old, const double *__restrict__ coef,
                                                 const double *__restrict__ vel, const FeTab *__restrict__ tab, HeatArgs a,
                                                 int mode, int trans, double *__restrict__ val, double *__restrict__ rhs,
                                                 double *__restrict__ E, double *__restrict__ Fe) {
  constexpr int N2 = DIM == 2 ? 9 : 27, NG = N2, DD = DIM * DIM, NC = DIM == 2 ? 4 : 6;
  // node-indexed LDS rows padded to an even length (zeros beyond N2): 16-byte loads of node pairs
  constexpr int NP = N2 + 1;
  // lane tiles: rows TI ti .. TI ti + TI - 1, columns TJ tj .. TJ tj + TJ - 1 (j < N2)
  constexpr int TI = DIM == 2 ? 1 : 3, TJ = DIM == 2 ? 3 : 4, NTI = N2 / TI, NTJ = (N2 + TJ - 1) / TJ;
  constexpr int NW = 4;  // waves per block
  __shared__ __align__(16) double s_R[NG * DIM * NP];  // [q][k][s]
  __shared__ __align__(16) double s_phi[NG * NP];      // [q][s]
  __shared__ __align__(16) double s_xx[NW][DIM * NP];
  __shared__ __align__(16) double s_vn[NW][ADV ? DIM * NP : 2];
  __shared__ __align__(16) double s_Tn[NW][NP];
  __shared__ __align__(16) double s_jb[NW][NG * DD];
  __shared__ __align__(16) double s_C[NW][NG * NC];
  __shared__ __align__(16) double s_vt[NW][ADV ? NG * DIM : 2];
  __shared__ double s_JxW[NW][NG], s_src[NW][NG];
  __shared__ int s_bcv[NW][N2];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  for (int t = tid; t < NG * DIM * NP; t += blockDim.x) {
    const int q = t / (DIM * NP), k = (t / NP) % DIM, sn = t % NP;
    s_R[t] = sn < N2 ? tab->dphi2[sn * NG + q + k * N2 * NG] : 0.;
  }
  for (int t = tid; t < NG * NP; t += blockDim.x) s_phi[t] = t % NP < N2 ? tab->phi2[(t % NP) * NG + t / NP] : 0.;
  for (int t = tid; t < NW * DIM; t += blockDim.x) {
    s_xx[t / DIM][(t % DIM) * NP + N2] = 0.;
    if (ADV) s_vn[t / DIM][(t % DIM) * NP + N2] = 0.;
  }
  if (tid < NW) s_Tn[tid][N2] = 0.;
  __syncthreads();
  const double tcdt = a.tc * a.rhocp / a.dt;
  double *xx = s_xx[wv], *vn = s_vn[wv], *Tn = s_Tn[wv], *jb = s_jb[wv], *Cq = s_C[wv], *vt = s_vt[wv];
  double *JxWq = s_JxW[wv], *srcq = s_src[wv];
  int *bcv = s_bcv[wv];
  const int ti = lane / NTJ, tj = lane % NTJ;
  const bool tile = lane < NTI * NTJ;
  const int nwaves = gridDim.x * NW;
  // sum_s a[s] b[s] over the padded node rows, two nodes per 16-byte load
  auto dotp = [](const double *x, const double *y) {
    const double2 *x2 = (const double2 *)x, *y2 = (const double2 *)y;
    double t0 = 0., t1 = 0.;
#pragma unroll 7
    for (int p = 0; p < NP / 2; p++) {
      const double2 u = x2[p], w = y2[p];
      t0 += u.x * w.x;
      t1 += u.y * w.y;
    }
    return t0 + t1;
  };
  for (int ei = xcd_block() * NW + wv; ei < nel; ei += nwaves) {
    const int e = elems ? elems[ei] : ei;
    // slots of this lane's entries and their current values (scatter into A)
    int64_t pp[TI][TJ];
    double pv[TI][TJ];
    if (!TOE && tile) {
#pragma unroll
      for (int u = 0; u < TI; u++)
#pragma unroll
        for (int v = 0; v < TJ; v++) {
          const int i = TI * ti + u, j = TJ * tj + v;
          pp[u][v] = j < N2 ? el_slot(elslot, elpos, e, N2, trans ? j : i, trans ? i : j) : -1;
          pv[u][v] = pp[u][v] >= 0 ? val[pp[u][v]] : 0.;
        }
    }
    if (lane < N2) {
      const int node = conn[(int64_t)e * N2 + lane];
      const int kd = dof_top[node];
      for (int d = 0; d < DIM; d++) {
        xx[lane + d * NP] = xyz[node + (int64_t)d * Ntop];
        if (ADV) vn[lane + d * NP] = vel ? vel[node + (int64_t)d * Ntop] : a.vel_const[d];
      }
      Tn[lane] = x_old[kd];
      bcv[lane] = bc0[kd];
    }
    wave_lds_sync();
    // 1. Jacobian entries J[q][c][k] = sum_s x_c,s dphi_s/dxi_k (q)
#pragma unroll 1
    for (int p = lane; p < NG * DD; p += 64) {
      const int q = p / DD, c = (p / DIM) % DIM, k = p % DIM;
      jb[p] = dotp(xx + c * NP, s_R + (q * DIM + k) * NP);
    }
    wave_lds_sync();
    // 2. per Gauss point
    if (lane < NG) {
      const int q = lane;
      const double *J = jb + q * DD;
      double M[DIM][DIM], det;
      if (DIM == 2) {
        det = J[0] * J[3] - J[2] * J[1];
        const double id = 1. / det;
        M[0][0] = J[3] * id; M[0][1] = -J[2] * id; M[1][0] = -J[1] * id; M[1][1] = J[0] * id;
      } else {
        // J[c*3+k] = dx_c/dxi_k; M[d][k] = (J^-1)[k][d]
        const double c00 = J[4] * J[8] - J[5] * J[7], c01 = J[5] * J[6] - J[3] * J[8], c02 = J[3] * J[7] - J[4] * J[6];
        det = J[0] * c00 + J[1] * c01 + J[2] * c02;
        const double id = 1. / det;
        M[0][0] = c00 * id; M[0][1] = c01 * id; M[0][2] = c02 * id;
        M[1][0] = (J[2] * J[7] - J[1] * J[8]) * id; M[1][1] = (J[0] * J[8] - J[2] * J[6]) * id;
        M[1][2] = (J[1] * J[6] - J[0] * J[7]) * id;
        M[2][0] = (J[1] * J[5] - J[2] * J[4]) * id; M[2][1] = (J[2] * J[3] - J[0] * J[5]) * id;
        M[2][2] = (J[0] * J[4] - J[1] * J[3]) * id;
      }
      const double JxW = det * tab->w[q];
      const double *ph = s_phi + q * NP;
      const double Tg = dotp(ph, Tn);
      double w[DIM];
      const double ke = coef ? coef[e] : 1.0;
      for (int d = 0; d < DIM; d++) w[d] = JxW * a.alpha0 * ke;
      if (ADV) {
        double vg[DIM];
        for (int d = 0; d < DIM; d++) vg[d] = dotp(ph, vn + d * NP);
        double mod2 = 1.e-20;
        for (int d = 0; d < DIM; d++) mod2 += vg[d] * vg[d];
        mod2 = sqrt(mod2);
        for (int d = 0; d < DIM; d++) {
          const double h = fabs(xx[d * NP + 2 + 4 * (DIM - 2)] - xx[d * NP + 0]);
          const double h1 = fabs(xx[d * NP + 3 + 4 * (DIM - 2)] - xx[d * NP + 1]);
          const double h_eff = h < h1 ? h1 : h;
          const double Pe = 0.5 * mod2 * h_eff / a.alpha0;
          const double fup = a.upwind * a.rhocp * 0.5 * (1. / tanh(Pe) - 1. / Pe) * h_eff / mod2;
          w[d] = JxW * (a.alpha0 * ke + fup * vg[d] * vg[d]);
        }
        for (int k = 0; k < DIM; k++) {
          double t = 0.;
          for (int d = 0; d < DIM; d++) t += vg[d] * M[d][k];
          vt[q * DIM + k] = a.rhocp * t;
        }
      }
      // C_q in (00, 01, 02, 11, 12, 22) order (2D: 00, 01, 11 and a zero pad)
      int c = 0;
      for (int k = 0; k < DIM; k++)
        for (int l = k; l < DIM; l++) {
          double t = 0.;
          for (int d = 0; d < DIM; d++) t += w[d] * M[d][k] * M[d][l];
          Cq[q * NC + c++] = t;
        }
      if (DIM == 2) Cq[q * NC + 3] = 0.;
      JxWq[q] = JxW;
      srcq[q] = JxW * (tcdt * Tg + a.source);
    }
    wave_lds_sync();
    // 3. element matrix tile
    if (tile) {
      double K[TI][TJ];
#pragma unroll
      for (int u = 0; u < TI; u++)
#pragma unroll
        for (int v = 0; v < TJ; v++) K[u][v] = 0.;
#pragma unroll 1
      for (int q = 0; q < NG; q++) {
        const double *R = s_R + q * DIM * NP;
        double C[DIM][DIM];
        {
          const double2 *c2 = (const double2 *)(Cq + q * NC);
          double cc[NC];
#pragma unroll
          for (int t = 0; t < NC / 2; t++) { const double2 u = c2[t]; cc[2 * t] = u.x; cc[2 * t + 1] = u.y; }
          int c = 0;
#pragma unroll
          for (int k = 0; k < DIM; k++)
#pragma unroll
            for (int l = k; l < DIM; l++) { C[k][l] = cc[c]; C[l][k] = cc[c]; c++; }
        }
        double rc[TJ][DIM];
#pragma unroll
        for (int k = 0; k < DIM; k++) {
          const double *Rk = R + k * NP + TJ * tj;
          if (TJ == 4) {
            const double2 u0 = ((const double2 *)Rk)[0], u1 = ((const double2 *)Rk)[1];
            rc[0][k] = u0.x; rc[1][k] = u0.y; rc[2 % TJ][k] = u1.x; rc[3 % TJ][k] = u1.y;
          } else {
#pragma unroll
            for (int v = 0; v < TJ; v++) rc[v][k] = Rk[v];
          }
        }
#pragma unroll
        for (int u = 0; u < TI; u++) {
          const int i = TI * ti + u;
          double H[DIM];
#pragma unroll
          for (int l = 0; l < DIM; l++) H[l] = 0.;
#pragma unroll
          for (int k = 0; k < DIM; k++) {
            const double r = R[k * NP + i];
#pragma unroll
            for (int l = 0; l < DIM; l++) H[l] += C[k][l] * r;
          }
#pragma unroll
          for (int v = 0; v < TJ; v++)
#pragma unroll
            for (int l = 0; l < DIM; l++) K[u][v] += H[l] * rc[v][l];
        }
        if (ADV) {
          const double *ph = s_phi + q * NP;
          double b[TJ];
#pragma unroll
          for (int v = 0; v < TJ; v++) {
            double t = tcdt * ph[TJ * tj + v];
#pragma unroll
            for (int k = 0; k < DIM; k++) t += vt[q * DIM + k] * rc[v][k];
            b[v] = t;
          }
          const double jw = JxWq[q];
#pragma unroll
          for (int u = 0; u < TI; u++) {
            const double ai = jw * ph[TI * ti + u];
#pragma unroll
            for (int v = 0; v < TJ; v++) K[u][v] += ai * b[v];
          }
        }
      }
      // 4. output
#pragma unroll
      for (int u = 0; u < TI; u++) {
        const int i = TI * ti + u;
        const bool keep = bcv[i] != 0;
#pragma unroll
        for (int v = 0; v < TJ; v++) {
          const int j = TJ * tj + v;
          if (j >= N2) continue;
          if (TOE) {
            E[((int64_t)e * N2 + (trans ? j : i)) * N2 + (trans ? i : j)] = keep ? K[u][v] : 0.;
          } else if (keep && pp[u][v] >= 0) {
            val[pp[u][v]] = pv[u][v] + K[u][v];
          }
        }
      }
    }
    if (mode == 1 && lane < N2) {
      const int i = lane;
      double F = 0.;
#pragma unroll 9
      for (int q = 0; q < NG; q++) F += s_phi[q * NP + i] * srcq[q];
      if (TOE) Fe[(int64_t)e * N2 + i] = bcv[i] != 0 ? F : 0.;
      else if (bcv[i] != 0) rhs[eldof[(int64_t)e * N2 + i]] += F;
    }
    wave_lds_sync();
  }
}

// Heat boundary faces: one thread per element of the colour (only elements
// with boundary sides do work).
// E != nullptr: the element-buffer assembly -- the face terms are added to the element's own block of E and Fe
// (every element of the level in one launch, elems == nullptr)
template <int DIM>
__global__ void k_heat_bd(int nel, const int *__restrict__ elems, const int *__restrict__ conn, const int *__restrict__ eldof,
                          const unsigned short *__restrict__ elpos, const int64_t *__restrict__ elslot, const unsigned char *__restrict__ bmask,
                          const double *__restrict__ xyz, int Ntop, const int *__restrict__ dof_top,
                          const int *__restrict__ bc0, const int *__restrict__ bc1, const double *__restrict__ x_old,
                          const FeTab *__restrict__ tab, HeatArgs a, int mode, int trans, double *__restrict__ val,
                          double *__restrict__ rhs, double *__restrict__ E, double *__restrict__ Fe) {
  constexpr int N2 = DIM == 2 ? 9 : 27, NB = DIM == 2 ? 3 : 9, NGB = NB, NS = 2 * DIM;
  int ei = blockIdx.x * blockDim.x + threadIdx.x;
  if (ei >= nel) return;
  int e = elems ? elems[ei] : ei;
  auto add = [&](int row, int col, double x) {
    if (E) { E[((int64_t)e * N2 + row) * N2 + col] += x; return; }
    const int64_t p = el_slot(elslot, elpos, e, N2, row, col);
    if (val && p >= 0) val[p] += x;
  };
  unsigned char bm = bmask[e];
  if (!bm) return;
  for (int s = 0; s < NS; s++) {
    if (!((bm >> s) & 1)) continue;
    int sur[NB];
    double xb[DIM * NB];
    for (int k = 0; k < NB; k++) {
      sur[k] = face_node<DIM>(s, k);
      int node = conn[(int64_t)e * N2 + sur[k]];
      for (int d = 0; d < DIM; d++) xb[k + d * NB] = xyz[node + (int64_t)d * Ntop];
    }
    int cnode = conn[(int64_t)e * N2 + sur[NB - 1]];
    int ckd = dof_top[cnode];
    if (bc0[ckd] == 0) {
      int bc_s = bc1[ckd];
      double Ipen = jac_sur<DIM, NB, NGB>(NGB - 1, xb, tab) / a.dt;
      for (int k = 0; k < NB; k++) {
        int lv = sur[k];
        if (val || E) add(lv, lv, Ipen);
        if (mode == 1) {
          int node = conn[(int64_t)e * N2 + lv];
          const double f = (bc_s % 2) * Ipen * x_old[dof_top[node]];
          if (Fe) Fe[(int64_t)e * N2 + lv] += f;
          else rhs[eldof[(int64_t)e * N2 + lv]] += f;
        }
      }
    } else if (bc1[ckd] != 0) {
      for (int qp = 0; qp < NGB; qp++) {
        double JxW = jac_sur<DIM, NB, NGB>(qp, xb, tab) * tab->bw[qp];
        for (int i = 0; i < NB; i++) {
          int lei = sur[i];
          int kd = dof_top[conn[(int64_t)e * N2 + lei]];
          int bcs = bc1[kd], bcv = bc0[kd];
          double dtxJxW = JxW * bcv;
          int bc_alpha = (bcs & 2) >> 1;
          if (bc_alpha == 0) continue;
          double phii = tab->bphi2[i * NGB + qp];
          for (int j = 0; j < NB; j++)
            add(trans ? sur[j] : lei, trans ? lei : sur[j], dtxJxW * bc_alpha * a.alpha0 * phii * tab->bphi2[j * NGB + qp]);
        }
      }
    }
  }
}

// ===========================================================================
// Navier-Stokes (Q2-Q1 coupled) volume kernel
// ===========================================================================
// One 256-thread workgroup per element.  After the Gauss-point phase (as in
// the heat kernel) the global gradients G[q][d][i] sit in LDS and every block
// of the element matrix is a short sum over the Gauss points of products of
// G rows:
//   (iv,iv): sum_q [ sum_k JxW cl_k G_ki G_kj + JxW cl_iv G_iv,i G_iv,j
//                    + JxW phi_i (tcdt phi_j + u_old.grad phi_j) ]
//   (iv,jv): sum_q JxW (nu + f_up,iv u_iv u_jv) G_jv,i G_iv,j      (iv != jv)
//   B^T:     -sum_q JxW psi_j G_iv,i        B: sum_q JxW psi_i G_jv,j
// (MGSolverNS.C:364-489).  With a.newton the velocity blocks also take the
// derivatives with respect to the advecting field (u = x_old) of the
// advection term and of the upwind coefficients f_d(|u|) u_d^2, f_iv u_iv u_jv:
//   (iv,m): sum_q JxW [ phi_i phi_s du_iv/dx_m + (sum_k G_k,i B_k^{iv,m}) phi_s ],
//   B_k^{iv,m} = D_km du_iv/dx_k + E^{iv}_km du_k/dx_iv,
//   D_km = f_k' (u_m/|u|) u_k^2 + 2 f_k u_k [k = m],
//   E^{iv}_km = f_iv' (u_m/|u|) u_iv u_k + f_iv (u_k [m = iv] + u_iv [m = k]),
// f_d = A_d L(c_d |u|) / |u|, L(P) = coth P - 1/P, so f_d' = A_d (c_d L'(P) |u| - L(P)) / |u|^2,
// L'(P) = 1/P^2 - 1/sinh^2 P.  This makes the matrix the Newton Jacobian of the
// steady residual A(u) u - b at u = x_old.  The DIM^2 velocity blocks are split into 3x3
// register tiles (1x1 in 2D), 81 per block; B^T / B entries are one per
// thread.  Velocity rows of Dirichlet DoFs are dropped (bc_vol), pressure
// rows are not.
// TOE: every entry of the element matrix (zeros included) written to E[e][row][col] (row / column swapped for
// the transposed matrix) and the element rhs to Fe[e][row], for all elements in one launch (elems == nullptr)
template <int DIM, bool TOE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_ns_vol(int nel, const int *__restrict__ elems, const int *__restrict__ conn,
                                                const int *__restrict__ eldof, const unsigned short *__restrict__ elpos, const int64_t *__restrict__ elslot,
                                                const double *__restrict__ xyz, int Ntop, const int *__restrict__ dof_top,
                                                const int *__restrict__ bc0, const double *__restrict__ x_old,
                                                const double *__restrict__ coef, const FeTab *__restrict__ tab, NsArgs a,
                                                int mode, int trans, double *__restrict__ val, double *__restrict__ rhs,
                                                double *__restrict__ E, double *__restrict__ Fe) {
  constexpr int N2 = DIM == 2 ? 9 : 27, N1 = DIM == 2 ? 4 : 8, NG = N2, NL = DIM * N2 + N1, TI = DIM == 2 ? 1 : 3;
  __shared__ double s_G[NG * DIM * N2];  // [q][d][i]
  __shared__ double s_phi[N2 * NG], s_psi[N1 * NG];
  __shared__ double s_xx[DIM * N2], s_uold[DIM * N2];
  __shared__ double s_JxW[NG], s_vel[NG * DIM], s_fup[NG * DIM], s_cl[NG * DIM], s_IJ[NG * DIM * DIM];
  __shared__ double s_Tn[N2], s_Tq[NG];
  __shared__ double s_du[NG * DIM * DIM];   // [q][iv][jv] = du_iv/dx_jv (Newton term only)
  __shared__ double s_fupd[NG * DIM];        // [q][d] = d f_d / d|u| (Newton term with upwinding)
  __shared__ double s_B[NG * DIM * DIM * DIM];  // [q][iv][m][k] (Newton term with upwinding)
  __shared__ int s_bcv[DIM * N2];
  const int tid = threadIdx.x;
  const double tcdt = a.tc / a.dt;
  for (int t = tid; t < N2 * NG; t += blockDim.x) s_phi[t] = tab->phi2[t];
  for (int t = tid; t < N1 * NG; t += blockDim.x) s_psi[t] = tab->phi1[t];
  for (int ei = xcd_block(); ei < nel; ei += gridDim.x) {
    const int e = elems ? elems[ei] : ei;
    // entry (row, col) of the element matrix: E (zero in Dirichlet rows) or added to A through the slot map
    auto put = [&](int row, int col, double x, bool keep) {
      if (TOE) {
        E[((int64_t)e * NL + (trans ? col : row)) * NL + (trans ? row : col)] = keep ? x : 0.;
      } else if (keep) {
        const int64_t p = el_slot(elslot, elpos, e, NL, trans ? col : row, trans ? row : col);
        if (p >= 0) val[p] += x;
      }
    };
    __syncthreads();
    if (tid < N2) {
      int node = conn[(int64_t)e * N2 + tid];
      for (int d = 0; d < DIM; d++) {
        s_xx[tid + d * N2] = xyz[node + (int64_t)d * Ntop];
        int kd = dof_top[node + d * Ntop];
        s_uold[tid + d * N2] = x_old[kd];
        s_bcv[tid + d * N2] = bc0[kd];
      }
      s_Tn[tid] = a.Tn ? a.Tn[node] : 0.0;
    }
    __syncthreads();
    const double nu = coef ? coef[e] : a.IRe;
    if (tid < NG) {
      const int qp = tid;
      double IJ[DIM * DIM];
      const double det = jac_qp<DIM, N2, NG>(qp, s_xx, tab->dphi2, IJ);
      s_JxW[qp] = det * tab->w[qp];
      double vg[DIM], Tq = 0.;
      for (int d = 0; d < DIM; d++) vg[d] = 0.;
#pragma unroll 3
      for (int sn = 0; sn < N2; sn++) {
        const double ph = s_phi[sn * NG + qp];
        for (int d = 0; d < DIM; d++) vg[d] += ph * s_uold[sn + d * N2];
        Tq += ph * s_Tn[sn];
      }
      s_Tq[qp] = Tq;
      double h_eff[DIM];
      for (int d = 0; d < DIM; d++) {
        double h = fabs(s_xx[d * N2 + 2 + 4 * (DIM - 2)] - s_xx[d * N2 + 0]);
        double h1 = fabs(s_xx[d * N2 + 3 + 4 * (DIM - 2)] - s_xx[d * N2 + 1]);
        h_eff[d] = h < h1 ? h1 : h;
      }
      double mod2 = 1.e-20;
      for (int d = 0; d < DIM; d++) mod2 += vg[d] * vg[d];
      mod2 = sqrt(mod2);
      for (int d = 0; d < DIM; d++) {
        const double Pe = 0.5 * mod2 * h_eff[d] / a.IRe;
        const double Lp = 1. / tanh(Pe) - 1. / Pe;
        const double fup = a.upwind * 0.5 * Lp * h_eff[d] / mod2;
        s_vel[qp * DIM + d] = vg[d];
        s_fup[qp * DIM + d] = fup;
        s_cl[qp * DIM + d] = nu + fup * vg[d] * vg[d];
        if (a.newton) {
          const double sh = sinh(Pe), dL = 1. / (Pe * Pe) - 1. / (sh * sh);
          const double Ad = a.upwind * 0.5 * h_eff[d], cd = 0.5 * h_eff[d] / a.IRe;
          s_fupd[qp * DIM + d] = Ad * (cd * dL * mod2 - Lp) / (mod2 * mod2);
        }
      }
      for (int t = 0; t < DIM * DIM; t++) s_IJ[qp * DIM * DIM + t] = IJ[t];
    }
    __syncthreads();
    for (int pq = tid; pq < NG * N2; pq += blockDim.x) {
      const int q = pq / N2, i = pq % N2;
      double gr[DIM];
      for (int d = 0; d < DIM; d++) gr[d] = tab->dphi2[i * NG + q + d * N2 * NG];
      for (int d = 0; d < DIM; d++) {
        double g = 0.;
        for (int j = 0; j < DIM; j++) g += s_IJ[q * DIM * DIM + j + d * DIM] * gr[j];
        s_G[(q * DIM + d) * N2 + i] = g;
      }
    }
    __syncthreads();
    if (a.newton) {
      for (int t = tid; t < NG * DIM * DIM; t += blockDim.x) {
        const int q = t / (DIM * DIM), iv = (t / DIM) % DIM, jv = t % DIM;
        double g = 0.;
        for (int sn = 0; sn < N2; sn++) g += s_G[(q * DIM + jv) * N2 + sn] * s_uold[sn + iv * N2];
        s_du[t] = g;
      }
      __syncthreads();
      if (a.upwind != 0.) {
        for (int t = tid; t < NG * DIM * DIM * DIM; t += blockDim.x) {
          const int q = t / (DIM * DIM * DIM), iv = (t / (DIM * DIM)) % DIM, m = (t / DIM) % DIM, k = t % DIM;
          const double *v = s_vel + q * DIM;
          double mod2 = 1.e-20;
          for (int d = 0; d < DIM; d++) mod2 += v[d] * v[d];
          mod2 = sqrt(mod2);
          const double um = v[m] / mod2;
          const double D = s_fupd[q * DIM + k] * um * v[k] * v[k] + (k == m ? 2. * s_fup[q * DIM + k] * v[k] : 0.);
          const double E = s_fupd[q * DIM + iv] * um * v[iv] * v[k] +
                           s_fup[q * DIM + iv] * ((m == iv ? v[k] : 0.) + (m == k ? v[iv] : 0.));
          s_B[t] = D * s_du[(q * DIM + iv) * DIM + k] + E * s_du[(q * DIM + k) * DIM + iv];
        }
        __syncthreads();
      }
    }
    // velocity blocks: DIM^2 blocks x 81 tiles
    for (int u = tid; u < DIM * DIM * 81; u += blockDim.x) {
      const int blk = u / 81, t = u % 81;
      const int iv = blk / DIM, jv = blk % DIM, ti = t / 9, tj = t % 9;
      double K[TI][TI];
#pragma unroll
      for (int x = 0; x < TI; x++)
#pragma unroll
        for (int y = 0; y < TI; y++) K[x][y] = 0.;
      if (iv == jv) {
#pragma unroll 1
        for (int q = 0; q < NG; q++) {
          const double jw = s_JxW[q];
          const double rc = tcdt + (a.newton ? s_du[(q * DIM + iv) * DIM + iv] : 0.);
          double bj[TI];
#pragma unroll
          for (int y = 0; y < TI; y++) bj[y] = rc * s_phi[(tj + 9 * y) * NG + q];
#pragma unroll
          for (int k = 0; k < DIM; k++) {
            const double *G = s_G + (q * DIM + k) * N2;
            const double w = jw * (s_cl[q * DIM + k] + (k == iv ? s_cl[q * DIM + iv] : 0.));
            const double vk = s_vel[q * DIM + k];
            double gi[TI], gj[TI];
#pragma unroll
            for (int x = 0; x < TI; x++) { gi[x] = w * G[ti + 9 * x]; gj[x] = G[tj + 9 * x]; }
#pragma unroll
            for (int y = 0; y < TI; y++) bj[y] += vk * gj[y];
#pragma unroll
            for (int x = 0; x < TI; x++)
#pragma unroll
              for (int y = 0; y < TI; y++) K[x][y] += gi[x] * gj[y];
          }
#pragma unroll
          for (int x = 0; x < TI; x++) {
            const double ai = jw * s_phi[(ti + 9 * x) * NG + q];
#pragma unroll
            for (int y = 0; y < TI; y++) K[x][y] += ai * bj[y];
          }
        }
      } else {
#pragma unroll 1
        for (int q = 0; q < NG; q++) {
          const double c = s_JxW[q] * (nu + s_fup[q * DIM + iv] * s_vel[q * DIM + iv] * s_vel[q * DIM + jv]);
          const double *Gi = s_G + (q * DIM + jv) * N2, *Gj = s_G + (q * DIM + iv) * N2;
          double gi[TI], gj[TI];
#pragma unroll
          for (int x = 0; x < TI; x++) { gi[x] = c * Gi[ti + 9 * x]; gj[x] = Gj[tj + 9 * x]; }
#pragma unroll
          for (int x = 0; x < TI; x++)
#pragma unroll
            for (int y = 0; y < TI; y++) K[x][y] += gi[x] * gj[y];
          if (a.newton) {
            const double c2 = s_JxW[q] * s_du[(q * DIM + iv) * DIM + jv];
#pragma unroll
            for (int x = 0; x < TI; x++) {
              const double ai = c2 * s_phi[(ti + 9 * x) * NG + q];
#pragma unroll
              for (int y = 0; y < TI; y++) K[x][y] += ai * s_phi[(tj + 9 * y) * NG + q];
            }
          }
        }
      }
      if (a.newton && a.upwind != 0.) {
        // derivative of the upwind coefficients: block (iv, m = jv)
#pragma unroll 1
        for (int q = 0; q < NG; q++) {
          double gi[TI];
#pragma unroll
          for (int x = 0; x < TI; x++) gi[x] = 0.;
#pragma unroll
          for (int k = 0; k < DIM; k++) {
            const double c = s_JxW[q] * s_B[((q * DIM + iv) * DIM + jv) * DIM + k];
            const double *G = s_G + (q * DIM + k) * N2;
#pragma unroll
            for (int x = 0; x < TI; x++) gi[x] += c * G[ti + 9 * x];
          }
#pragma unroll
          for (int x = 0; x < TI; x++)
#pragma unroll
            for (int y = 0; y < TI; y++) K[x][y] += gi[x] * s_phi[(tj + 9 * y) * NG + q];
        }
      }
#pragma unroll
      for (int x = 0; x < TI; x++) {
        const int row = iv * N2 + ti + 9 * x;
        const bool keep = s_bcv[row] != 0;
        if (!TOE && !keep) continue;
#pragma unroll
        for (int y = 0; y < TI; y++) put(row, jv * N2 + tj + 9 * y, K[x][y], keep);
      }
    }
    // B^T (velocity rows, pressure columns) and B (pressure rows, velocity columns): B = -(B^T)^T entry for entry
    // (the same products in the same order), so each pair is formed once; the B^T entry is dropped in Dirichlet
    // velocity rows, the B entry is always kept
    constexpr int NBT = DIM * N2 * N1;
    for (int u = tid; u < NBT; u += blockDim.x) {
      const int row = u / N1, j = u % N1, iv = row / N2, i = row % N2;
      double K = 0.;
#pragma unroll 3
      for (int q = 0; q < NG; q++) K -= s_JxW[q] * s_psi[j * NG + q] * s_G[(q * DIM + iv) * N2 + i];
      put(row, DIM * N2 + j, K, s_bcv[row] != 0);
      put(DIM * N2 + j, row, -K, true);
    }
    // the pressure-pressure block is zero (no stabilisation)
    if (TOE)
      for (int u = tid; u < N1 * N1; u += blockDim.x) put(DIM * N2 + u / N1, DIM * N2 + u % N1, 0., false);
    if (mode == 1 && tid < DIM * N2) {
      const int iv = tid / N2, i = tid % N2;
      const bool keep = s_bcv[i + iv * N2] != 0;
      double F = 0.;
      if (keep) {
        for (int qp = 0; qp < NG; qp++) {
          const double phii = s_phi[i * NG + qp];
          F += s_JxW[qp] * (tcdt * s_vel[qp * DIM + iv] * phii + a.IFr * a.dirg[iv] * phii + a.src[iv] * phii +
                            a.buoy[iv] * s_Tq[qp] * phii);
        }
      }
      if (TOE) Fe[(int64_t)e * NL + tid] = F;
      else if (keep) rhs[eldof[(int64_t)e * NL + tid]] += F;
    }
    if (TOE && mode == 1 && tid < N1) Fe[(int64_t)e * NL + DIM * N2 + tid] = 0.;
  }
}

// NS boundary faces (penalty Dirichlet with normal/tangential bits; Neumann tangential/normal terms)
template <int DIM>
__global__ void k_ns_bd(int nel, const int *__restrict__ elems, const int *__restrict__ conn, const int *__restrict__ eldof,
                        const unsigned short *__restrict__ elpos, const int64_t *__restrict__ elslot, const unsigned char *__restrict__ bmask, const double *__restrict__ xyz,
                        int Ntop, const int *__restrict__ dof_top, const int *__restrict__ bc0, const int *__restrict__ bc1,
                        const double *__restrict__ x_old, const FeTab *__restrict__ tab, NsArgs a, int mode, int trans,
                        double *__restrict__ val, double *__restrict__ rhs, double *__restrict__ E, double *__restrict__ Fe) {
  constexpr int N2 = DIM == 2 ? 9 : 27, N1 = DIM == 2 ? 4 : 8, NB = DIM == 2 ? 3 : 9, NGB = NB, NS = 2 * DIM;
  constexpr int NL = DIM * N2 + N1;
  // one thread per (element, velocity component): a thread writes only the rows of its own component
  // (entries (row_iv, col_jv) with jv >= iv, or their transposed positions), so the threads of one
  // element never touch the same entry
  const int tix = blockIdx.x * blockDim.x + threadIdx.x;
  const int ei = tix / DIM, iv = tix % DIM;
  if (ei >= nel) return;
  int e = elems ? elems[ei] : ei;
  unsigned char bm = bmask[e];
  if (!bm) return;
  auto add = [&](int row, int col, double x) {
    if (E) { E[((int64_t)e * NL + row) * NL + col] += x; return; }
    const int64_t p = el_slot(elslot, elpos, e, NL, row, col);
    if (p >= 0) val[p] += x;
  };
  for (int s = 0; s < NS; s++) {
    if (!((bm >> s) & 1)) continue;
    int sur[NB], gn[NB];
    double xb[DIM * NB], nrm[DIM];
    for (int k = 0; k < NB; k++) {
      sur[k] = face_node<DIM>(s, k);
      gn[k] = conn[(int64_t)e * N2 + sur[k]];
      for (int d = 0; d < DIM; d++) xb[k + d * NB] = xyz[gn[k] + (int64_t)d * Ntop];
    }
    // MGFE::normal_g
    if (DIM == 2) {
      nrm[0] = xb[1 + NB] - xb[NB]; nrm[1] = -(xb[1] - xb[0]);
    } else {
      double t1[3], t3[3];
      for (int i = 0; i < 3; i++) { t1[i] = xb[1 + i * NB] - xb[i * NB]; t3[i] = xb[3 + i * NB] - xb[i * NB]; }
      nrm[0] = t1[1] * t3[2] - t1[2] * t3[1];
      nrm[DIM > 1 ? 1 : 0] = t1[2] * t3[0] - t1[0] * t3[2];
      nrm[DIM > 2 ? 2 : 0] = t1[0] * t3[1] - t1[1] * t3[0];
    }
    double mm = 0.;
    for (int d = 0; d < DIM; d++) mm += nrm[d] * nrm[d];
    mm = sqrt(mm);
    for (int d = 0; d < DIM; d++) nrm[d] /= mm;
    {
      int ckd = dof_top[gn[NB - 1] + iv * Ntop];
      if (bc0[ckd] == 0) {
        double Ipen = jac_sur<DIM, NB, NGB>(NGB - 1, xb, tab) / a.dt;
        for (int i = 0; i < NB; i++) {
          int row = sur[i] + iv * N2;
          int kd = dof_top[gn[i] + iv * Ntop];
          int bcs = bc1[kd];
          int bc_rhs = (bcs & 4) >> 2, bc_tg = (bcs & 2) >> 1, bc_n = bcs % 2;
          double vo[DIM], vdotn = 0.;
          for (int k = 0; k < DIM; k++) { vo[k] = x_old[dof_top[gn[i] + k * Ntop]]; vdotn += nrm[k] * vo[k]; }
          if (mode == 1) {
            const double f = bc_rhs * Ipen * ((1 - bc_n) * (1 - bc_tg) * vo[iv] + bc_n * vdotn * nrm[iv] +
                                              bc_tg * (vo[iv] - vdotn * nrm[iv]));
            if (Fe) Fe[(int64_t)e * NL + row] += f;
            else rhs[eldof[(int64_t)e * NL + row]] += f;
          }
          if (!val && !E) continue;
          add(row, row, ((1 - bc_n) * (1 - bc_tg) + bc_tg) * Ipen + Ipen * nrm[iv] * nrm[iv] * (bc_n - bc_tg));
          for (int jv = iv + 1; jv < DIM; jv++) {
            const int col = sur[i] + jv * N2;
            add(trans ? col : row, trans ? row : col, Ipen * nrm[jv] * nrm[iv] * (bc_n - bc_tg));
          }
        }
      } else if (bc1[ckd] != 0 && (val || E)) {
        for (int qp = 0; qp < NGB; qp++) {
          double JxW = jac_sur<DIM, NB, NGB>(qp, xb, tab) * tab->bw[qp];
          for (int i = 0; i < NB; i++) {
            int row = sur[i] + iv * N2;
            int kd = dof_top[gn[i] + iv * Ntop];
            int bcs = bc1[kd], bcv = bc0[kd];
            double dtJxW = JxW * (bcv % 2);
            int bc_tg = (bcs & 2) >> 1, bc_n = bcs % 2;
            if (bc_tg == 0 && bc_n == 0) continue;
            double phii = tab->bphi2[i * NGB + qp];
            for (int j = 0; j < NB; j++) {
              double phij = tab->bphi2[j * NGB + qp];
              add(row, row, dtJxW * bc_tg * phij * phii + dtJxW * nrm[iv] * nrm[iv] * phij * phii * (bc_n - bc_tg));
              for (int jv = iv + 1; jv < DIM; jv++) {
                const int col = sur[j] + jv * N2;
                add(trans ? col : row, trans ? row : col, dtJxW * nrm[jv] * nrm[iv] * phij * phii * (bc_n - bc_tg));
              }
            }
          }
        }
      }
    }
  }
}

// ===========================================================================
// Heat functional / gradient kernel (top level, one workgroup per element):
//   u_e = Klap_e T_e with Klap_e(i,j) = sum_q JxW_q grad phi_i . grad phi_j
//   diss[e] = T_e . u_e ;  adj[e] = sum_i bc_vol_i lam_i u_i
//   y[dof_i] += ysc * coef_e * u_i   (colour by colour, when y != null)
// ===========================================================================
template <int DIM>
__global__ __launch_bounds__(64) void k_heat_lap(int nel, const int *__restrict__ elems, const int *__restrict__ conn,
                                                 const int *__restrict__ eldof, const double *__restrict__ xyz, int Ntop,
                                                 const int *__restrict__ dof_top, const int *__restrict__ bc0,
                                                 const double *__restrict__ T, const double *__restrict__ lam,
                                                 const double *__restrict__ coef, const FeTab *__restrict__ tab,
                                                 double *__restrict__ diss, double *__restrict__ adj,
                                                 double *__restrict__ y, double ysc) {
  constexpr int N2 = DIM == 2 ? 9 : 27, NG = N2;
  __shared__ double s_xx[DIM * N2], s_T[N2], s_lam[N2], s_gT[NG * DIM], s_JxW[NG];
  __shared__ double s_dg[NG * N2 * DIM];
  __shared__ int s_bcv[N2], s_dof[N2];
  __shared__ double s_red[2 * N2];
  const int tid = threadIdx.x;
  for (int ei = xcd_block(); ei < nel; ei += gridDim.x) {
    const int e = elems[ei];
    __syncthreads();
    if (tid < N2) {
      int node = conn[(int64_t)e * N2 + tid];
      int kd = dof_top[node];
      for (int d = 0; d < DIM; d++) s_xx[tid + d * N2] = xyz[node + (int64_t)d * Ntop];
      s_T[tid] = T[kd];
      s_lam[tid] = lam ? lam[kd] : 0.0;
      s_bcv[tid] = bc0[kd];
      s_dof[tid] = eldof[(int64_t)e * N2 + tid];
    }
    __syncthreads();
    if (tid < NG) {
      const int qp = tid;
      double IJ[DIM * DIM];
      double det = jac_qp<DIM, N2, NG>(qp, s_xx, tab->dphi2, IJ);
      s_JxW[qp] = det * tab->w[qp];
      double g[DIM];
      for (int d = 0; d < DIM; d++) g[d] = 0.;
      for (int sn = 0; sn < N2; sn++) {
        double gr[DIM];
        for (int d = 0; d < DIM; d++) gr[d] = tab->dphi2[sn * NG + qp + d * N2 * NG];
        for (int i = 0; i < DIM; i++) {
          double sum = 0.;
          for (int j = 0; j < DIM; j++) sum += IJ[j + i * DIM] * gr[j];
          s_dg[(qp * N2 + sn) * DIM + i] = sum;
          g[i] += sum * s_T[sn];
        }
      }
      for (int d = 0; d < DIM; d++) s_gT[qp * DIM + d] = g[d];
    }
    __syncthreads();
    if (tid < N2) {
      const int i = tid;
      double u = 0.;
      for (int qp = 0; qp < NG; qp++) {
        double t = 0.;
        for (int d = 0; d < DIM; d++) t += s_dg[(qp * N2 + i) * DIM + d] * s_gT[qp * DIM + d];
        u += s_JxW[qp] * t;
      }
      s_red[i] = s_T[i] * u;
      s_red[N2 + i] = s_bcv[i] * s_lam[i] * u;
      if (y) y[s_dof[i]] += ysc * (coef ? coef[e] : 1.0) * u;
    }
    __syncthreads();
    if (tid == 0) {
      double a0 = 0., a1 = 0.;
      for (int i = 0; i < N2; i++) { a0 += s_red[i]; a1 += s_red[N2 + i]; }
      if (diss) diss[e] = a0;
      if (adj) adj[e] = a1;
    }
  }
}

// ===========================================================================
// Navier-Stokes viscous functional / gradient kernel (top level):
//   w_(iv,i) = sum_q JxW_q sum_k d_k phi_i (d_k u_iv + d_iv u_k)   ( = dA/dnu_e u )
//   diss[e]  = sum_q JxW_q sum_{iv,k} (d_k u_iv)^2                 ( int_e |grad u|^2 )
//   adj[e]   = sum_(iv,i) bc_vol lam_(iv,i) w_(iv,i)
//   y[dof]  += ysc * nu_e * sum_q JxW_q d_k phi_i d_k u_iv          ( 2 nu K_lap u with ysc = 2 )
// ===========================================================================
template <int DIM>
__global__ __launch_bounds__(128) void k_ns_visc(int nel, const int *__restrict__ elems, const int *__restrict__ conn,
                                                const int *__restrict__ eldof, const double *__restrict__ xyz, int Ntop,
                                                const int *__restrict__ dof_top, const int *__restrict__ bc0,
                                                const double *__restrict__ U, const double *__restrict__ lam,
                                                const double *__restrict__ coef, double nu0, const FeTab *__restrict__ tab,
                                                double *__restrict__ diss, double *__restrict__ adj,
                                                double *__restrict__ y, double ysc) {
  constexpr int N2 = DIM == 2 ? 9 : 27, N1 = DIM == 2 ? 4 : 8, NG = N2, NL = DIM * N2 + N1;
  // one thread per velocity row (DIM * N2 <= 81) and per quadrature point: launched with 128 threads
  static_assert(DIM * N2 <= 128, "k_ns_visc needs a thread per velocity row");
  __shared__ double s_xx[DIM * N2], s_u[DIM * N2], s_lam[DIM * N2], s_JxW[NG];
  __shared__ double s_gu[NG * DIM * DIM];  // [qp][iv][k] = d_k u_iv
  __shared__ double s_dg[NG * N2 * DIM];
  __shared__ int s_bcv[DIM * N2], s_dof[DIM * N2];
  __shared__ double s_red[2 * DIM * N2 + NG];
  const int tid = threadIdx.x;
  for (int ei = xcd_block(); ei < nel; ei += gridDim.x) {
    const int e = elems[ei];
    __syncthreads();
    if (tid < N2) {
      int node = conn[(int64_t)e * N2 + tid];
      for (int d = 0; d < DIM; d++) {
        s_xx[tid + d * N2] = xyz[node + (int64_t)d * Ntop];
        int kd = dof_top[node + (int64_t)d * Ntop];
        s_u[tid + d * N2] = U[kd];
        s_lam[tid + d * N2] = lam ? lam[kd] : 0.0;
        s_bcv[tid + d * N2] = bc0[kd];
        s_dof[tid + d * N2] = eldof[(int64_t)e * NL + tid + d * N2];
      }
    }
    __syncthreads();
    if (tid < NG) {
      const int qp = tid;
      double IJ[DIM * DIM];
      double det = jac_qp<DIM, N2, NG>(qp, s_xx, tab->dphi2, IJ);
      s_JxW[qp] = det * tab->w[qp];
      double g[DIM * DIM];
      for (int t = 0; t < DIM * DIM; t++) g[t] = 0.;
      for (int sn = 0; sn < N2; sn++) {
        double gr[DIM];
        for (int d = 0; d < DIM; d++) gr[d] = tab->dphi2[sn * NG + qp + d * N2 * NG];
        for (int k = 0; k < DIM; k++) {
          double sum = 0.;
          for (int j = 0; j < DIM; j++) sum += IJ[j + k * DIM] * gr[j];
          s_dg[(qp * N2 + sn) * DIM + k] = sum;
          for (int iv = 0; iv < DIM; iv++) g[iv * DIM + k] += sum * s_u[sn + iv * N2];
        }
      }
      double q2 = 0.;
      for (int t = 0; t < DIM * DIM; t++) { s_gu[qp * DIM * DIM + t] = g[t]; q2 += g[t] * g[t]; }
      s_red[2 * DIM * N2 + qp] = s_JxW[qp] * q2;
    }
    __syncthreads();
    const double nu = coef ? coef[e] : nu0;
    if (tid < DIM * N2) {
      const int iv = tid / N2, i = tid % N2;
      double w = 0., lapu = 0.;
      for (int qp = 0; qp < NG; qp++) {
        double t = 0., l = 0.;
        for (int k = 0; k < DIM; k++) {
          const double dk = s_dg[(qp * N2 + i) * DIM + k];
          t += dk * (s_gu[qp * DIM * DIM + iv * DIM + k] + s_gu[qp * DIM * DIM + k * DIM + iv]);
          l += dk * s_gu[qp * DIM * DIM + iv * DIM + k];
        }
        w += s_JxW[qp] * t;
        lapu += s_JxW[qp] * l;
      }
      s_red[tid] = s_bcv[tid] * s_lam[tid] * w;
      if (y) y[s_dof[tid]] += ysc * nu * lapu;
    }
    __syncthreads();
    if (tid == 0) {
      double a0 = 0., a1 = 0.;
      for (int t = 0; t < DIM * N2; t++) a1 += s_red[t];
      for (int q = 0; q < NG; q++) a0 += s_red[2 * DIM * N2 + q];
      if (diss) diss[e] = a0;
      if (adj) adj[e] = a1;
    }
  }
}

// ===========================================================================
// Nusselt number of a differentially heated box (top level):
//   J = sum_e int_e (u_a T / kappa - k_e dT/dx_a)          (a = axis of the heated walls)
// the heat flux through the box along axis a in units of the conductive one,
// averaged over the box (at steady state the flux through every plane
// x_a = const is the same, so this is the hot-wall Nusselt number).
//   Jel[e] = J_e,  dJdk[e] = -int_e dT/dx_a,
//   y[dof_i] += sum_q JxW (u_a phi_i / kappa - k_e d phi_i / dx_a)      (dJ/dT, colour by colour)
// ===========================================================================
template <int DIM>
__global__ __launch_bounds__(64) void k_heat_nusselt(int nel, const int *__restrict__ elems, const int *__restrict__ conn,
                                                     const int *__restrict__ eldof, const double *__restrict__ xyz, int Ntop,
                                                     const int *__restrict__ dof_top, const double *__restrict__ vel,
                                                     const double *__restrict__ T, const double *__restrict__ coef,
                                                     double kappa, int axis, const FeTab *__restrict__ tab,
                                                     double *__restrict__ Jel, double *__restrict__ dJdk,
                                                     double *__restrict__ y) {
  constexpr int N2 = DIM == 2 ? 9 : 27, NG = N2;
  __shared__ double s_xx[DIM * N2], s_T[N2], s_u[N2];
  __shared__ double s_gx[NG * N2], s_JxW[NG], s_uq[NG], s_Tq[NG], s_dT[NG];
  __shared__ int s_dof[N2];
  const int tid = threadIdx.x;
  for (int ei = xcd_block(); ei < nel; ei += gridDim.x) {
    const int e = elems[ei];
    __syncthreads();
    if (tid < N2) {
      const int node = conn[(int64_t)e * N2 + tid];
      for (int d = 0; d < DIM; d++) s_xx[tid + d * N2] = xyz[node + (int64_t)d * Ntop];
      s_T[tid] = T[dof_top[node]];
      s_u[tid] = vel ? vel[node + (int64_t)axis * Ntop] : 0.0;
      s_dof[tid] = eldof[(int64_t)e * N2 + tid];
    }
    __syncthreads();
    const double ke = coef ? coef[e] : 1.0;
    if (tid < NG) {
      const int qp = tid;
      double IJ[DIM * DIM];
      const double det = jac_qp<DIM, N2, NG>(qp, s_xx, tab->dphi2, IJ);
      s_JxW[qp] = det * tab->w[qp];
      double Tq = 0., uq = 0., dT = 0.;
      for (int sn = 0; sn < N2; sn++) {
        const double ph = tab->phi2[sn * NG + qp];
        Tq += ph * s_T[sn];
        uq += ph * s_u[sn];
        double g = 0.;
        for (int j = 0; j < DIM; j++) g += IJ[j + axis * DIM] * tab->dphi2[sn * NG + qp + j * N2 * NG];
        s_gx[qp * N2 + sn] = g;
        dT += g * s_T[sn];
      }
      s_uq[qp] = uq;
      s_Tq[qp] = Tq;
      s_dT[qp] = dT;
    }
    __syncthreads();
    if (tid == 0) {
      double a0 = 0., a1 = 0.;
      for (int q = 0; q < NG; q++) {
        a0 += s_JxW[q] * (s_uq[q] * s_Tq[q] / kappa - ke * s_dT[q]);
        a1 -= s_JxW[q] * s_dT[q];
      }
      if (Jel) Jel[e] = a0;
      if (dJdk) dJdk[e] = a1;
    }
    if (y && tid < N2) {
      double acc = 0.;
      for (int q = 0; q < NG; q++)
        acc += s_JxW[q] * (s_uq[q] * tab->phi2[tid * NG + q] / kappa - ke * s_gx[q * N2 + tid]);
      y[s_dof[tid]] += acc;
    }
  }
}

// ===========================================================================
// Coupling terms of the Boussinesq system, on the heat system's top level,
// node-space outputs (colour by colour: elements of one colour share no node):
//   mode 0: out[j]            += scale sum_q JxW phi_j T_h                     dNu/du_axis (scale = 1/kappa)
//   mode 1: out[m * Ntop + j] += sum_q JxW phi_j (scale lam_h dT_h/dx_m + sum_d D_dm dlam_h/dx_d dT_h/dx_d)
//           = (dR_T/du)^T lam (scale = rho c_p; lam_h interpolating lam on the rows that carry the volume
//           terms, bc_vol != 0); D_dm = f_d' (u_m/|u|) u_d^2 + 2 f_d u_d [d = m] is the derivative of the upwind
//           diffusion f_d(|u|) u_d^2 of MGSolverT.C:290-300 (f_d = upwind rho c_p 0.5 L(Pe_d) h_d / |u|,
//           Pe_d = 0.5 |u| h_d / alpha0, L(P) = coth P - 1/P)
//   mode 2: out[j]            += scale sum_q JxW phi_j m_h                     M m (m nodal, Q2 mass matrix)
// ===========================================================================
template <int DIM>
__global__ __launch_bounds__(64) void k_coupling(int nel, const int *__restrict__ elems, const int *__restrict__ conn,
                                                 const double *__restrict__ xyz, int Ntop, const int *__restrict__ dof_top,
                                                 const int *__restrict__ bc0, const double *__restrict__ T,
                                                 const double *__restrict__ lam, const double *__restrict__ m, double scale,
                                                 int mode, const double *__restrict__ vel, double upwind, double rhocp,
                                                 double alpha0, const FeTab *__restrict__ tab, double *__restrict__ out) {
  constexpr int N2 = DIM == 2 ? 9 : 27, NG = N2;
  __shared__ double s_xx[DIM * N2], s_T[N2], s_l[N2], s_u[DIM * N2];
  __shared__ double s_JxW[NG], s_a[NG], s_g[NG * DIM], s_c[NG * DIM];
  __shared__ int s_node[N2];
  const int tid = threadIdx.x;
  for (int ei = xcd_block(); ei < nel; ei += gridDim.x) {
    const int e = elems[ei];
    __syncthreads();
    if (tid < N2) {
      const int node = conn[(int64_t)e * N2 + tid];
      s_node[tid] = node;
      for (int d = 0; d < DIM; d++) s_xx[tid + d * N2] = xyz[node + (int64_t)d * Ntop];
      const int kd = dof_top[node];
      s_T[tid] = mode == 2 ? m[node] : T[kd];
      s_l[tid] = mode == 1 ? (bc0[kd] != 0 ? lam[kd] : 0.0) : 0.0;
      for (int d = 0; d < DIM; d++) s_u[tid + d * N2] = (mode == 1 && vel) ? vel[node + (int64_t)d * Ntop] : 0.0;
    }
    __syncthreads();
    if (tid < NG) {
      const int qp = tid;
      double IJ[DIM * DIM];
      const double det = jac_qp<DIM, N2, NG>(qp, s_xx, tab->dphi2, IJ);
      s_JxW[qp] = det * tab->w[qp];
      double a = 0., lq = 0., g[DIM], gl[DIM], v[DIM];
      for (int d = 0; d < DIM; d++) { g[d] = 0.; gl[d] = 0.; v[d] = 0.; }
      for (int sn = 0; sn < N2; sn++) {
        const double ph = tab->phi2[sn * NG + qp];
        a += ph * s_T[sn];
        lq += ph * s_l[sn];
        if (mode == 1)
          for (int d = 0; d < DIM; d++) {
            double gd = 0.;
            for (int j = 0; j < DIM; j++) gd += IJ[j + d * DIM] * tab->dphi2[sn * NG + qp + j * N2 * NG];
            g[d] += gd * s_T[sn];
            gl[d] += gd * s_l[sn];
            v[d] += ph * s_u[sn + d * N2];
          }
      }
      s_a[qp] = mode == 1 ? lq : a;
      for (int d = 0; d < DIM; d++) s_g[qp * DIM + d] = g[d];
      if (mode == 1) {
        // derivative of the upwind diffusion: c_m = sum_d D_dm dlam/dx_d dT/dx_d
        double mod2 = 1.e-20;
        for (int d = 0; d < DIM; d++) mod2 += v[d] * v[d];
        mod2 = sqrt(mod2);
        double f[DIM], fd[DIM];
        for (int d = 0; d < DIM; d++) {
          const double h = fabs(s_xx[d * N2 + 2 + 4 * (DIM - 2)] - s_xx[d * N2 + 0]);
          const double h1 = fabs(s_xx[d * N2 + 3 + 4 * (DIM - 2)] - s_xx[d * N2 + 1]);
          const double he = h < h1 ? h1 : h;
          const double Pe = 0.5 * mod2 * he / alpha0, Lp = 1. / tanh(Pe) - 1. / Pe, sh = sinh(Pe);
          const double Ad = upwind * rhocp * 0.5 * he, cd = 0.5 * he / alpha0;
          f[d] = Ad * Lp / mod2;
          fd[d] = Ad * (cd * (1. / (Pe * Pe) - 1. / (sh * sh)) * mod2 - Lp) / (mod2 * mod2);
        }
        for (int mm = 0; mm < DIM; mm++) {
          double c = 0.;
          for (int d = 0; d < DIM; d++) {
            const double D = fd[d] * (v[mm] / mod2) * v[d] * v[d] + (d == mm ? 2. * f[d] * v[d] : 0.);
            c += D * gl[d] * g[d];
          }
          s_c[qp * DIM + mm] = upwind != 0. ? c : 0.;
        }
      }
    }
    __syncthreads();
    if (tid < N2) {
      const int j = tid;
      if (mode == 1) {
        for (int d = 0; d < DIM; d++) {
          double acc = 0.;
          for (int q = 0; q < NG; q++)
            acc += s_JxW[q] * tab->phi2[j * NG + q] * (scale * s_a[q] * s_g[q * DIM + d] + s_c[q * DIM + d]);
          out[(int64_t)d * Ntop + s_node[j]] += acc;
        }
      } else {
        double acc = 0.;
        for (int q = 0; q < NG; q++) acc += s_JxW[q] * tab->phi2[j * NG + q] * s_a[q];
        out[s_node[j]] += scale * acc;
      }
    }
  }
}

namespace fmk {

int coupling(hipStream_t s, int dim, int nel, const int *elems, const int *conn, const double *xyz, int Ntop,
             const int *dof_top, const int *bc0, const double *T, const double *lam, const double *m, double scale, int mode,
             const double *vel, double upwind, double rhocp, double alpha0, const FeTab *tab, double *out) {
  if (nel <= 0) return 0;
  unsigned g = (unsigned)(nel < 8192 ? nel : 8192);
  if (dim == 2)
    hipLaunchKernelGGL(k_coupling<2>, dim3(g), dim3(64), 0, s, nel, elems, conn, xyz, Ntop, dof_top, bc0, T, lam, m, scale,
                       mode, vel, upwind, rhocp, alpha0, tab, out);
  else
    hipLaunchKernelGGL(k_coupling<3>, dim3(g), dim3(64), 0, s, nel, elems, conn, xyz, Ntop, dof_top, bc0, T, lam, m, scale,
                       mode, vel, upwind, rhocp, alpha0, tab, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int heat_nusselt(hipStream_t s, int dim, int nel, const int *elems, const int *conn, const int *eldof, const double *xyz,
                 int Ntop, const int *dof_top, const double *vel, const double *T, const double *coef, double kappa,
                 int axis, const FeTab *tab, double *Jel, double *dJdk, double *y) {
  if (nel <= 0) return 0;
  unsigned g = (unsigned)(nel < 8192 ? nel : 8192);
  if (dim == 2)
    hipLaunchKernelGGL(k_heat_nusselt<2>, dim3(g), dim3(64), 0, s, nel, elems, conn, eldof, xyz, Ntop, dof_top, vel, T,
                       coef, kappa, axis, tab, Jel, dJdk, y);
  else
    hipLaunchKernelGGL(k_heat_nusselt<3>, dim3(g), dim3(64), 0, s, nel, elems, conn, eldof, xyz, Ntop, dof_top, vel, T,
                       coef, kappa, axis, tab, Jel, dJdk, y);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int ns_visc(hipStream_t s, int dim, int nel, const int *elems, const int *conn, const int *eldof, const double *xyz,
            int Ntop, const int *dof_top, const int *bc0, const double *U, const double *lam, const double *coef,
            double nu0, const FeTab *tab, double *diss, double *adj, double *y, double ysc) {
  if (nel <= 0) return 0;
  unsigned g = (unsigned)(nel < 8192 ? nel : 8192);
  if (dim == 2)
    hipLaunchKernelGGL(k_ns_visc<2>, dim3(g), dim3(128), 0, s, nel, elems, conn, eldof, xyz, Ntop, dof_top, bc0, U, lam,
                       coef, nu0, tab, diss, adj, y, ysc);
  else
    hipLaunchKernelGGL(k_ns_visc<3>, dim3(g), dim3(128), 0, s, nel, elems, conn, eldof, xyz, Ntop, dof_top, bc0, U, lam,
                       coef, nu0, tab, diss, adj, y, ysc);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int heat_lap(hipStream_t s, int dim, int nel, const int *elems, const int *conn, const int *eldof, const double *xyz,
             int Ntop, const int *dof_top, const int *bc0, const double *T, const double *lam, const double *coef,
             const FeTab *tab, double *diss, double *adj, double *y, double ysc) {
  if (nel <= 0) return 0;
  unsigned g = (unsigned)(nel < 8192 ? nel : 8192);
  if (dim == 2)
    hipLaunchKernelGGL(k_heat_lap<2>, dim3(g), dim3(64), 0, s, nel, elems, conn, eldof, xyz, Ntop, dof_top, bc0, T, lam,
                       coef, tab, diss, adj, y, ysc);
  else
    hipLaunchKernelGGL(k_heat_lap<3>, dim3(g), dim3(64), 0, s, nel, elems, conn, eldof, xyz, Ntop, dof_top, bc0, T, lam,
                       coef, tab, diss, adj, y, ysc);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// blocks of the one-wave-per-element kernels (four waves per block)
static inline unsigned nblk_w(int nel) {
  int b = (nel + 3) / 4;
  b = b < 8192 ? b : 8192;
  return (unsigned)(b < 1 ? 1 : b);
}
static inline unsigned nblk(int nel) {
  int b = nel < 4096 ? nel : 4096;
  return (unsigned)(b < 1 ? 1 : b);
}

// xbc: Dirichlet data for the boundary-face penalty RHS (nullptr: x_old, as in the reference);
// bd_only: boundary kernel alone (with val == nullptr: the boundary RHS alone)
int heat_assemble(hipStream_t s, int dim, int nel, const int *elems, const int *conn, const int *eldof,
                  const unsigned short *elpos, const int64_t *elslot, const unsigned char *bmask, const double *xyz, int Ntop, const int *dof_top,
                  const int *bc0, const int *bc1, const double *x_old, const double *xbc, const double *coef,
                  const double *vel, const FeTab *tab, HeatArgs a, int mode, int trans, int bd_only, double *val,
                  double *rhs) {
  if (nel <= 0) return 0;
  const double *xb = xbc ? xbc : x_old;
  // the mass / advection part of the element matrix is zero without a time term and without a velocity
  const bool adv = a.tc * a.rhocp != 0. || vel || a.vel_const[0] != 0. || a.vel_const[1] != 0. || a.vel_const[2] != 0.;
  if (dim == 2) {
    if (!bd_only) {
      auto kern = adv ? k_heat_el<2, true, false> : k_heat_el<2, false, false>;
      hipLaunchKernelGGL(kern, dim3(nblk_w(nel)), dim3(256), 0, s, nel, elems, conn, eldof, elpos, elslot, xyz, Ntop, dof_top,
                         bc0, x_old, coef, vel, tab, a, mode, trans, val, rhs, nullptr, nullptr);
    }
    hipLaunchKernelGGL(k_heat_bd<2>, dim3((nel + 127) / 128), dim3(128), 0, s, nel, elems, conn, eldof, elpos, elslot, bmask,
                       xyz, Ntop, dof_top, bc0, bc1, xb, tab, a, mode, trans, val, rhs, nullptr, nullptr);
  } else {
    if (!bd_only) {
      auto kern = adv ? k_heat_el<3, true, false> : k_heat_el<3, false, false>;
      hipLaunchKernelGGL(kern, dim3(nblk_w(nel)), dim3(256), 0, s, nel, elems, conn, eldof, elpos, elslot, xyz, Ntop, dof_top,
                         bc0, x_old, coef, vel, tab, a, mode, trans, val, rhs, nullptr, nullptr);
    }
    hipLaunchKernelGGL(k_heat_bd<3>, dim3((nel + 127) / 128), dim3(128), 0, s, nel, elems, conn, eldof, elpos, elslot, bmask,
                       xyz, Ntop, dof_top, bc0, bc1, xb, tab, a, mode, trans, val, rhs, nullptr, nullptr);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int ns_assemble(hipStream_t s, int dim, int nel, const int *elems, const int *conn, const int *eldof,
                const unsigned short *elpos, const int64_t *elslot, const unsigned char *bmask, const double *xyz, int Ntop, const int *dof_top,
                const int *bc0, const int *bc1, const double *x_old, const double *xbc, const double *coef,
                const FeTab *tab, NsArgs a, int mode, int trans, int bd_only, double *val, double *rhs) {
  if (nel <= 0) return 0;
  const double *xb = xbc ? xbc : x_old;
  if (dim == 2) {
    if (!bd_only)
      hipLaunchKernelGGL((k_ns_vol<2, false>), dim3(nblk(nel)), dim3(256), 0, s, nel, elems, conn, eldof, elpos, elslot, xyz,
                         Ntop, dof_top, bc0, x_old, coef, tab, a, mode, trans, val, rhs, nullptr, nullptr);
    hipLaunchKernelGGL(k_ns_bd<2>, dim3((2 * nel + 127) / 128), dim3(128), 0, s, nel, elems, conn, eldof, elpos, elslot, bmask,
                       xyz, Ntop, dof_top, bc0, bc1, xb, tab, a, mode, trans, val, rhs, nullptr, nullptr);
  } else {
    if (!bd_only)
      hipLaunchKernelGGL((k_ns_vol<3, false>), dim3(nblk(nel)), dim3(256), 0, s, nel, elems, conn, eldof, elpos, elslot, xyz,
                         Ntop, dof_top, bc0, x_old, coef, tab, a, mode, trans, val, rhs, nullptr, nullptr);
    hipLaunchKernelGGL(k_ns_bd<3>, dim3((3 * nel + 127) / 128), dim3(128), 0, s, nel, elems, conn, eldof, elpos, elslot, bmask,
                       xyz, Ntop, dof_top, bc0, bc1, xb, tab, a, mode, trans, val, rhs, nullptr, nullptr);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// ---- element-buffer assembly ------------------------------------------------------------------------------
// All elements of the level in one launch each: the element matrices (E) and right-hand sides (Fe), then the
// boundary faces added to the same element blocks; egather sums them into the SELL slots.
int heat_assemble_e(hipStream_t s, int dim, int nel, const int *conn, const unsigned char *bmask, const double *xyz, int Ntop,
                    const int *dof_top, const int *bc0, const int *bc1, const double *x_old, const double *xbc,
                    const double *coef, const double *vel, const FeTab *tab, HeatArgs a, int mode, int trans, double *E,
                    double *Fe) {
  if (nel <= 0) return 0;
  const double *xb = xbc ? xbc : x_old;
  const bool adv = a.tc * a.rhocp != 0. || vel || a.vel_const[0] != 0. || a.vel_const[1] != 0. || a.vel_const[2] != 0.;
  if (dim == 2) {
    auto kern = adv ? k_heat_el<2, true, true> : k_heat_el<2, false, true>;
    hipLaunchKernelGGL(kern, dim3(nblk_w(nel)), dim3(256), 0, s, nel, nullptr, conn, nullptr, nullptr, nullptr, xyz, Ntop,
                       dof_top, bc0, x_old, coef, vel, tab, a, mode, trans, nullptr, nullptr, E, Fe);
    hipLaunchKernelGGL(k_heat_bd<2>, dim3((nel + 127) / 128), dim3(128), 0, s, nel, nullptr, conn, nullptr, nullptr, nullptr,
                       bmask, xyz, Ntop, dof_top, bc0, bc1, xb, tab, a, mode, trans, nullptr, nullptr, E, Fe);
  } else {
    auto kern = adv ? k_heat_el<3, true, true> : k_heat_el<3, false, true>;
    hipLaunchKernelGGL(kern, dim3(nblk_w(nel)), dim3(256), 0, s, nel, nullptr, conn, nullptr, nullptr, nullptr, xyz, Ntop,
                       dof_top, bc0, x_old, coef, vel, tab, a, mode, trans, nullptr, nullptr, E, Fe);
    hipLaunchKernelGGL(k_heat_bd<3>, dim3((nel + 127) / 128), dim3(128), 0, s, nel, nullptr, conn, nullptr, nullptr, nullptr,
                       bmask, xyz, Ntop, dof_top, bc0, bc1, xb, tab, a, mode, trans, nullptr, nullptr, E, Fe);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int ns_assemble_e(hipStream_t s, int dim, int nel, const int *conn, const unsigned char *bmask, const double *xyz, int Ntop,
                  const int *dof_top, const int *bc0, const int *bc1, const double *x_old, const double *xbc,
                  const double *coef, const FeTab *tab, NsArgs a, int mode, int trans, double *E, double *Fe) {
  if (nel <= 0) return 0;
  const double *xb = xbc ? xbc : x_old;
  if (dim == 2) {
    hipLaunchKernelGGL((k_ns_vol<2, true>), dim3(nblk(nel)), dim3(256), 0, s, nel, nullptr, conn, nullptr, nullptr, nullptr,
                       xyz, Ntop, dof_top, bc0, x_old, coef, tab, a, mode, trans, nullptr, nullptr, E, Fe);
    hipLaunchKernelGGL(k_ns_bd<2>, dim3((2 * nel + 127) / 128), dim3(128), 0, s, nel, nullptr, conn, nullptr, nullptr, nullptr,
                       bmask, xyz, Ntop, dof_top, bc0, bc1, xb, tab, a, mode, trans, nullptr, nullptr, E, Fe);
  } else {
    hipLaunchKernelGGL((k_ns_vol<3, true>), dim3(nblk(nel)), dim3(256), 0, s, nel, nullptr, conn, nullptr, nullptr, nullptr,
                       xyz, Ntop, dof_top, bc0, x_old, coef, tab, a, mode, trans, nullptr, nullptr, E, Fe);
    hipLaunchKernelGGL(k_ns_bd<3>, dim3((3 * nel + 127) / 128), dim3(128), 0, s, nel, nullptr, conn, nullptr, nullptr, nullptr,
                       bmask, xyz, Ntop, dof_top, bc0, bc1, xb, tab, a, mode, trans, nullptr, nullptr, E, Fe);
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// One block per row block (lg lanes of one slice).  Everything the block reads is staged in LDS first with
// independent loads -- its element rows (E), its contribution stream (cidx), the count and stream start of each
// of its slots -- so a block waits for memory once; then every slot is the sum of its contributions in element
// order, written in FP64 and FP32 (each group of lg lanes at one position is contiguous), and the rhs of each row
// the sum of its element rows' Fe entries.  LDS layout (bytes): rows [0, oc), stream [oc, os), starts [os, on),
// counts [on, on + W lg), then the block's element row ids (8-byte aligned).
__global__ __launch_bounds__(256) void k_egather(int lg, int nl, int oc, int os, int on, const int64_t *__restrict__ sl_off,
                                                 const int *__restrict__ sl_w, const int *__restrict__ perm,
                                                 const int64_t *__restrict__ lseg, const int *__restrict__ seg,
                                                 const int64_t *__restrict__ bbase, const unsigned char *__restrict__ cnt,
                                                 const unsigned short *__restrict__ cst, const unsigned short *__restrict__ cidx,
                                                 const double *__restrict__ E, const double *__restrict__ Fe,
                                                 double *__restrict__ v, float *__restrict__ vf, double *__restrict__ rhs) {
  extern __shared__ double smem[];
  unsigned char *sb = (unsigned char *)smem;
  double *se = smem;
  unsigned short *sc = (unsigned short *)(sb + oc), *ss = (unsigned short *)(sb + os);
  unsigned char *sn = sb + on;
  const int64_t g = xcd_block();  // the row blocks of a slice on one XCD: their partial lines meet in one L2
  const int gpb = 64 / lg;
  const int64_t sl = g / gpb;
  const int l0 = (int)(g % gpb) * lg;
  const int64_t s0 = lseg[g * lg], s1 = lseg[(g + 1) * lg];
  const int nseg = (int)(s1 - s0);
  const int64_t c0 = bbase[g], c1 = bbase[g + 1];
  const int nc = (int)(c1 - c0);
  const int W = sl_w[sl], nit = W * lg;
  const int64_t off = sl_off[sl];
  const int tid = threadIdx.x;
  // the loads of each stage are issued in batches of 8 per thread before their LDS stores
  constexpr int B = 8, NT = 256;
  int *sg = (int *)(sb + on + ((nit + 7) / 8) * 8);  // element row ids (after the counts)
  for (int t0 = tid; t0 < nseg; t0 += B * NT) {
    int r[B];
#pragma unroll
    for (int u = 0; u < B; u++) r[u] = t0 + u * NT < nseg ? seg[s0 + t0 + u * NT] : 0;
#pragma unroll
    for (int u = 0; u < B; u++) if (t0 + u * NT < nseg) sg[t0 + u * NT] = r[u];
  }
  for (int t0 = tid; t0 < nc; t0 += B * NT) {
    unsigned short r[B];
#pragma unroll
    for (int u = 0; u < B; u++) r[u] = t0 + u * NT < nc ? cidx[c0 + t0 + u * NT] : 0;
#pragma unroll
    for (int u = 0; u < B; u++) if (t0 + u * NT < nc) sc[t0 + u * NT] = r[u];
  }
  for (int t0 = tid; t0 < nit; t0 += B * NT) {
    unsigned short r[B];
    unsigned char q[B];
#pragma unroll
    for (int u = 0; u < B; u++) {
      const int t = t0 + u * NT;
      const int64_t slot = off + 64 * (int64_t)(t / lg) + l0 + t % lg;
      r[u] = t < nit ? cst[slot] : 0;
      q[u] = t < nit ? cnt[slot] : 0;
    }
#pragma unroll
    for (int u = 0; u < B; u++) if (t0 + u * NT < nit) { ss[t0 + u * NT] = r[u]; sn[t0 + u * NT] = q[u]; }
  }
  __syncthreads();
  for (int t0 = tid; t0 < nseg * nl; t0 += B * NT) {
    double r[B];
#pragma unroll
    for (int u = 0; u < B; u++) {
      const int t = t0 + u * NT;
      r[u] = t < nseg * nl ? E[(int64_t)sg[t / nl] * nl + t % nl] : 0.;
    }
#pragma unroll
    for (int u = 0; u < B; u++) if (t0 + u * NT < nseg * nl) se[t0 + u * NT] = r[u];
  }
  if (rhs && tid < lg) {
    const int row = perm[64 * sl + l0 + tid];
    if (row >= 0) {
      double f = 0.;
      for (int64_t q = lseg[g * lg + tid]; q < lseg[g * lg + tid + 1]; q++) f += Fe[sg[q - s0]];
      rhs[row] = f;
    }
  }
  __syncthreads();
#pragma unroll 4
  for (int t = tid; t < nit; t += NT) {
    const int64_t slot = off + 64 * (int64_t)(t / lg) + l0 + t % lg;
    const int c = sn[t], st = ss[t];
    double sum = 0.;
    for (int m = 0; m < c; m++) sum += se[sc[st + m]];
    v[slot] = sum;
    vf[slot] = (float)sum;
  }
}

int egather(hipStream_t s, int64_t nblk, int lg, int nl, int oc, int os, int on, size_t lds, const int64_t *sl_off,
            const int *sl_w, const int *perm, const int64_t *lseg, const int *seg, const int64_t *bbase,
            const unsigned char *cnt, const unsigned short *cst, const unsigned short *cidx, const double *E,
            const double *Fe, double *v, float *vf, double *rhs) {
  if (nblk <= 0) return 0;
  if (nblk > 0x7fffffff) return -1;
  hipLaunchKernelGGL(k_egather, dim3((unsigned)nblk), dim3(256), lds, s, lg, nl, oc, os, on, sl_off, sl_w, perm, lseg, seg,
                     bbase, cnt, cst, cidx, E, Fe, v, vf, rhs);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace fmk
