// vanka.hip -- element-patch Vanka smoother for the Q2-Q1 Navier-Stokes
// saddle-point systems (gfx950).
//
// The reference keeps Vanka only as experimental LASPACK code
// (contrib/laspack/itersolv.c:156-798, VankaNSIter) reachable through
// VANKANSM (contrib/matrix/laspack_linear_solverM.C:95-101); this build
// defines its own: every element is a patch holding all its velocity and
// pressure DoFs (22 in 2D, 89 in 3D).  Setup extracts the patch block of the
// assembled global matrix (through the element->CSR map) and inverts it in
// FP64 in LDS (Gauss-Jordan, partial pivoting).  A sweep forms r = b - A x
// once and adds omega * w_i * (K_e^-1 r_e)_i to x, colour by colour (no two
// elements of a colour share a DoF, so the scatter needs no atomics);
// w_i = 1 / (number of patches holding DoF i).  The local inverses are
// computed in FP64 and stored transposed (so the threads of a patch read
// them coalesced) in FP64 and as an FP32 copy: the FP64 cycle applies the
// FP64 inverse, the FP32 cycle the FP32 copy with FP32 accumulation (as it
// applies the FP32 copy of the FP64 coarse-level inverse).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdlib>
#include <cstdint>
#include <vector>
#include "fm_internal.h"
#include "fm_kernels.h"
#include "fm_device.h"

// Inverse of every element patch matrix K_e (NL x NL, gathered from the level matrix through the element ->
// slot map), FP64, by exchange steps: the step on pivot (r, k), a = M[r][k], swaps the roles of unknown k and
// equation r,
//   M[r][k] = 1/a, M[r][j] = -M[r][j]/a, M[i][k] = M[i][k]/a, M[i][j] -= M[i][k] M[r][j]/a,
// with r chosen by partial pivoting among the rows not used yet (no row is moved; ties go to the lower row);
// afterwards (K_e^-1)[k][piv[q]] = M[piv[k]][q].
//
// The steps run in blocks of TR columns.  The matrix lives in registers: thread (ty, tx) of a G x G grid owns
// the TR x TR tile at rows ty*TR, columns tx*TR (96 x 96 for the 89 x 89 patch).  For block kt the wavefront
// holding tile column kt takes the block's TR columns through LDS (rows lane and lane + 64 per lane) and runs
// their TR exchange steps alone: pivot search by a wavefront arg-max (DPP inside rows of 16 lanes, a swizzle
// across row pairs, lanes 0 and 32), the pivot row broadcast by readlane, no barrier.  Exchange steps compose,
// so for the pivot rows P and the panel columns K of the block,
//   M'[i][j] = (i in P ? 0 : M[i][j]) - sum_q M'[i][K_q] M[P_q][j]        (j outside K),
// where M'[., K] is the factored panel; every tile takes that rank-TR update after the pivot rows M[P_q][.]
// are published (two barriers per block instead of two per step).  Entries -delta in the published pivot
// rows at the panel's own columns make the same update leave the factored panel in the panel's tiles.
// The gather reads the element map for rows and columns up to G*TR (the map is padded by FM_ELMAP_PAD).
// waves_per_eu(3): 168 registers, three 256-thread patches per CU.

// lane l's copy of v (l uniform)
__device__ __forceinline__ double readlane_d(double v, int l) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// one arg-max stage against the lane that the DPP control CTRL names (ties to the smaller index)
template <int CTRL>
__device__ __forceinline__ void argmax_dpp(double &v, int &i) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, CTRL, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
  const int oi = __builtin_amdgcn_mov_dpp(i, CTRL, 0xF, 0xF, false);
  const double ov = __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
  const bool take = (ov > v) | ((ov == v) & (oi < i));
  v = take ? ov : v;
  i = take ? oi : i;
}

// largest v over the wavefront, ties to the smallest i; every lane gets the result
__device__ __forceinline__ void wave_argmax(double &v, int &i) {
  argmax_dpp<0xB1>(v, i);              // quad_perm [1,0,3,2]
  argmax_dpp<0x4E>(v, i);              // quad_perm [2,3,0,1]
  argmax_dpp<0x141>(v, i);             // row_half_mirror
  argmax_dpp<0x140>(v, i);             // row_mirror: every lane holds its row's result
  {                                    // rows 0/1 and 2/3: swizzle xor 16 inside each half
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_ds_swizzle((int)(unsigned)u, 0x401F);
    const unsigned hi = (unsigned)__builtin_amdgcn_ds_swizzle((int)(unsigned)(u >> 32), 0x401F);
    const int oi = __builtin_amdgcn_ds_swizzle(i, 0x401F);
    const double ov = __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    const bool take = (ov > v) | ((ov == v) & (oi < i));
    v = take ? ov : v;
    i = take ? oi : i;
  }
  const double v0 = readlane_d(v, 0), v1 = readlane_d(v, 32);
  const int i0 = __builtin_amdgcn_readlane(i, 0), i1 = __builtin_amdgcn_readlane(i, 32);
  const bool second = (v1 > v0) | ((v1 == v0) & (i1 < i0));
  v = second ? v1 : v0;
  i = second ? i1 : i0;
}

// LDS writes of this wavefront visible to its own later LDS reads (no other wavefront involved)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <int NL, int G, int TR>
__global__ __launch_bounds__(G * G) __attribute__((amdgpu_waves_per_eu(3))) void k_vanka_invert(
    int nel, const unsigned short *__restrict__ elpos, const int64_t *__restrict__ elslot,
    const double *__restrict__ Av, double *__restrict__ vinv, float *__restrict__ vinvf, int *__restrict__ status) {
  constexpr int NP = G * TR, NT = G * G, NW = NT / 64, NR = (NP + 63) / 64, NB = (NL + TR - 1) / TR;
  constexpr int GW = 64 / G;           // tile columns per wavefront
  static_assert(NP >= NL && NT % 64 == 0 && 64 % G == 0 && NP - 1 + (NP - 1) * NL < NL * NL + FM_ELMAP_PAD,
                "tile grid");
  __shared__ __align__(16) double s_pan[2][TR][NP];   // the block's columns after its exchange steps
  __shared__ __align__(16) double s_a12[2][TR][NP];   // the block's pivot rows before them
  __shared__ double s_red[NW];
  __shared__ __align__(16) int s_step[NP];              // step of pivot row r
  __shared__ int s_piv[NP], s_flag;                    // pivot row of step k, singular
  const int tid = threadIdx.x, tx = tid / G, ty = tid % G, lane = tid & 63, wid = tid >> 6;
  const int i0 = ty * TR, j0 = tx * TR;
  for (int e = xcd_block(); e < nel; e += gridDim.x) {
    const int64_t base = (int64_t)e * NL * NL;
    double m[TR][TR];
    double amax = 0.;
    {
      // branch-free gather in halves of the tile: all element-map loads of a half before its value loads;
      // entries outside the patch or the pattern read slot 0 and are replaced by 0
      constexpr int TH = (TR + 1) / 2;
      const unsigned short *ep = elpos + base + i0 * NL + j0;
      const int64_t *sp = elslot + (int64_t)e * NL + i0;
#pragma unroll
      for (int h = 0; h < TR; h += TH) {
        unsigned short kk[TH][TR];
        int64_t rs[TH];
#pragma unroll
        for (int a = 0; a < TH; a++) {
          rs[a] = sp[h + a];
#pragma unroll
          for (int b = 0; b < TR; b++) kk[a][b] = ep[(h + a) * NL + b];
        }
#pragma unroll
        for (int a = 0; a < TH; a++)
#pragma unroll
          for (int b = 0; b < TR; b++) {
            const bool ok = i0 + h + a < NL && j0 + b < NL && rs[a] >= 0 && kk[a][b] != 0xFFFFu;
            const double v = Av[ok ? rs[a] + 64 * (int64_t)kk[a][b] : 0];
            m[h + a][b] = ok ? v : 0.;
            amax = fmax(amax, fabs(m[h + a][b]));
          }
      }
    }
    for (int off = 32; off > 0; off >>= 1) amax = fmax(amax, __shfl_xor(amax, off, 64));
    __syncthreads();                   // the previous patch's readers of s_piv / s_step / s_flag are done
    if (lane == 0) s_red[wid] = amax;
    for (int t = tid; t < NP; t += NT) s_step[t] = -1;
    if (tid == 0) s_flag = 0;
    __syncthreads();
    double mx = 0.;
#pragma unroll
    for (int w = 0; w < NW; w++) mx = fmax(mx, s_red[w]);
    const double tiny = 1e-14 * mx;
    for (int kt = 0; kt < NB; kt++) {
      const int k0 = kt * TR, nk = min(TR, NL - k0), buf = kt & 1;
      // the wavefront holding tile column kt runs the block's steps (none once the patch is singular: the
      // remaining block updates then work on stale data and the patch is written as zeros)
      if (wid == kt / GW && !s_flag) {
        if (tx == kt)
#pragma unroll
          for (int b = 0; b < TR; b++)
#pragma unroll
            for (int a = 0; a < TR; a++) s_pan[buf][b][i0 + a] = m[a][b];
        wave_lds_sync();
        double p[NR][TR];
        bool fr[NR];                   // row lane + 64 s free (inside the patch, not a pivot row yet)
#pragma unroll
        for (int s = 0; s < NR; s++) {
          const int row = lane + 64 * s;
          const bool in = row < NP;
          const int rc = in ? row : NP - 1;   // every read issued unconditionally, before any use
          double v[TR];
#pragma unroll
          for (int b = 0; b < TR; b++) v[b] = s_pan[buf][b][rc];
          const int st = s_step[rc];
#pragma unroll
          for (int b = 0; b < TR; b++) p[s][b] = in ? v[b] : 0.;
          fr[s] = row < NL && st < 0;
        }
        bool sing = false;
#pragma unroll
        for (int kb = 0; kb < TR; kb++) {
          if (kb < nk && !sing) {
            double bv = -1.;
            int bi = NP;
#pragma unroll
            for (int s = 0; s < NR; s++) {
              const double v = fr[s] ? fabs(p[s][kb]) : -1.;
              if (v > bv) { bv = v; bi = lane + 64 * s; }
            }
            wave_argmax(bv, bi);
            if (!(bv > tiny)) {
              sing = true;
            } else {
              const int r = __builtin_amdgcn_readfirstlane(bi), rl = r & 63, rsl = r >> 6;
              double rr[TR];
#pragma unroll
              for (int b = 0; b < TR; b++) {
                double v = p[0][b];
#pragma unroll
                for (int s = 1; s < NR; s++) v = rsl == s ? p[s][b] : v;
                rr[b] = readlane_d(v, rl);
              }
              double ra = __builtin_amdgcn_rcp(rr[kb]);   // two Newton steps on the hardware reciprocal
              ra = fma(ra, fma(-rr[kb], ra, 1.0), ra);
              ra = fma(ra, fma(-rr[kb], ra, 1.0), ra);
#pragma unroll
              for (int s = 0; s < NR; s++) {
                const bool isp = lane + 64 * s == r;
                const double c = p[s][kb] * ra;
#pragma unroll
                for (int b = 0; b < TR; b++)
                  if (b != kb) p[s][b] = isp ? -rr[b] * ra : fma(-c, rr[b], p[s][b]);
                p[s][kb] = isp ? ra : c;
                fr[s] = fr[s] && !isp;
              }
              if (lane == 0) { s_piv[k0 + kb] = r; s_step[r] = k0 + kb; }
            }
          }
        }
#pragma unroll
        for (int s = 0; s < NR; s++) {
          const int row = lane + 64 * s;
          if (row < NP)
#pragma unroll
            for (int b = 0; b < TR; b++) s_pan[buf][b][row] = p[s][b];
        }
        if (sing && lane == 0) s_flag = 1;
      }
      __syncthreads();                 // factored panel, pivots and flag published
      // the pivot rows outside the panel's columns; -delta at the panel's columns
      int q[TR];                       // q[a] in [0, nk): row i0 + a is the block's pivot row q
      if constexpr (TR % 2 == 0) {     // all reads before the writes (8-byte pairs: i0 = ty TR is even)
        const int2 *sp2 = reinterpret_cast<const int2 *>(s_step + i0);
#pragma unroll
        for (int a = 0; a < TR; a += 2) {
          const int2 v = sp2[a / 2];
          q[a] = v.x - k0;
          q[a + 1] = v.y - k0;
        }
      } else {
#pragma unroll
        for (int a = 0; a < TR; a++) q[a] = s_step[i0 + a] - k0;
      }
#pragma unroll
      for (int a = 0; a < TR; a++) {
        if (tx != kt && q[a] >= 0 && q[a] < nk)
#pragma unroll
          for (int b = 0; b < TR; b++) s_a12[buf][q[a]][j0 + b] = m[a][b];
      }
      if (tid < TR * TR && tid / TR < nk) s_a12[buf][tid / TR][k0 + tid % TR] = tid / TR == tid % TR ? -1. : 0.;
      for (int t = nk * NP + tid; t < TR * NP; t += NT) s_a12[buf][t / NP][t % NP] = 0.;
      __syncthreads();                 // the pivot rows published
#pragma unroll
      for (int a = 0; a < TR; a++) {
        const double keep = (tx == kt || (q[a] >= 0 && q[a] < nk)) ? 0. : 1.;
#pragma unroll
        for (int b = 0; b < TR; b++) m[a][b] *= keep;
      }
#pragma unroll
      for (int c = 0; c < TR; c++) {
        double lc[TR], ar[TR];
#pragma unroll
        for (int a = 0; a < TR; a++) lc[a] = s_pan[buf][c][i0 + a];
#pragma unroll
        for (int b = 0; b < TR; b++) ar[b] = s_a12[buf][c][j0 + b];
#pragma unroll
        for (int a = 0; a < TR; a++)
#pragma unroll
          for (int b = 0; b < TR; b++) m[a][b] = fma(-lc[a], ar[b], m[a][b]);
      }
    }
    if (s_flag) {
      if (tid == 0) atomicAdd(status, 1);
      for (int t = tid; t < NL * NL; t += NT) {
        if (vinv) vinv[base + t] = 0.0;
        if (vinvf) vinvf[base + t] = 0.f;
      }
      continue;
    }
    // transposed store: (K_e^-1)[step[i]][piv[j]] = M[i][j] at piv[j] * NL + step[i]
#pragma unroll
    for (int a = 0; a < TR; a++)
#pragma unroll
      for (int b = 0; b < TR; b++) {
        const int i = i0 + a, j = j0 + b;
        if (i < NL && j < NL) {
          const int64_t o = base + (int64_t)s_piv[j] * NL + s_step[i];
          if (vinv) vinv[o] = m[a][b];
          if (vinvf) vinvf[o] = (float)m[a][b];
        }
      }
  }
}

// x[dof] += omega * w[dof] * (K_e^-1 r_e) for the elements of one colour; the
// inverse is stored transposed (vinv[e][b][a] = (K_e^-1)_ab) in the cycle's precision T.
// 3D patches (NL = 89): 128 threads per element, thread a owning row a; for every column b the threads read
// the NL contiguous entries vinv[e][b][0 .. NL) (two waves per element keep more loads in flight than one:
// 267 against 322 ms for the C3 channel solve).
template <int NL, typename T>
__global__ __launch_bounds__(128) void k_vanka_apply(int nel, const int *__restrict__ elems, const int *__restrict__ eldof,
                                                     const T *__restrict__ vinv, const T *__restrict__ r,
                                                     T *__restrict__ x, const double *__restrict__ wt, double omega) {
  __shared__ T re[NL];
  __shared__ int dof[NL];
  const int tid = threadIdx.x;
  for (int ei = xcd_block(); ei < nel; ei += gridDim.x) {
    const int e = elems[ei];
    __syncthreads();
    for (int a = tid; a < NL; a += blockDim.x) {
      int d = eldof[(int64_t)e * NL + a];
      dof[a] = d;
      re[a] = r[d];
    }
    __syncthreads();
    const T *Ki = vinv + (int64_t)e * NL * NL;
    for (int a = tid; a < NL; a += blockDim.x) {
      T s = 0;
      for (int b = 0; b < NL; b++) s += Ki[b * NL + a] * re[b];
      const int d = dof[a];
      x[d] = x[d] + (T)(omega * wt[d]) * s;
    }
  }
}

// 2D patches (NL = 22): one wavefront per element, four per block (a 128-thread block would leave 106
// threads idle); lane l owns rows l and l + 64.
template <int NL, typename T>
__global__ __launch_bounds__(256) void k_vanka_apply_w(int nel, const int *__restrict__ elems, const int *__restrict__ eldof,
                                                     const T *__restrict__ vinv, const T *__restrict__ r,
                                                     T *__restrict__ x, const double *__restrict__ wt, double omega) {
  static_assert(NL <= 128, "two rows per lane");
  __shared__ T re[4][NL];
  __shared__ int dof[4][NL];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  for (int eb = 4 * xcd_block(); eb < nel; eb += 4 * (int)gridDim.x) {
    const int ei = eb + w;
    const bool act = ei < nel;
    const int e = act ? elems[ei] : 0;
    __syncthreads();  // the previous group's readers of re / dof are done
    if (act)
      for (int a = l; a < NL; a += 64) {
        const int d = eldof[(int64_t)e * NL + a];
        dof[w][a] = d;
        re[w][a] = r[d];
      }
    __syncthreads();
    if (!act) continue;
    const T *Ki = vinv + (int64_t)e * NL * NL;
    const bool r0 = l < NL, r1 = l + 64 < NL;
    T s0 = 0, s1 = 0;
#pragma unroll 8
    for (int b = 0; b < NL; b++) {
      const T rb = re[w][b];
      if (r0) s0 += Ki[b * NL + l] * rb;
      if (r1) s1 += Ki[b * NL + 64 + l] * rb;
    }
    if (r0) {
      const int d = dof[w][l];
      x[d] = x[d] + (T)(omega * wt[d]) * s0;
    }
    if (r1) {
      const int d = dof[w][l + 64];
      x[d] = x[d] + (T)(omega * wt[d]) * s1;
    }
  }
}

__global__ void k_count_mult(int nel, int nl, const int *__restrict__ eldof, double *__restrict__ cnt) {
  int64_t total = (int64_t)nel * nl;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[eldof[t]], 1.0);
}
__global__ void k_recip(int64_t n, double *w) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    w[i] = w[i] > 0. ? 1.0 / w[i] : 0.0;
}


// patch inverses of the matrix with values Av (A or A^T: same pattern) into *store
static int vanka_invert_level(fm_system *s, int L, const double *Av, double **store, float **storef) {
  Level &l = s->lev[L];
  hipStream_t st = s->ctx->stream;
  const int nl = s->nloc;
  if (s->nq1 == 0) { fm_set_error("Vanka needs a Q2-Q1 (velocity-pressure) system"); return FM_EINVAL; }
  // the inverses are kept in the precisions the system's store mask names (bit 0 FP64, bit 1 FP32); a store
  // left out is released
  const bool k64 = (s->vanka_store & 1) != 0, k32 = (s->vanka_store & 2) != 0;
  if (!k64 && *store) { (void)hipFree(*store); *store = nullptr; }
  if (!k32 && *storef) { (void)hipFree(*storef); *storef = nullptr; }
  if (k64 && !*store) {
    if (hipMalloc((void **)store, (size_t)l.nel * nl * nl * sizeof(double)) != hipSuccess) { fm_set_error("Vanka alloc"); return FM_ENOMEM; }
  }
  if (k32 && !*storef) {
    if (hipMalloc((void **)storef, (size_t)l.nel * nl * nl * sizeof(float)) != hipSuccess) { fm_set_error("Vanka alloc"); return FM_ENOMEM; }
  }
  if (!l.vwt) {
    if (hipMalloc((void **)&l.vwt, l.n * sizeof(double)) != hipSuccess) { fm_set_error("Vanka alloc"); return FM_ENOMEM; }
  }
  if (!l.vone) {
    if (hipMalloc((void **)&l.vone, l.n * sizeof(double)) != hipSuccess) { fm_set_error("Vanka alloc"); return FM_ENOMEM; }
    if (fmk::fill<double>(st, l.n, l.vone, 1.0) != 0) return FM_EHIP;
  }
  int *status = nullptr;
  if (hipMalloc((void **)&status, sizeof(int)) != hipSuccess) return FM_ENOMEM;
  FM_HIP(hipMemsetAsync(status, 0, sizeof(int), st));
  FM_HIP(hipMemsetAsync(l.vwt, 0, l.n * sizeof(double), st));
  unsigned g = (unsigned)(l.nel < 16384 ? l.nel : 16384);
  if (s->mesh->dim == 2)
    hipLaunchKernelGGL((k_vanka_invert<22, 8, 3>), dim3(g), dim3(64), 0, st, l.nel, l.elpos, l.elslot, Av, *store, *storef, status);
  else
    hipLaunchKernelGGL((k_vanka_invert<89, 16, 6>), dim3(g), dim3(256), 0, st, l.nel, l.elpos, l.elslot, Av, *store, *storef, status);
  FM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_count_mult, dim3(1024), dim3(256), 0, st, l.nel, nl, l.eldofi, l.vwt);
  hipLaunchKernelGGL(k_recip, dim3(1024), dim3(256), 0, st, l.n, l.vwt);
  int h_status = 0;
  FM_HIP(hipMemcpyAsync(&h_status, status, sizeof(int), hipMemcpyDeviceToHost, st));
  FM_HIP(hipStreamSynchronize(st));
  (void)hipFree(status);
  if (h_status > 0) {
    fm_set_error("Vanka: " + std::to_string(h_status) + " singular element patch(es) on level " + std::to_string(L));
    return FM_ESINGULAR;
  }
  return 0;
}

// xz: x is zero on entry, so the first residual is b (copied, no product)
// grid cap of the 3D patch application (FEMUS_VANKA_GRID, default 131072 blocks): one block per element up to
// the cap.  C5 at 48^3 (13,824 elements per colour on the top level): 643 ms per coupled step with a cap of 8192
// (each block walking ~2 elements), 628 ms with one block per element.
static int g_vanka_grid = [] {
  const char *e = getenv("FEMUS_VANKA_GRID");
  const int v = e ? atoi(e) : 0;
  return v >= 256 ? v : 131072;
}();

// colour order of the post-smoothing sweeps (FEMUS_VANKA_POST_ORDER): 1 the reverse of the pre-smoothing
// order (a symmetric V-cycle), 0 the same order; by default reversed for 3D patches and the same for 2D ones,
// as measured: 96x24x24 channel (Vanka 2/2) 20 FGMRES iterations reversed against 23, Boussinesq cube 48^3
// (3/3) 15 against 16 (96^3: 19 either way); 2D cavity 32^2 (3/3) 16 reversed against 13, its adjoint gradient
// 806 against 753 ms.  Reversing every other sweep instead applies one colour twice in a row at each turn
// (33 iterations on the channel).
static int g_vanka_post_order = [] {
  const char *e = getenv("FEMUS_VANKA_POST_ORDER");
  return e ? (atoi(e) != 0) : -1;
}();

template <typename T>
static int vanka_sweep(fm_system *s, int L, const CsrView<T> &A, const T *vinv, T *x, const T *b, T *r, int its,
                       double omega, int colored, bool xz, bool post) {
  Level &l = s->lev[L];
  hipStream_t st = s->ctx->stream;
  // additive: one residual per sweep, damped and weighted by 1/multiplicity;
  // coloured: residual refreshed before every colour, undamped weights 1
  const double *wt = l.vwt;
  bool fresh = xz;  // x still zero: the residual is b
  const bool rev_post = g_vanka_post_order >= 0 ? g_vanka_post_order == 1 : s->mesh->dim == 3;
  // colour c > = 0: only the rows that colour's elements touch (whole SELL slices, see Level::cslist)
  auto residual = [&](int c) -> int {
    if (fresh) {
      fresh = false;
      return fmk::copy<T>(st, l.n, b, r) == 0 ? 0 : FM_EHIP;
    }
    if (c >= 0 && l.cslist && c + 1 < (int)l.cs_off.size()) {
      CsrView<T> Ac = A;
      Ac.slist = l.cslist + l.cs_off[c];
      Ac.nlist = l.cs_off[c + 1] - l.cs_off[c];
      return fmk::resid(st, Ac, b, x, r);
    }
    return fmk::resid(st, A, b, x, r);
  };
  for (int it = 0; it < its; it++) {
    if (!colored) {
      int rc = residual(-1);
      if (rc) return rc;
    }
    for (int cc = 0; cc < l.ncolor; cc++) {
      const int c = post && rev_post ? l.ncolor - 1 - cc : cc;
      int n = l.color_off[c + 1] - l.color_off[c];
      if (n <= 0) continue;
      if (colored) {
        int rc = residual(c);
        if (rc) return rc;
      }
      const double *w = colored ? l.vone : wt;
      if (s->mesh->dim == 2)
        hipLaunchKernelGGL((k_vanka_apply_w<22, T>), dim3((unsigned)std::min((n + 3) / 4, 8192)), dim3(256), 0, st, n,
                           l.color_elems + l.color_off[c], l.eldofi, vinv, r, x, w, omega);
      else
        hipLaunchKernelGGL((k_vanka_apply<89, T>), dim3((unsigned)std::min(n, g_vanka_grid)), dim3(128), 0, st, n,
                           l.color_elems + l.color_off[c], l.eldofi, vinv, r, x, w, omega);
    }
  }
  return hipGetLastError() == hipSuccess ? 0 : FM_EHIP;
}

int fm_vanka_invert(fm_system *s, int L, const double *Av, double **store, float **storef) {
  return vanka_invert_level(s, L, Av, store, storef);
}
int fm_vanka_apply64(fm_system *s, int L, const CsrView<double> &A, const double *vinv, double *x, const double *b,
                     double *r, int its, double omega, int colored, bool xz, bool post) {
  return vanka_sweep<double>(s, L, A, vinv, x, b, r, its, omega, colored, xz, post);
}
int fm_vanka_apply32(fm_system *s, int L, const CsrView<float> &A, const float *vinv, float *x, const float *b,
                     float *r, int its, double omega, int colored, bool xz, bool post) {
  return vanka_sweep<float>(s, L, A, vinv, x, b, r, its, omega, colored, xz, post);
}

// Diagnostic: one element's patch inverse as stored for the smoother (FP64, transposition undone) and the
// patch DoFs in reference numbering, local order [u, v, (w), p].
extern "C" int fm_vanka_patch(const fm_system *s, int L, int e, int *dofs, double *kinv) {
  if (!s || !dofs || !kinv || L < 0 || L >= (int)s->lev.size()) return fm_einval(__func__);
  const Level &l = s->lev[L];
  if (!s->mg_ready || !l.vinv) { fm_set_error("no Vanka inverses: run fm_mg_setup with the Vanka smoother"); return FM_ESTATE; }
  if (e < 0 || e >= l.nel) return fm_einval(__func__);
  const int nl = s->nloc;
  std::vector<double> t((size_t)nl * nl);
  hipStream_t st = s->ctx->stream;
  FM_HIP(hipSetDevice(s->ctx->device));
  FM_HIP(hipMemcpyAsync(dofs, l.eldof + (int64_t)e * nl, nl * sizeof(int), hipMemcpyDeviceToHost, st));
  FM_HIP(hipMemcpyAsync(t.data(), l.vinv + (int64_t)e * nl * nl, t.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  FM_HIP(hipStreamSynchronize(st));
  for (int a = 0; a < nl; a++)
    for (int b = 0; b < nl; b++) kinv[a * nl + b] = t[(size_t)b * nl + a];
  return FM_OK;
}
