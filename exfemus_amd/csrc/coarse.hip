// coarse.hip -- dense FP64 inverse of the coarsest multigrid level on the device.
//
// The reference solves its coarsest level with the smoother's Krylov method
// (60 iterations, src/MGSolverBase.C:223); here the level matrix is inverted
// once per assembly and every coarse solve is one dense mat-vec.
//
// Method: blocked in-place exchange (sweep) inversion.  The exchange on pivot
// (r, c), a = M[r][c], swaps the roles of unknown x_c and equation r:
//   M[r][c] = 1/a,  M[r][j] = -M[r][j]/a,  M[i][c] = M[i][c]/a,
//   M[i][j] -= M[i][c] M[r][j] / a                       (i != r, j != c).
// After exchanges on (r_p, p) for every column p, x_p = sum_q M[r_p][q] y_{r_q}.
// Pivot rows are chosen by partial pivoting among the rows not used yet; no
// row is moved (r_p is recorded instead).  Columns are taken in panels of NB:
//   1. the panel's columns are copied to an n x NB column-major buffer and its
//      NB exchanges run there, one launch each (the launch also finds the next
//      column's pivot: a 64-bit atomic max over (|value| as float, row));
//   2. the rest of the matrix takes the panel's block exchange (rows R = the
//      panel's pivot rows, columns K = the panel) as one rank-NB update
//        M[:, J] -= L W,   W = M[R, J] before the update,
//        L = panel result with 1 added at (r_q, q),
//      which gives M'[R,J] = -P M[R,J] on the pivot rows (P = M'[R,K]) and
//      M'[I,J] = M[I,J] - M'[I,K] M[R,J] on the others;
//   3. the panel result is copied back into columns K.
// A pinned unknown (enclosed flow: the pressure is defined up to a constant)
// is processed last; when its pivot is below the singularity threshold its
// exchange is skipped, which solves the system with that unknown set to 0
// and the one equation left without a pivot dropped.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <vector>
#include "fm_internal.h"

namespace {

constexpr int NB = 64;     // panel width

struct SwInfo {
  int singular, pinned;
};

__device__ __forceinline__ unsigned long long sw_key(double v, int i) {
  // presence bit | |v| rounded to float (bits 62..32) | ~row (ties -> lowest row)
  const unsigned fb = __float_as_uint((float)fabs(v));
  return (1ull << 63) | ((unsigned long long)(fb & 0x7fffffffu) << 32) | (unsigned long long)(0xffffffffu - (unsigned)i);
}
__device__ __forceinline__ int sw_row(unsigned long long key) {
  return key ? (int)(0xffffffffu - (unsigned)(key & 0xffffffffull)) : -1;
}
__device__ __forceinline__ void sw_block_max(unsigned long long cand, unsigned long long *slot) {
  __shared__ unsigned long long s_best[4];
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long v = __shfl_xor(cand, o, 64);
    cand = v > cand ? v : cand;
  }
  if ((threadIdx.x & 63) == 0) s_best[threadIdx.x >> 6] = cand;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long b = s_best[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) b = s_best[w] > b ? s_best[w] : b;
    if (b) atomicMax(slot, b);
  }
}

// M[rank[i]][colpos[rank[ci]]] = A[i][ci] (the level matrix in internal order, columns permuted)
__global__ void k_sw_init(int n, const int64_t *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ v,
                          const int *__restrict__ rank, const int *__restrict__ colpos, double *__restrict__ M) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int64_t r = rank[i];
    for (int64_t q = rp[i]; q < rp[i + 1]; q++) M[r * n + colpos[rank[ci[q]]]] = v[q];
  }
}

// P[t][i] = M[i][k0 + t] (t < bw): 64 x 64 tiles transposed through LDS
__global__ __launch_bounds__(256) void k_sw_gather(int n, int k0, int bw, const double *__restrict__ M, double *__restrict__ P,
                                                   const SwInfo *info) {
  if (info->singular) return;
  __shared__ double s[64][65];
  const int i0 = blockIdx.x * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int ii = e >> 6, t = e & 63;
    const int i = i0 + ii;
    s[ii][t] = (i < n && t < bw) ? M[(int64_t)i * n + k0 + t] : 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int t = e >> 6, ii = e & 63;
    const int i = i0 + ii;
    if (i < n && t < bw) P[(int64_t)t * n + i] = s[ii][t];
  }
}
// M[i][k0 + t] = P[t][i]
__global__ __launch_bounds__(256) void k_sw_scatter(int n, int k0, int bw, double *__restrict__ M, const double *__restrict__ P,
                                                    const SwInfo *info) {
  if (info->singular) return;
  __shared__ double s[64][65];
  const int i0 = blockIdx.x * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int t = e >> 6, ii = e & 63;
    const int i = i0 + ii;
    s[ii][t] = (i < n && t < bw) ? P[(int64_t)t * n + i] : 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int ii = e >> 6, t = e & 63;
    const int i = i0 + ii;
    if (i < n && t < bw) M[(int64_t)i * n + k0 + t] = s[ii][t];
  }
}

// pivot candidate of panel column 0 among the rows without a pivot
__global__ __launch_bounds__(256) void k_sw_argmax(int n, const double *__restrict__ P, const int *__restrict__ rowstep,
                                                   unsigned long long *slot, const SwInfo *info) {
  if (info->singular) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  unsigned long long c = 0;
  if (i < n && rowstep[i] < 0) c = sw_key(P[i], i);
  sw_block_max(c, slot);
}

// exchange step p = k0 + t on the panel (Pin -> Pout); grid (ceil(n / 256), bw)
__global__ __launch_bounds__(256) void k_sw_step(int n, int p, int t, int bw, const double *__restrict__ Pin,
                                                 double *__restrict__ Pout, int *__restrict__ rowstep, int *__restrict__ prow,
                                                 unsigned long long *__restrict__ keys, double thr, int pin_step,
                                                 SwInfo *info) {
  if (info->singular) return;
  const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
  const int r = sw_row(keys[p]);
  const double a = r >= 0 ? Pin[(int64_t)t * n + r] : 0.0;
  const bool ok = r >= 0 && fabs(a) > thr;      // false for NaN
  unsigned long long cand = 0;
  if (i < n) {
    const double pij = Pin[(int64_t)j * n + i];
    double out = pij;
    if (ok) {
      const double ra = 1.0 / a;
      const double pic = Pin[(int64_t)t * n + i], prj = Pin[(int64_t)j * n + r];
      if (i == r) out = (j == t) ? ra : -prj * ra;
      else out = (j == t) ? pic * ra : pij - pic * (prj * ra);
    }
    Pout[(int64_t)j * n + i] = out;
    if (ok && j == t + 1 && i != r && rowstep[i] < 0) cand = sw_key(out, i);
  }
  if (ok && j == t + 1) sw_block_max(cand, keys + p + 1);
  if (blockIdx.x == 0 && j == 0 && threadIdx.x == 0) {
    if (ok) {
      rowstep[r] = p;
      prow[p] = r;
    } else {
      prow[p] = -1;
      if (p == pin_step) info->pinned = 1;
      else info->singular = 1;
    }
  }
}

// W[q][j] = M[prow[k0 + q]][j] (0 for a skipped pivot); grid (ceil(n / 256), bw)
__global__ __launch_bounds__(256) void k_sw_rows(int n, int k0, const double *__restrict__ M, const int *__restrict__ prow,
                                                 double *__restrict__ W, const SwInfo *info) {
  if (info->singular) return;
  const int j = blockIdx.x * 256 + threadIdx.x, q = blockIdx.y;
  if (j >= n) return;
  const int r = prow[k0 + q];
  W[(int64_t)q * n + j] = r >= 0 ? M[(int64_t)r * n + j] : 0.0;
}

// M[i][j] -= sum_q L[i][q] W[q][j] for every column j outside the panel;
// L[i][q] = P[q][i] + (rowstep[i] == k0 + q).  128 x 128 output tile per
// workgroup, 8 x 8 per thread (rows ty + 16 a, columns tx + 16 b), the 64-deep
// panel taken in two chunks of 32 through LDS.  Grid (ceil(n / 128), ceil(n / 128)).
constexpr int UT = 128, UK = 32;
__global__ __launch_bounds__(256) void k_sw_update(int n, int k0, int bw, double *__restrict__ M, const double *__restrict__ P,
                                                   const double *__restrict__ W, const int *__restrict__ rowstep,
                                                   const SwInfo *info) {
  if (info->singular) return;
  const int j0 = blockIdx.x * UT, i0 = blockIdx.y * UT;
  __shared__ double sL[UK][UT], sW[UK][UT];
  const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
  double acc[8][8] = {};
  for (int c0 = 0; c0 < NB; c0 += UK) {
    __syncthreads();
    for (int e = threadIdx.x; e < UK * UT; e += 256) {
      const int q = c0 + e / UT, x = e % UT;
      const int i = i0 + x, j = j0 + x;
      double l = 0., w = 0.;
      if (q < bw) {
        if (i < n) l = P[(int64_t)q * n + i] + (rowstep[i] == k0 + q ? 1.0 : 0.0);
        if (j < n) w = W[(int64_t)q * n + j];
      }
      sL[e / UT][x] = l;
      sW[e / UT][x] = w;
    }
    __syncthreads();
#pragma unroll 2
    for (int q = 0; q < UK; q++) {
      double lv[8], wv[8];
#pragma unroll
      for (int a = 0; a < 8; a++) lv[a] = sL[q][ty + 16 * a];
#pragma unroll
      for (int b = 0; b < 8; b++) wv[b] = sW[q][tx + 16 * b];
#pragma unroll
      for (int a = 0; a < 8; a++)
#pragma unroll
        for (int b = 0; b < 8; b++) acc[a][b] = fma(lv[a], wv[b], acc[a][b]);
    }
  }
#pragma unroll
  for (int a = 0; a < 8; a++) {
    const int i = i0 + ty + 16 * a;
    if (i >= n) continue;
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const int j = j0 + tx + 16 * b;
      if (j < n && (j < k0 || j >= k0 + bw)) M[(int64_t)i * n + j] -= acc[a][b];
    }
  }
}

// Ainv[d][t] = M[prow[colpos[d]]][rowstep[t]] (0 where either index is missing), FP64 and FP32 copies;
// grid (ceil(n / 256), n)
__global__ __launch_bounds__(256) void k_sw_output(int n, const double *__restrict__ M, const int *__restrict__ prow,
                                                   const int *__restrict__ rowstep, const int *__restrict__ colpos,
                                                   double *__restrict__ Ainv, float *__restrict__ Ainvf) {
  const int t = blockIdx.x * 256 + threadIdx.x, d = blockIdx.y;
  if (t >= n) return;
  const int r = prow[colpos[d]], q = rowstep[t];
  const double v = (r >= 0 && q >= 0) ? M[(int64_t)r * n + q] : 0.0;
  Ainv[(int64_t)d * n + t] = v;
  Ainvf[(int64_t)d * n + t] = (float)v;
}

}  // namespace

// Dense inverse of the level matrix A (CSR rows in reference order) in the level's internal order (rank).
// pin: internal index of the unknown that may be pinned (-1: none).  thr: singularity threshold on |pivot|.
int fm_dense_inverse(hipStream_t st, const DevCSR &A, const int *d_rank, int pin, double thr, double **dAinv,
                     float **dAinvf) {
  const int n = (int)A.nr;
  if (n <= 0 || n > 65535) { fm_set_error("coarse level size outside 1..65535"); return FM_EINVAL; }
  const size_t nn = (size_t)n * n;
  if (!*dAinv && hipMalloc((void **)dAinv, nn * sizeof(double)) != hipSuccess) return FM_ENOMEM;
  if (!*dAinvf && hipMalloc((void **)dAinvf, nn * sizeof(float)) != hipSuccess) return FM_ENOMEM;
  std::vector<int> colpos(n);
  for (int i = 0; i < n; i++) colpos[i] = i;
  if (pin >= 0) std::swap(colpos[pin], colpos[n - 1]);
  double *M = nullptr, *P0 = nullptr, *P1 = nullptr, *W = nullptr;
  int *ibuf = nullptr;
  unsigned long long *keys = nullptr;
  SwInfo *info = nullptr;
  auto cleanup = [&]() {
    (void)hipFree(M); (void)hipFree(P0); (void)hipFree(P1); (void)hipFree(W);
    (void)hipFree(ibuf); (void)hipFree(keys); (void)hipFree(info);
  };
  if (hipMalloc((void **)&M, nn * sizeof(double)) != hipSuccess || hipMalloc((void **)&P0, (size_t)n * NB * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&P1, (size_t)n * NB * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&W, (size_t)n * NB * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&ibuf, 3 * (size_t)n * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&keys, (size_t)n * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc((void **)&info, sizeof(SwInfo)) != hipSuccess) {
    cleanup();
    return FM_ENOMEM;
  }
  int *rowstep = ibuf, *prow = ibuf + n, *d_colpos = ibuf + 2 * n;
  int rc = 0;
  const SwInfo h0{0, 0};
  if (hipMemsetAsync(M, 0, nn * sizeof(double), st) != hipSuccess ||
      hipMemsetAsync(ibuf, 0xff, 2 * (size_t)n * sizeof(int), st) != hipSuccess ||
      hipMemcpyAsync(d_colpos, colpos.data(), n * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemsetAsync(keys, 0, (size_t)n * sizeof(unsigned long long), st) != hipSuccess ||
      hipMemcpyAsync(info, &h0, sizeof(SwInfo), hipMemcpyHostToDevice, st) != hipSuccess) {
    cleanup();
    return FM_EHIP;
  }
  const unsigned g256 = (unsigned)((n + 255) / 256), g64 = (unsigned)((n + 63) / 64);
  hipLaunchKernelGGL(k_sw_init, dim3(std::min(1024u, g256)), dim3(256), 0, st, n, A.rp, A.ci, A.v, d_rank, d_colpos, M);
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int bw = std::min(NB, n - k0);
    hipLaunchKernelGGL(k_sw_gather, dim3(g64), dim3(256), 0, st, n, k0, bw, M, P0, info);
    hipLaunchKernelGGL(k_sw_argmax, dim3(g256), dim3(256), 0, st, n, P0, rowstep, keys + k0, info);
    double *pin_buf = P0, *pout = P1;
    for (int t = 0; t < bw; t++) {
      hipLaunchKernelGGL(k_sw_step, dim3(g256, (unsigned)bw), dim3(256), 0, st, n, k0 + t, t, bw, pin_buf, pout, rowstep,
                         prow, keys, thr, pin >= 0 ? n - 1 : -1, info);
      std::swap(pin_buf, pout);
    }
    hipLaunchKernelGGL(k_sw_rows, dim3(g256, (unsigned)bw), dim3(256), 0, st, n, k0, M, prow, W, info);
    const unsigned gu = (unsigned)((n + UT - 1) / UT);
    hipLaunchKernelGGL(k_sw_update, dim3(gu, gu), dim3(256), 0, st, n, k0, bw, M, pin_buf, W, rowstep, info);
    hipLaunchKernelGGL(k_sw_scatter, dim3(g64), dim3(256), 0, st, n, k0, bw, M, pin_buf, info);
  }
  SwInfo h{0, 0};
  if (hipGetLastError() != hipSuccess || hipMemcpyAsync(&h, info, sizeof(SwInfo), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    rc = FM_EHIP;
  } else if (h.singular) {
    fm_set_error("coarse matrix singular");
    rc = FM_ESINGULAR;
  } else {
    hipLaunchKernelGGL(k_sw_output, dim3(g256, (unsigned)n), dim3(256), 0, st, n, M, prow, rowstep, d_colpos, *dAinv, *dAinvf);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) rc = FM_EHIP;
  }
  cleanup();
  return rc;
}
