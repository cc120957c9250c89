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
#include <cstdint>
#include <vector>
#include "fm_internal.h"
#include "fm_kernels.h"

template <int NL>
__global__ __launch_bounds__(256) void k_vanka_invert(int nel, const int *__restrict__ elpos, const double *__restrict__ Av,
                                                      double *__restrict__ vinv, float *__restrict__ vinvf,
                                                      int *__restrict__ status) {
  __shared__ double M[NL * NL];
  __shared__ int piv[NL];
  __shared__ double s_best, s_red[4];
  __shared__ int s_bi;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  for (int e = blockIdx.x; e < nel; e += gridDim.x) {
    __syncthreads();
    const int64_t base = (int64_t)e * NL * NL;
    double amax = 0.;
    for (int t = tid; t < NL * NL; t += blockDim.x) {
      int p = elpos[base + t];
      double v = p >= 0 ? Av[p] : 0.0;
      M[t] = v;
      amax = fmax(amax, fabs(v));
    }
    // block maximum of |M| (the singularity threshold)
    for (int off = 32; off > 0; off >>= 1) amax = fmax(amax, __shfl_xor(amax, off, 64));
    if (lane == 0) s_red[wid] = amax;
    __syncthreads();
    if (tid == 0) {
      double m = 0.;
      for (int w = 0; w < (int)(blockDim.x >> 6); w++) m = fmax(m, s_red[w]);
      s_best = m;
    }
    __syncthreads();
    const double tiny = 1e-14 * s_best;
    bool singular = false;
    for (int k = 0; k < NL; k++) {
      // partial pivoting: the first wave finds the largest |M[i][k]|, i >= k (first index on ties)
      if (tid < 64) {
        int bi = NL;
        double bv = -1.;
        for (int i = k + lane; i < NL; i += 64) {
          const double v = fabs(M[i * NL + k]);
          if (v > bv) { bv = v; bi = i; }
        }
        for (int off = 32; off > 0; off >>= 1) {
          const double ov = __shfl_xor(bv, off, 64);
          const int oi = __shfl_xor(bi, off, 64);
          if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
        }
        if (tid == 0) {
          s_bi = bi;
          piv[k] = bi;
          s_best = bv;
        }
      }
      __syncthreads();
      if (!(s_best > tiny)) { singular = true; break; }
      const int bi = s_bi;
      if (bi != k)
        for (int j = tid; j < NL; j += blockDim.x) {
          double t = M[k * NL + j]; M[k * NL + j] = M[bi * NL + j]; M[bi * NL + j] = t;
        }
      __syncthreads();
      const double inv = 1.0 / M[k * NL + k];
      __syncthreads();
      for (int j = tid; j < NL; j += blockDim.x) M[k * NL + j] = (j == k ? 1.0 : M[k * NL + j]) * inv;
      __syncthreads();
      // eliminate column k from the other rows; column k of those rows becomes -f * inv
      for (int t = tid; t < NL * NL; t += blockDim.x) {
        const int i = t / NL, j = t % NL;
        if (i == k) continue;
        const double f = M[i * NL + k];
        if (j == k) continue;
        M[t] -= f * M[k * NL + j];
      }
      __syncthreads();
      for (int i = tid; i < NL; i += blockDim.x)
        if (i != k) M[i * NL + k] = -M[i * NL + k] * M[k * NL + k];
      __syncthreads();
    }
    if (singular) {
      if (tid == 0) atomicAdd(status, 1);
      for (int t = tid; t < NL * NL; t += blockDim.x) { vinv[base + t] = 0.0; vinvf[base + t] = 0.f; }
      continue;
    }
    // undo the row interchanges as column interchanges, last to first
    for (int k = NL - 1; k >= 0; k--) {
      const int p = piv[k];
      if (p != k)
        for (int i = tid; i < NL; i += blockDim.x) {
          double t = M[i * NL + k]; M[i * NL + k] = M[i * NL + p]; M[i * NL + p] = t;
        }
      __syncthreads();
    }
    // transposed: entry (i, j) of the inverse at j * NL + i
    for (int t = tid; t < NL * NL; t += blockDim.x) {
      const int i = t / NL, j = t % NL;
      vinv[base + j * NL + i] = M[t];
      vinvf[base + j * NL + i] = (float)M[t];
    }
  }
}

// x[dof] += omega * w[dof] * (K_e^-1 r_e) for the elements of one colour; the
// inverse is stored transposed (vinv[e][b][a] = (K_e^-1)_ab) in the cycle's precision T
template <int NL, typename T>
__global__ __launch_bounds__(128) void k_vanka_apply(int nel, const int *__restrict__ elems, const int *__restrict__ eldof,
                                                     const T *__restrict__ vinv, const T *__restrict__ r,
                                                     T *__restrict__ x, const double *__restrict__ wt, double omega) {
  __shared__ T re[NL];
  __shared__ int dof[NL];
  const int tid = threadIdx.x;
  for (int ei = blockIdx.x; ei < nel; ei += gridDim.x) {
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
  if (!*store) {
    if (hipMalloc((void **)store, (size_t)l.nel * nl * nl * sizeof(double)) != hipSuccess) { fm_set_error("Vanka alloc"); return FM_ENOMEM; }
  }
  if (!*storef) {
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
  unsigned g = (unsigned)(l.nel < 8192 ? l.nel : 8192);
  if (s->mesh->dim == 2)
    hipLaunchKernelGGL(k_vanka_invert<22>, dim3(g), dim3(256), 0, st, l.nel, l.elpos, Av, *store, *storef, status);
  else
    hipLaunchKernelGGL(k_vanka_invert<89>, dim3(g), dim3(256), 0, st, l.nel, l.elpos, Av, *store, *storef, status);
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
template <typename T>
static int vanka_sweep(fm_system *s, int L, const CsrView<T> &A, const T *vinv, T *x, const T *b, T *r, int its,
                       double omega, int colored, bool xz) {
  Level &l = s->lev[L];
  hipStream_t st = s->ctx->stream;
  // additive: one residual per sweep, damped and weighted by 1/multiplicity;
  // coloured: residual refreshed before every colour, undamped weights 1
  const double *wt = l.vwt;
  bool fresh = xz;  // x still zero: the residual is b
  // colour c > = 0: only the rows that colour's elements touch (whole SELL slices, see Level::cslist)
  auto residual = [&](int c) -> int {
    if (fresh) {
      fresh = false;
      return hipMemcpyAsync(r, b, l.n * sizeof(T), hipMemcpyDeviceToDevice, st) == hipSuccess ? 0 : FM_EHIP;
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
    for (int c = 0; c < l.ncolor; c++) {
      int n = l.color_off[c + 1] - l.color_off[c];
      if (n <= 0) continue;
      if (colored) {
        int rc = residual(c);
        if (rc) return rc;
      }
      unsigned g = (unsigned)(n < 8192 ? n : 8192);
      const double *w = colored ? l.vone : wt;
      if (s->mesh->dim == 2)
        hipLaunchKernelGGL((k_vanka_apply<22, T>), dim3(g), dim3(128), 0, st, n, l.color_elems + l.color_off[c], l.eldofi,
                           vinv, r, x, w, omega);
      else
        hipLaunchKernelGGL((k_vanka_apply<89, T>), dim3(g), dim3(128), 0, st, n, l.color_elems + l.color_off[c], l.eldofi,
                           vinv, r, x, w, omega);
    }
  }
  return hipGetLastError() == hipSuccess ? 0 : FM_EHIP;
}

int fm_vanka_invert(fm_system *s, int L, const double *Av, double **store, float **storef) {
  return vanka_invert_level(s, L, Av, store, storef);
}
int fm_vanka_apply64(fm_system *s, int L, const CsrView<double> &A, const double *vinv, double *x, const double *b,
                     double *r, int its, double omega, int colored, bool xz) {
  return vanka_sweep<double>(s, L, A, vinv, x, b, r, its, omega, colored, xz);
}
int fm_vanka_apply32(fm_system *s, int L, const CsrView<float> &A, const float *vinv, float *x, const float *b,
                     float *r, int its, double omega, int colored, bool xz) {
  return vanka_sweep<float>(s, L, A, vinv, x, b, r, its, omega, colored, xz);
}
