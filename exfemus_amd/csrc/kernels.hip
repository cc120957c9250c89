// kernels.hip -- gfx950 kernels of the FEMuS hot path (wave64, CDNA4).
//
//  * CSR SpMV family (y = Ax, r = b - Ax, fused Chebyshev step, fused
//    prolongation-correction): LPR lanes of a 64-wide wavefront own one row,
//    partial sums are reduced with cross-lane shuffles.  Rows are
//    independent; values/columns stream once from HBM (the roofline bound).
//    Replaces LaspackVectorM::resid / Mul_QV (contrib/matrix/laspack_vectorM.C:147-189,
//    contrib/laspack/operats.c:782-1047).
//  * Element assembly of the heat (Q2) and Navier-Stokes (Q2-Q1) systems:
//    one workgroup per element, shape tables and global gradients in LDS,
//    elements processed colour by colour (2^dim colours: no two elements of
//    a colour share a node) so the scatter into CSR needs no atomics.
//    Restates MGSolT::GenMatRhs / MGSolNS::GenMatRhs
//    (template_appl/nse/SRC/MGSolverT.C:93-525, MGSolverNS.C:108-717).
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdlib>
#include <algorithm>
#include "fm_kernels.h"

#define WAVE 64

// ---------------------------------------------------------------------------
// SpMV core with epilogues
// ---------------------------------------------------------------------------
enum { EP_Y = 0, EP_RES = 1, EP_CHEB = 2, EP_ADD = 3, EP_JAC = 4, EP_RESD = 5 };

template <typename T>
struct EpArgs {
  T *y;          // EP_Y: y; EP_RES: r; EP_CHEB: r (in/out); EP_ADD: x (in/out); EP_JAC: x (in/out)
  const T *b;    // EP_RES / EP_JAC: b
  T *x;          // EP_CHEB: x (in/out)
  T *dnew;       // EP_CHEB: new direction
  const T *dold; // EP_CHEB: old direction (also the gathered vector)
  const T *dinv; // EP_CHEB / EP_JAC
  T c1, c2;      // EP_CHEB: dnew = c1*dold + c2*dinv*r ; EP_ADD: x += c1*(Px) ; EP_JAC: omega ; EP_RESD: d = c1*dinv*r
  const double *cond;  // EP_ADD: if non-null, apply only when cond[0] < cond[1] (reference rule)
  int xold;      // EP_CHEB: x += dold + dnew instead of x += dnew (dold not yet added to x)
};

template <typename T, int EP>
__device__ inline void apply_ep(int64_t row, T sum, const T *__restrict__ xg, const EpArgs<T> &ep) {
  if (EP == EP_Y) {
    ep.y[row] = sum;
  } else if (EP == EP_RES) {
    ep.y[row] = ep.b[row] - sum;
  } else if (EP == EP_CHEB) {
    T r = ep.y[row] - sum;
    ep.y[row] = r;
    const T dol = ep.dold[row];
    T dn = ep.c1 * dol + ep.c2 * ep.dinv[row] * r;
    ep.dnew[row] = dn;
    ep.x[row] += ep.xold ? dol + dn : dn;
  } else if (EP == EP_ADD) {
    if (ep.cond == nullptr || ep.cond[0] < ep.cond[1]) ep.y[row] += ep.c1 * sum;
  } else if (EP == EP_RESD) {
    // r = b - A x and the first smoothing direction d = c1 dinv r (x is the gathered vector: not written here)
    const T r = ep.b[row] - sum;
    ep.y[row] = r;
    ep.dnew[row] = ep.c1 * ep.dinv[row] * r;
  } else if (EP == EP_JAC) {
    // x_new = x + omega * dinv * (b - A x); written to dnew (x read via xg)
    ep.dnew[row] = xg[row] + ep.c1 * ep.dinv[row] * (ep.b[row] - sum);
  }
}

// CSR: LPR lanes per row, shuffle reduction
template <int LPR, typename T, int EP>
__global__ __launch_bounds__(256) void k_spmv(int64_t nrows, const int64_t *__restrict__ rp,
                                              const int *__restrict__ ci, const T *__restrict__ v,
                                              const T *__restrict__ xg, EpArgs<T> ep) {
  const int lane = threadIdx.x % LPR;
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / LPR;
  if (row >= nrows) return;
  const int64_t q0 = rp[row], q1 = rp[row + 1];
  T sum = 0;
  for (int64_t q = q0 + lane; q < q1; q += LPR) sum += v[q] * xg[ci[q]];
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, LPR);
  if (lane != 0) return;
  apply_ep<T, EP>(row, sum, xg, ep);
}

// SELL-64: one wavefront per slice, one lane per row, the entries of a row
// summed in increasing internal column order; matrix values and column
// offsets are streamed with non-temporal loads so that the gathered vector
// stays resident in L2 / Infinity Cache.  The slice index is made wave-uniform
// (readfirstlane) so that the slice header and the per-position column bases
// are scalar loads shared by the 64 lanes.  A slice decodes column
// base[k] + c8[...] (8-bit slices), base[k] + c16[slot] (16-bit), or reads
// absolute columns from c32 (32-bit).  With xcd_chunk > 0 the block
// index is remapped so that the blocks the hardware places on one XCD (block
// id modulo 8) walk one contiguous range of slices.
// matrix stream loads: non-temporal (NT) for matrices that are read once per pass and would only evict the
// gathered vector, ordinary for matrices small enough to stay in the Infinity Cache between the back-to-back
// products of one smoothing burst
template <bool NT, typename P>
__device__ __forceinline__ P ldm(const P *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

// one lane's row of a slice whose column offsets are BITS wide (8 or 16), packed 32 / BITS positions per
// 32-bit word: the word of group g (positions g * PER ..) of this lane is cw[64 g]
template <typename T, int BITS, int UNR, bool NT>
__device__ __forceinline__ T sell_row_packed(const T *__restrict__ vp, const unsigned *__restrict__ cw,
                                             const int *__restrict__ bp, int w, const T *__restrict__ xg) {
  constexpr int PER = 32 / BITS;
  constexpr unsigned MASK = (1u << BITS) - 1u;
  static_assert(UNR % PER == 0, "the unroll must cover whole words");
  T sum = 0;
  int k = 0;
  for (; k + UNR <= w; k += UNR) {
    T v[UNR];
    unsigned p[UNR / PER];
#pragma unroll
    for (int u = 0; u < UNR; u++) v[u] = ldm<NT>(vp + (k + u) * 64);
#pragma unroll
    for (int g = 0; g < UNR / PER; g++) p[g] = ldm<NT>(cw + (k / PER + g) * 64);
    int c[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) c[u] = bp[k + u] + (int)((p[u / PER] >> (BITS * (u % PER))) & MASK);
    T xv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) xv[u] = xg[c[u]];
#pragma unroll
    for (int u = 0; u < UNR; u++) sum += v[u] * xv[u];
  }
  // the last w mod UNR positions (Q2 rows hold 27, 45, 75 or 125 entries) as one batch under wave-uniform
  // guards, so that their loads are in flight together instead of one dependent chain per position; the
  // terms are added in the same order as above
  const int rem = w - k;
  if (rem > 0) {
    T v[UNR], xv[UNR];
    unsigned p[UNR / PER];
#pragma unroll
    for (int u = 0; u < UNR; u++)
      if (u < rem) v[u] = ldm<NT>(vp + (k + u) * 64);
#pragma unroll
    for (int g = 0; g < UNR / PER; g++)
      if (g * PER < rem) p[g] = ldm<NT>(cw + (k / PER + g) * 64);
#pragma unroll
    for (int u = 0; u < UNR; u++)
      if (u < rem) xv[u] = xg[bp[k + u] + (int)((p[u / PER] >> (BITS * (u % PER))) & MASK)];
#pragma unroll
    for (int u = 0; u < UNR; u++)
      if (u < rem) sum += v[u] * xv[u];
  }
  return sum;
}

// one lane's row of a lane-predicted slice over the positions [k0, k1): the column of position k is
// bp[k] + lane + a 2-bit residual, bits 2 (k % 16) of the lane's word cw[64 (k / 16)].  A batch's columns need
// only that word and the scalar bases, so the word of the next 16 positions is loaded one chunk ahead and a
// batch issues its value loads and its gathers together (one HBM round trip per batch, not HBM then L2).
template <typename T, int UNR, bool NT>
__device__ __forceinline__ T sell_row_pred(const T *__restrict__ vp, const unsigned *__restrict__ cw,
                                           const int *__restrict__ bp, int lane, int k0, int k1,
                                           const T *__restrict__ xg) {
  static_assert(16 % UNR == 0, "a batch lies inside one residual word");
  T sum = 0;
  if (k0 >= k1) return sum;
  unsigned p = ldm<NT>(cw + (k0 >> 4) * 64);
  int k = k0;
  while (k < k1) {
    const int kend = min(k1, ((k >> 4) + 1) << 4);   // end of this word's positions
    unsigned pn = 0;
    if (kend < k1) pn = ldm<NT>(cw + (kend >> 4) * 64);
    for (; k + UNR <= kend; k += UNR) {
      T v[UNR], xv[UNR];
#pragma unroll
      for (int u = 0; u < UNR; u++) v[u] = ldm<NT>(vp + (k + u) * 64);
#pragma unroll
      for (int u = 0; u < UNR; u++) xv[u] = xg[bp[k + u] + lane + (int)((p >> (2 * ((k + u) & 15))) & 3u)];
#pragma unroll
      for (int u = 0; u < UNR; u++) sum += v[u] * xv[u];
    }
    const int rem = kend - k;  // the word's last positions as one guarded batch
    if (rem > 0) {
      T v[UNR], xv[UNR];
#pragma unroll
      for (int u = 0; u < UNR; u++)
        if (u < rem) v[u] = ldm<NT>(vp + (k + u) * 64);
#pragma unroll
      for (int u = 0; u < UNR; u++)
        if (u < rem) xv[u] = xg[bp[k + u] + lane + (int)((p >> (2 * ((k + u) & 15))) & 3u)];
#pragma unroll
      for (int u = 0; u < UNR; u++)
        if (u < rem) sum += v[u] * xv[u];
      k = kend;
    }
    p = pn;
  }
  return sum;
}

// sum over the positions [k0, k1) of one lane's row of slice s (k0 a multiple of UNR, so that it is also a
// multiple of the 4 or 2 positions packed per 32-bit word of 8- or 16-bit offsets)
template <typename T, int UNR, bool NT>
__device__ __forceinline__ T sell_range(int s, int lane, int k0, int k1, const int64_t *__restrict__ sl_off,
                                        const int64_t *__restrict__ sl_c32, const unsigned short *__restrict__ c16,
                                        const int *__restrict__ base, const int *__restrict__ c32,
                                        const int64_t *__restrict__ sl_c8, const unsigned char *__restrict__ c8,
                                        const int64_t *__restrict__ sl_c16, const int64_t *__restrict__ sl_c2,
                                        const unsigned *__restrict__ c2, const T *__restrict__ vals,
                                        const T *__restrict__ xg) {
  const int64_t off = sl_off[s];
  const int64_t o2 = sl_c2[s];
  if (o2 >= 0) return sell_row_pred<T, UNR, NT>(vals + off + lane, c2 + o2 + lane, base + (off >> 6), lane, k0, k1, xg);
  const int64_t o32 = sl_c32[s];
  const int64_t o8 = sl_c8[s], o16 = sl_c16[s];
  const T *vp = vals + off + lane + (int64_t)k0 * 64;
  const int *bp = base + (off >> 6) + k0;
  const int w = k1 - k0;
  if (w <= 0) return T(0);
  if (o8 >= 0)
    return sell_row_packed<T, 8, UNR, NT>(vp, reinterpret_cast<const unsigned *>(c8 + o8) + lane + (k0 / 4) * 64, bp, w, xg);
  if (o16 >= 0)
    return sell_row_packed<T, 16, UNR, NT>(vp, reinterpret_cast<const unsigned *>(c16 + o16) + lane + (k0 / 2) * 64, bp,
                                       w, xg);
  const int *cp = c32 + o32 + lane + (int64_t)k0 * 64;
  T sum = 0;
  int k = 0;
  for (; k + UNR <= w; k += UNR) {
    T v[UNR];
    int c[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) v[u] = ldm<NT>(vp + (k + u) * 64);
#pragma unroll
    for (int u = 0; u < UNR; u++) c[u] = ldm<NT>(cp + (k + u) * 64);
    T xv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) xv[u] = xg[c[u]];
#pragma unroll
    for (int u = 0; u < UNR; u++) sum += v[u] * xv[u];
  }
  const int rem = w - k;  // the remainder as one guarded batch, as in sell_row_packed
  if (rem > 0) {
    T v[UNR], xv[UNR];
    int c[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++)
      if (u < rem) {
        v[u] = ldm<NT>(vp + (k + u) * 64);
        c[u] = ldm<NT>(cp + (k + u) * 64);
      }
#pragma unroll
    for (int u = 0; u < UNR; u++)
      if (u < rem) xv[u] = xg[c[u]];
#pragma unroll
    for (int u = 0; u < UNR; u++)
      if (u < rem) sum += v[u] * xv[u];
  }
  return sum;
}

template <typename T, int EP, int UNR>
__global__ __launch_bounds__(256) void k_sell(int nsl, const int64_t *__restrict__ sl_off,
                                              const int64_t *__restrict__ sl_c32, const int *__restrict__ sl_w,
                                              const int *__restrict__ perm, const unsigned short *__restrict__ c16,
                                              const int *__restrict__ base, const int *__restrict__ c32,
                                              const int64_t *__restrict__ sl_c8, const unsigned char *__restrict__ c8,
                                              const int64_t *__restrict__ sl_c16, const int64_t *__restrict__ sl_c2,
                                              const unsigned *__restrict__ c2,
                                              const T *__restrict__ vals, const T *__restrict__ xg, EpArgs<T> ep,
                                              int xcd_chunk, const int *__restrict__ slist) {
  int blk = blockIdx.x;
  if (xcd_chunk > 0) blk = (int)(blockIdx.x % 8) * xcd_chunk + (int)(blockIdx.x / 8);
  const int w_idx = __builtin_amdgcn_readfirstlane(blk * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6));
  if (w_idx >= nsl) return;  // nsl: number of slices, or of list entries when slist is given
  const int s = slist ? __builtin_amdgcn_readfirstlane(slist[w_idx]) : w_idx;
  const int lane = threadIdx.x & 63;
  const int row = perm[(int64_t)s * 64 + lane];
  const T sum = sell_range<T, UNR, true>(s, lane, 0, sl_w[s], sl_off, sl_c32, c16, base, c32, sl_c8, c8, sl_c16, sl_c2, c2,
                                         vals, xg);
  if (row < 0) return;
  apply_ep<T, EP>(row, sum, xg, ep);
}

// The same product for matrices with few slices (the coarser multigrid levels), where one wave per slice
// leaves the chip mostly idle and a launch costs the latency of one wave walking its rows' ~100 positions
// in sequence: S waves (one block of 64 S threads) share a slice, wave j summing the positions
// [j c, (j + 1) c) with c = ceil(w / S) rounded up to the unroll; the partial sums meet in LDS and wave 0
// adds them in wave order and applies the epilogue.
template <typename T, int EP, int UNR, int S, bool NT>
__global__ __launch_bounds__(64 * S) void k_sell_split(int nsl, const int64_t *__restrict__ sl_off,
                                                       const int64_t *__restrict__ sl_c32, const int *__restrict__ sl_w,
                                                       const int *__restrict__ perm,
                                                       const unsigned short *__restrict__ c16,
                                                       const int *__restrict__ base, const int *__restrict__ c32,
                                                       const int64_t *__restrict__ sl_c8,
                                                       const unsigned char *__restrict__ c8,
                                                       const int64_t *__restrict__ sl_c16,
                                                       const int64_t *__restrict__ sl_c2, const unsigned *__restrict__ c2,
                                                       const T *__restrict__ vals, const T *__restrict__ xg,
                                                       EpArgs<T> ep, const int *__restrict__ slist) {
  __shared__ T part[S - 1][64];
  const int w_idx = blockIdx.x;
  if (w_idx >= nsl) return;  // whole block
  const int s = slist ? __builtin_amdgcn_readfirstlane(slist[w_idx]) : w_idx;
  const int lane = threadIdx.x & 63;
  const int j = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int w = sl_w[s];
  const int c = ((w + S - 1) / S + UNR - 1) / UNR * UNR;
  const int k0 = min(w, j * c), k1 = min(w, k0 + c);
  T sum = sell_range<T, UNR, NT>(s, lane, k0, k1, sl_off, sl_c32, c16, base, c32, sl_c8, c8, sl_c16, sl_c2, c2, vals,
                                   xg);
  if (j > 0) part[j - 1][lane] = sum;
  __syncthreads();
  if (j > 0) return;
#pragma unroll
  for (int q = 0; q < S - 1; q++) sum += part[q][lane];
  const int row = perm[(int64_t)s * 64 + lane];
  if (row < 0) return;
  apply_ep<T, EP>(row, sum, xg, ep);
}

// launch-shape knobs of the SELL kernel (read once from the environment):
// FEMUS_SELL_UNROLL = 4 | 8 | 16 (default 8; the split kernel runs 16 when it is 16, else 8),
// FEMUS_SELL_XCD = 0 | 1 (default 1),
// FEMUS_SELL_SPLIT = W (default 16384; 0 = off): a product over n slices with n < W runs k_sell_split with
// S = the smallest of 2, 4, 8 for which n S >= W (8 at most).
// Measured on MI355X, 3D heat hex27 64^3 top matrix (135M nnz) in the
// node-lexicographic internal order: contiguous per-XCD slice ranges take the
// FP64 SpMV from 389 to 301 us and the FP32 one from 249 to 187 us (each XCD's
// L2 then holds the x window of its own range); unroll 8 beats 4 by ~2%.
// FEMUS_SELL_LIST_SPLIT = 0 | 1 (default 1): products restricted to a slice list (the residual of one Vanka
// colour) with W to 4 W entries run with S = 2.  Measured: the C5 step at 48^3 (top-level colour lists of
// ~19,600 slices) 627 -> 611 ms mixed, 985 -> 941 ms FP64; at 96^3 (~154,000 slices) no gain.
// FEMUS_SELL_TEMPORAL_MB = M (default 200; 0 = never): split-slice products over a matrix whose values and
// column offsets take at most M MB read it with ordinary loads, so that it stays in the 256 MB Infinity Cache
// across the back-to-back products of a smoothing burst; larger matrices stream with non-temporal loads.
static int g_sell_unroll = -1, g_sell_xcd = -1, g_sell_split = -1, g_sell_list2 = -1;
static int64_t g_sell_temporal = -1;
static void sell_knobs() {
  if (g_sell_unroll >= 0) return;
  const char *u = getenv("FEMUS_SELL_UNROLL");
  const char *x = getenv("FEMUS_SELL_XCD");
  const char *p = getenv("FEMUS_SELL_SPLIT");
  const char *q = getenv("FEMUS_SELL_LIST_SPLIT");
  g_sell_unroll = (u && atoi(u) == 4) ? 4 : ((u && atoi(u) == 16) ? 16 : 8);
  g_sell_xcd = (x && atoi(x) == 0) ? 0 : 1;
  if (g_sell_split < 0) g_sell_split = p ? std::max(0, atoi(p)) : 16384;
  g_sell_list2 = (q && atoi(q) == 0) ? 0 : 1;
  const char *tm = getenv("FEMUS_SELL_TEMPORAL_MB");
  g_sell_temporal = (int64_t)(tm ? std::max(0, atoi(tm)) : 200) << 20;
}

template <typename T, int EP, int S>
static void launch_split(hipStream_t s, const CsrView<T> &A, const T *x, const EpArgs<T> &ep, int nsl) {
  // ordinary loads for a matrix whose values and offsets fit the cache budget (FEMUS_SELL_TEMPORAL_MB)
  const bool nt = A.nslot * (int64_t)(sizeof(T) + 1) > g_sell_temporal;
  if (g_sell_unroll == 16)
    hipLaunchKernelGGL((k_sell_split<T, EP, 16, S, true>), dim3((unsigned)nsl), dim3(64 * S), 0, s, nsl, A.sl_off,
                       A.sl_c32, A.sl_w, A.perm, A.c16, A.base, A.c32, A.sl_c8, A.c8, A.sl_c16, A.sl_c2, A.c2, A.sv, x, ep,
                       A.slist);
  else if (nt)
    hipLaunchKernelGGL((k_sell_split<T, EP, 8, S, true>), dim3((unsigned)nsl), dim3(64 * S), 0, s, nsl, A.sl_off,
                       A.sl_c32, A.sl_w, A.perm, A.c16, A.base, A.c32, A.sl_c8, A.c8, A.sl_c16, A.sl_c2, A.c2, A.sv, x, ep,
                       A.slist);
  else
    hipLaunchKernelGGL((k_sell_split<T, EP, 8, S, false>), dim3((unsigned)nsl), dim3(64 * S), 0, s, nsl, A.sl_off,
                       A.sl_c32, A.sl_w, A.perm, A.c16, A.base, A.c32, A.sl_c8, A.c8, A.sl_c16, A.sl_c2, A.c2, A.sv, x, ep,
                       A.slist);
}

template <typename T, int EP>
static int launch_spmv(hipStream_t s, const CsrView<T> &A, const T *x, const EpArgs<T> &ep) {
  if (A.nr <= 0) return 0;
  const int block = 256;
  if (A.nsl > 0) {
    sell_knobs();
    const int64_t nwork = A.slist ? A.nlist : A.nsl;
    if (nwork <= 0) return 0;
    const int nsl = (int)nwork;
    if (nwork < g_sell_split) {
      const int S = nwork * 2 >= g_sell_split ? 2 : (nwork * 4 >= g_sell_split ? 4 : 8);
      if (S == 2) launch_split<T, EP, 2>(s, A, x, ep, nsl);
      else if (S == 4) launch_split<T, EP, 4>(s, A, x, ep, nsl);
      else launch_split<T, EP, 8>(s, A, x, ep, nsl);
      return hipGetLastError() == hipSuccess ? 0 : -2;
    }
    if (A.slist && g_sell_list2 && nwork < 4 * (int64_t)g_sell_split) {
      launch_split<T, EP, 2>(s, A, x, ep, nsl);
      return hipGetLastError() == hipSuccess ? 0 : -2;
    }
    int64_t nb = (nwork + 3) / 4;
    int64_t chunk = 0;
    if (g_sell_xcd && nb >= 64) {
      chunk = (nb + 7) / 8;
      nb = chunk * 8;  // padded grid: surplus blocks map past the last slice and exit
    }
    dim3 grid((unsigned)nb);
    if (g_sell_unroll == 8)
      hipLaunchKernelGGL((k_sell<T, EP, 8>), grid, dim3(block), 0, s, nsl, A.sl_off, A.sl_c32, A.sl_w, A.perm, A.c16,
                         A.base, A.c32, A.sl_c8, A.c8, A.sl_c16, A.sl_c2, A.c2, A.sv, x, ep, (int)chunk, A.slist);
    else if (g_sell_unroll == 16)
      hipLaunchKernelGGL((k_sell<T, EP, 16>), grid, dim3(block), 0, s, nsl, A.sl_off, A.sl_c32, A.sl_w, A.perm, A.c16,
                         A.base, A.c32, A.sl_c8, A.c8, A.sl_c16, A.sl_c2, A.c2, A.sv, x, ep, (int)chunk, A.slist);
    else
      hipLaunchKernelGGL((k_sell<T, EP, 4>), grid, dim3(block), 0, s, nsl, A.sl_off, A.sl_c32, A.sl_w, A.perm, A.c16,
                         A.base, A.c32, A.sl_c8, A.c8, A.sl_c16, A.sl_c2, A.c2, A.sv, x, ep, (int)chunk, A.slist);
    return hipGetLastError() == hipSuccess ? 0 : -2;
  }
  int64_t threads = A.nr * A.lpr;
  dim3 grid((unsigned)((threads + block - 1) / block));
  int64_t n = A.nr;
  const int64_t *rp = A.rp;
  const int *ci = A.ci;
  const T *v = A.v;
  switch (A.lpr) {
    case 4: hipLaunchKernelGGL((k_spmv<4, T, EP>), grid, dim3(block), 0, s, n, rp, ci, v, x, ep); break;
    case 8: hipLaunchKernelGGL((k_spmv<8, T, EP>), grid, dim3(block), 0, s, n, rp, ci, v, x, ep); break;
    case 16: hipLaunchKernelGGL((k_spmv<16, T, EP>), grid, dim3(block), 0, s, n, rp, ci, v, x, ep); break;
    case 32: hipLaunchKernelGGL((k_spmv<32, T, EP>), grid, dim3(block), 0, s, n, rp, ci, v, x, ep); break;
    default: hipLaunchKernelGGL((k_spmv<64, T, EP>), grid, dim3(block), 0, s, n, rp, ci, v, x, ep); break;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

namespace fmk {

int sell_split(int W) {
  sell_knobs();
  const int old = g_sell_split;
  if (W >= 0) g_sell_split = W;
  return old;
}

int lanes_for(int64_t nnz, int64_t n) {
  double avg = n > 0 ? (double)nnz / (double)n : 1.0;
  if (avg <= 6) return 4;
  if (avg <= 14) return 8;
  if (avg <= 40) return 16;
  if (avg <= 100) return 32;
  return 64;
}

template <typename T> static int spmv_t(hipStream_t s, const CsrView<T> &A, const T *x, T *y) {
  EpArgs<T> ep{}; ep.y = y;
  return launch_spmv<T, EP_Y>(s, A, x, ep);
}
template <typename T> static int resid_t(hipStream_t s, const CsrView<T> &A, const T *b, const T *x, T *r) {
  EpArgs<T> ep{}; ep.y = r; ep.b = b;
  return launch_spmv<T, EP_RES>(s, A, x, ep);
}
template <typename T> static int cheb_t(hipStream_t s, const CsrView<T> &A, T *r, T *x, const T *dold, T *dnew,
                                        const T *dinv, T c1, T c2, int xold) {
  EpArgs<T> ep{}; ep.y = r; ep.x = x; ep.dold = dold; ep.dnew = dnew; ep.dinv = dinv; ep.c1 = c1; ep.c2 = c2;
  ep.xold = xold;
  return launch_spmv<T, EP_CHEB>(s, A, dold, ep);
}
template <typename T> static int resid_d_t(hipStream_t s, const CsrView<T> &A, const T *b, const T *x, T *r, T *d,
                                           const T *dinv, T a) {
  EpArgs<T> ep{}; ep.y = r; ep.b = b; ep.dnew = d; ep.dinv = dinv; ep.c1 = a;
  return launch_spmv<T, EP_RESD>(s, A, x, ep);
}
template <typename T> static int addmul_t(hipStream_t s, const CsrView<T> &P, const T *xc, T *x, T alpha,
                                          const double *cond) {
  EpArgs<T> ep{}; ep.y = x; ep.c1 = alpha; ep.cond = cond;
  return launch_spmv<T, EP_ADD>(s, P, xc, ep);
}
template <typename T> static int jacobi_t(hipStream_t s, const CsrView<T> &A, const T *b, const T *x, T *xn,
                                          const T *dinv, T omega) {
  EpArgs<T> ep{}; ep.b = b; ep.dnew = xn; ep.dinv = dinv; ep.c1 = omega;
  return launch_spmv<T, EP_JAC>(s, A, x, ep);
}

int spmv(hipStream_t s, const CsrView<double> &A, const double *x, double *y) { return spmv_t(s, A, x, y); }
int spmv(hipStream_t s, const CsrView<float> &A, const float *x, float *y) { return spmv_t(s, A, x, y); }
int resid(hipStream_t s, const CsrView<double> &A, const double *b, const double *x, double *r) { return resid_t(s, A, b, x, r); }
int resid(hipStream_t s, const CsrView<float> &A, const float *b, const float *x, float *r) { return resid_t(s, A, b, x, r); }
int cheb_step(hipStream_t s, const CsrView<double> &A, double *r, double *x, const double *dold, double *dnew,
              const double *dinv, double c1, double c2, int xold) {
  return cheb_t(s, A, r, x, dold, dnew, dinv, c1, c2, xold);
}
int cheb_step(hipStream_t s, const CsrView<float> &A, float *r, float *x, const float *dold, float *dnew,
              const float *dinv, float c1, float c2, int xold) {
  return cheb_t(s, A, r, x, dold, dnew, dinv, c1, c2, xold);
}
int resid_d(hipStream_t s, const CsrView<double> &A, const double *b, const double *x, double *r, double *d,
            const double *dinv, double a) { return resid_d_t(s, A, b, x, r, d, dinv, a); }
int resid_d(hipStream_t s, const CsrView<float> &A, const float *b, const float *x, float *r, float *d,
            const float *dinv, float a) { return resid_d_t(s, A, b, x, r, d, dinv, a); }
int addmul(hipStream_t s, const CsrView<double> &P, const double *xc, double *x, double alpha, const double *cond) {
  return addmul_t(s, P, xc, x, alpha, cond);
}
int addmul(hipStream_t s, const CsrView<float> &P, const float *xc, float *x, float alpha, const double *cond) {
  return addmul_t(s, P, xc, x, alpha, cond);
}
int jacobi_step(hipStream_t s, const CsrView<double> &A, const double *b, const double *x, double *xn,
                const double *dinv, double omega) { return jacobi_t(s, A, b, x, xn, dinv, omega); }
int jacobi_step(hipStream_t s, const CsrView<float> &A, const float *b, const float *x, float *xn,
                const float *dinv, float omega) { return jacobi_t(s, A, b, x, xn, dinv, omega); }

// ---------------------------------------------------------------------------
// vector kernels
// ---------------------------------------------------------------------------
template <typename T> __global__ void k_fill(int64_t n, T *x, T a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) x[i] = a;
}
// y = x (a kernel, not a copy command: a captured V-cycle graph then holds kernel nodes only)
template <typename T> __global__ void k_copy(int64_t n, const T *x, T *y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i] = x[i];
}
// x = a*x + b*y  (both same type)
template <typename T> __global__ void k_axpby(int64_t n, T a, T *x, T b, const T *y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = a * x[i] + b * y[i];
}
// y = a * dinv .* x / theta   (first Chebyshev direction; also x += y)
template <typename T> __global__ void k_cheb_first(int64_t n, const T *dinv, const T *r, T *d, T *x, T a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    T di = a * dinv[i] * r[i];
    d[i] = di;
    x[i] += di;
  }
}
// the first Chebyshev sweep from a zero iterate: r = b, d = a * dinv .* b, x = d
template <typename T> __global__ void k_cheb_start(int64_t n, const T *dinv, const T *b, T *r, T *d, T *x, T a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T bi = b[i];
    const T di = a * dinv[i] * bi;
    r[i] = bi;
    d[i] = di;
    x[i] = di;
  }
}
// y64 = (double) x32
__global__ void k_f32_to_f64(int64_t n, const float *a, double *b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) b[i] = (double)a[i];
}
__global__ void k_f64_to_f32(int64_t n, const double *a, float *b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) b[i] = (float)a[i];
}
// x64 += (double) e32
__global__ void k_add_f32(int64_t n, double *x, const float *e) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) x[i] += (double)e[i];
}
// scaled conversion: b32 = (float)(s * r64) with s = 1 / scale[0]
__global__ void k_f64_to_f32_scaled(int64_t n, const double *a, float *b, const double *scale) {
  double sinv = scale[0] > 0. ? 1.0 / scale[0] : 1.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) b[i] = (float)(a[i] * sinv);
}
// x64 += scale[0] * e32
__global__ void k_add_f32_scaled(int64_t n, double *x, const float *e, const double *scale) {
  double s = scale[0] > 0. ? scale[0] : 1.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) x[i] += s * (double)e[i];
}

static inline dim3 vgrid(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}
template <typename T> int fill(hipStream_t s, int64_t n, T *x, T a) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_fill<T>, vgrid(n), dim3(256), 0, s, n, x, a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
template int fill<double>(hipStream_t, int64_t, double *, double);
template int fill<float>(hipStream_t, int64_t, float *, float);
template <typename T> int copy(hipStream_t s, int64_t n, const T *x, T *y) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_copy<T>, vgrid(n), dim3(256), 0, s, n, x, y);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
template int copy<double>(hipStream_t, int64_t, const double *, double *);
template int copy<float>(hipStream_t, int64_t, const float *, float *);
template <typename T> int axpby(hipStream_t s, int64_t n, T a, T *x, T b, const T *y) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_axpby<T>, vgrid(n), dim3(256), 0, s, n, a, x, b, y);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
template int axpby<double>(hipStream_t, int64_t, double, double *, double, const double *);
template int axpby<float>(hipStream_t, int64_t, float, float *, float, const float *);
template <typename T> int cheb_first(hipStream_t s, int64_t n, const T *dinv, const T *r, T *d, T *x, T a) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_cheb_first<T>, vgrid(n), dim3(256), 0, s, n, dinv, r, d, x, a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
template int cheb_first<double>(hipStream_t, int64_t, const double *, const double *, double *, double *, double);
template int cheb_first<float>(hipStream_t, int64_t, const float *, const float *, float *, float *, float);
template <typename T> int cheb_start(hipStream_t s, int64_t n, const T *dinv, const T *b, T *r, T *d, T *x, T a) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_cheb_start<T>, vgrid(n), dim3(256), 0, s, n, dinv, b, r, d, x, a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
template int cheb_start<double>(hipStream_t, int64_t, const double *, const double *, double *, double *, double *, double);
template int cheb_start<float>(hipStream_t, int64_t, const float *, const float *, float *, float *, float *, float);
int from_f32(hipStream_t s, int64_t n, const float *a, double *b) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_f32_to_f64, vgrid(n), dim3(256), 0, s, n, a, b);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
int to_f32(hipStream_t s, int64_t n, const double *a, float *b, const double *scale) {
  if (n <= 0) return 0;
  if (scale) hipLaunchKernelGGL(k_f64_to_f32_scaled, vgrid(n), dim3(256), 0, s, n, a, b, scale);
  else hipLaunchKernelGGL(k_f64_to_f32, vgrid(n), dim3(256), 0, s, n, a, b);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
int add_f32(hipStream_t s, int64_t n, double *x, const float *e, const double *scale) {
  if (n <= 0) return 0;
  if (scale) hipLaunchKernelGGL(k_add_f32_scaled, vgrid(n), dim3(256), 0, s, n, x, e, scale);
  else hipLaunchKernelGGL(k_add_f32, vgrid(n), dim3(256), 0, s, n, x, e);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// ---- reductions: partial sums per block into scratch, finished by one block
template <typename T> __global__ __launch_bounds__(256) void k_dot_partial(int64_t n, const T *x, const T *y, double *part) {
  double s = 0.;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += (double)x[i] * (double)y[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  __shared__ double ws[4];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}
__global__ __launch_bounds__(256) void k_sum_final(int nb, const double *part, double *out, int do_sqrt) {
  double s = 0.;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += part[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  __shared__ double ws[4];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) { double t = ws[0] + ws[1] + ws[2] + ws[3]; out[0] = do_sqrt ? sqrt(t) : t; }
}
template <typename T> int dot_dev(hipStream_t s, int64_t n, const T *x, const T *y, double *part, double *out, int do_sqrt) {
  int nb = 1024;
  hipLaunchKernelGGL(k_dot_partial<T>, dim3(nb), dim3(256), 0, s, n, x, y, part);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s, nb, part, out, do_sqrt);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
template int dot_dev<double>(hipStream_t, int64_t, const double *, const double *, double *, double *, int);
template int dot_dev<float>(hipStream_t, int64_t, const float *, const float *, double *, double *, int);

// scalar helpers on device: out = a[0] / b[0] (relative norms for the reference rule)
__global__ void k_ratio(const double *a, const double *b, double *out) {
  if (threadIdx.x == 0) out[0] = b[0] != 0. ? a[0] / b[0] : 1.0;
}
int ratio(hipStream_t s, const double *a, const double *b, double *out) {
  hipLaunchKernelGGL(k_ratio, dim3(1), dim3(64), 0, s, a, b, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// ---- dense coarse solve: y = Ainv * b (row-major n x n); one wave per row
template <typename T> __global__ __launch_bounds__(256) void k_dense_mv(int n, const T *__restrict__ M, const T *__restrict__ b, T *__restrict__ y) {
  int row = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x % WAVE;
  if (row >= n) return;
  T s = 0;
  for (int j = lane; j < n; j += WAVE) s += M[(int64_t)row * n + j] * b[j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) y[row] = s;
}
template <typename T> int dense_mv(hipStream_t s, int n, const T *M, const T *b, T *y) {
  int64_t threads = (int64_t)n * WAVE;
  hipLaunchKernelGGL(k_dense_mv<T>, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, n, M, b, y);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
template int dense_mv<double>(hipStream_t, int, const double *, const double *, double *);
template int dense_mv<float>(hipStream_t, int, const float *, const float *, float *);

// ---- diagonal inverse
// dinv[i] = 1 / A_ii read at the diagonal's SELL slot dslot[i] (internal row i), FP64 and FP32
__global__ void k_dinv_slot(int64_t n, const int64_t *__restrict__ dslot, const double *__restrict__ sv, double *d64,
                            float *d32) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = dslot[i];
    const double d = q >= 0 ? sv[q] : 0.0;
    const double r = d != 0. ? 1.0 / d : 0.0;
    d64[i] = r;
    d32[i] = (float)r;
  }
}
int dinv_slot(hipStream_t s, int64_t n, const int64_t *dslot, const double *sv, double *d64, float *d32) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_dinv_slot, vgrid(n), dim3(256), 0, s, n, dslot, sv, d64, d32);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// ---- SELL <-> CSR entry correspondence: slot (s, k, l) of internal row i = perm[64 s + l] holds CSR entry
// rbase[i] + pos[slot] (pos = 0xFFFF: padding).  One wave per slice, four slices per block.
#define SELL_WALK(M, BODY)                                                                       \
  const int64_t s_ = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);                              \
  if (s_ >= (M).nsl) return;                                                                     \
  const int l_ = threadIdx.x & 63;                                                               \
  const int row = (M).perm[s_ * 64 + l_];                                                        \
  if (row < 0) return;                                                                           \
  const int w_ = (M).sl_w[s_];                                                                   \
  const int64_t rb = (M).rbase[row];                                                             \
  for (int k = 0; k < w_; k++) {                                                                 \
    const int64_t slot = (M).sl_off[s_] + (int64_t)k * 64 + l_;                                  \
    const unsigned t = (M).pos[slot];                                                            \
    if (t == 0xFFFFu) continue;                                                                  \
    BODY                                                                                         \
  }

// CSR values from the SELL copy: cv[rbase + pos] = sv[slot]
__global__ void k_sell2csr(SellMap m, const double *__restrict__ sv, double *cv) {
  SELL_WALK(m, cv[rb + t] = sv[slot];)
}
int sell2csr(hipStream_t s, const SellMap &m, const double *sv, double *cv) {
  if (m.nsl <= 0) return 0;
  hipLaunchKernelGGL(k_sell2csr, dim3((unsigned)((m.nsl + 3) / 4)), dim3(256), 0, s, m, sv, cv);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
// SELL values (FP64 and FP32) from the CSR values; padding slots are zero
__global__ void k_csr2sell(SellMap m, const double *__restrict__ cv, double *__restrict__ sv, float *__restrict__ svf) {
  const int64_t s_ = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s_ >= m.nsl) return;
  const int l_ = threadIdx.x & 63;
  const int row = m.perm[s_ * 64 + l_];
  const int w_ = m.sl_w[s_];
  const int64_t rb = row >= 0 ? m.rbase[row] : 0;
  for (int k = 0; k < w_; k++) {
    const int64_t slot = m.sl_off[s_] + (int64_t)k * 64 + l_;
    const unsigned t = m.pos[slot];
    const double v = (row >= 0 && t != 0xFFFFu) ? cv[rb + t] : 0.0;
    sv[slot] = v;
    svf[slot] = (float)v;
  }
}
int csr2sell(hipStream_t s, const SellMap &m, const double *csr_v, double *sv, float *svf) {
  if (m.nsl <= 0) return 0;
  hipLaunchKernelGGL(k_csr2sell, dim3((unsigned)((m.nsl + 3) / 4)), dim3(256), 0, s, m, csr_v, sv, svf);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
// kof[rbase + pos] = k: the SELL position of every CSR entry (entries of a row are at most 65,535)
__global__ void k_sell_kof(SellMap m, unsigned short *kof) {
  SELL_WALK(m, kof[rb + t] = (unsigned short)k;)
}
int sell_kof(hipStream_t s, const SellMap &m, unsigned short *kof) {
  if (m.nsl <= 0) return 0;
  hipLaunchKernelGGL(k_sell_kof, dim3((unsigned)((m.nsl + 3) / 4)), dim3(256), 0, s, m, kof);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
#undef SELL_WALK
// dslot[i] = SELL slot of the diagonal entry of internal row i (CSR row ord[i], columns sorted), -1 if absent
__global__ void k_diag_slot(int64_t n, const int64_t *__restrict__ rp, const int *__restrict__ ci,
                            const int *__restrict__ ord, const int64_t *__restrict__ rslot,
                            const unsigned short *__restrict__ kof, int64_t *dslot) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int r = ord[i];
    int64_t lo = rp[r], hi = rp[r + 1] - 1, p = -1;
    while (lo <= hi) {
      const int64_t mid = (lo + hi) >> 1;
      const int c = ci[mid];
      if (c == r) { p = rslot[i] + 64 * (int64_t)kof[mid]; break; }
      if (c < r) lo = mid + 1; else hi = mid - 1;
    }
    dslot[i] = p;
  }
}
int diag_slot(hipStream_t s, int64_t n, const int64_t *rp, const int *ci, const int *ord, const int64_t *rslot,
              const unsigned short *kof, int64_t *dslot) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_diag_slot, vgrid(n), dim3(256), 0, s, n, rp, ci, ord, rslot, kof, dslot);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
// elslot[t] = rslot[eldofi[t]]: SELL slot of position 0 of the row of each element DoF (-1: no DoF)
__global__ void k_elslot(int64_t n, const int *__restrict__ eldofi, const int64_t *__restrict__ rslot, int64_t *elslot) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int d = eldofi[i];
    elslot[i] = d >= 0 ? rslot[d] : -1;
  }
}
int elslot(hipStream_t s, int64_t n, const int *eldofi, const int64_t *rslot, int64_t *out) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_elslot, vgrid(n), dim3(256), 0, s, n, eldofi, rslot, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
// nodal field from a level vector in internal order: out[v * N + k] = x[rank[dof[k + v * N]]] for the nvar
// variables v (0 where the node carries no such DoF)
__global__ void k_node_gather(int64_t N, int nvar, const int *__restrict__ dof, const int *__restrict__ rank,
                              const double *__restrict__ x, double *__restrict__ out) {
  const int64_t total = N * nvar;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int d = dof[t];
    out[t] = d >= 0 ? x[rank[d]] : 0.0;
  }
}
int node_gather(hipStream_t s, int64_t N, int nvar, const int *dof, const int *rank, const double *x, double *out) {
  if (N <= 0 || nvar <= 0) return 0;
  hipLaunchKernelGGL(k_node_gather, vgrid(N * nvar), dim3(256), 0, s, N, nvar, dof, rank, x, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
// dst[i] = src[idx[i]]  (reference <-> internal DoF order)
__global__ void k_gather(int64_t n, const int *__restrict__ idx, const double *__restrict__ src, double *__restrict__ dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[idx[i]];
}
int gather(hipStream_t s, int64_t n, const int *idx, const double *src, double *dst) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_gather, vgrid(n), dim3(256), 0, s, n, idx, src, dst);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
__global__ void k_cast_vals(int64_t n, const double *v, float *vf) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) vf[i] = (float)v[i];
}
int cast_vals(hipStream_t s, int64_t n, const double *v, float *vf) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_cast_vals, vgrid(n), dim3(256), 0, s, n, v, vf);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// ---- element map: SELL position (within its row) of every element-matrix entry, found by binary search in the
// sorted CSR row and translated through kof; 0xFFFF where the pattern has no such entry
__global__ void k_elpos(int nel, int nloc, const int *__restrict__ eldof, const int64_t *__restrict__ rp,
                        const int *__restrict__ ci, const unsigned short *__restrict__ kof, unsigned short *__restrict__ pos) {
  int64_t total = (int64_t)nel * nloc * nloc;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    int64_t e = t / ((int64_t)nloc * nloc);
    int ab = (int)(t - e * nloc * nloc), a = ab / nloc, b = ab % nloc;
    int row = eldof[e * nloc + a], col = eldof[e * nloc + b];
    unsigned short k = 0xFFFF;
    if (row >= 0 && col >= 0) {
      int64_t lo = rp[row], hi = rp[row + 1] - 1;
      while (lo <= hi) {
        int64_t mid = (lo + hi) >> 1;
        int c = ci[mid];
        if (c == col) { k = kof[mid]; break; }
        if (c < col) lo = mid + 1; else hi = mid - 1;
      }
    }
    pos[t] = k;
  }
}
int elpos(hipStream_t s, int nel, int nloc, const int *eldof, const int64_t *rp, const int *ci,
          const unsigned short *kof, unsigned short *pos) {
  int64_t total = (int64_t)nel * nloc * nloc;
  int64_t b = (total + 255) / 256; if (b > 65536) b = 65536;
  hipLaunchKernelGGL(k_elpos, dim3((unsigned)b), dim3(256), 0, s, nel, nloc, eldof, rp, ci, kof, pos);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// STREAM-style copy (16-byte nontemporal loads and stores): the achievable-bandwidth yardstick.
// Measured on MI355X over 1 GiB: one element per thread per iteration with ~4 iterations per thread
// (64K blocks of 256) reaches ~6.1 TB/s; 8K blocks reach ~5.1 TB/s; 4-way unrolling does not help.
typedef double dvec2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_copy16(int64_t n16, const dvec2 *__restrict__ a, dvec2 *__restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
    __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}
int copy16(hipStream_t s, int64_t n16, const void *a, void *b) {
  int64_t nb = (n16 + 1023) / 1024;
  nb = nb < 1 ? 1 : (nb > 65536 ? 65536 : nb);
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)nb), dim3(256), 0, s, n16, (const dvec2 *)a, (dvec2 *)b);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace fmk
