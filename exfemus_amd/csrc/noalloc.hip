// noalloc.hip -- entry points that allocate nothing: their scratch memory is a fixed device array of this
// library, resolved once per device at context creation.
#include <hip/hip_runtime.h>
#include <mutex>
#include <string>
#include "fm_internal.h"
#include "fm_kernels.h"

// Scratch of fm_heat_nusselt_noalloc: per-element Nusselt contributions [0, MAXEL), per-element conductivity
// partials [MAXEL, 2 MAXEL), the functional's sum at 2 MAXEL.  16 MB per device, held until the process ends.
#define FM_NOALLOC_MAXEL (1 << 20)
static __device__ __align__(256) double scratch[2 * FM_NOALLOC_MAXEL + 64];

// Device address of scratch on each device (the array's name is only a host-side handle; every device holds
// its own copy at its own address).
static std::mutex g_mu;
static double *g_ptr[256] = {nullptr};

// Called by context creation with `device` current: resolves the device's copy of scratch once.
int fm_noalloc_attach(int device) {
  if (device < 0 || device >= 256) return fm_einval(__func__);
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return 0;
  void *p = nullptr;
  FM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_ptr[device] = static_cast<double *>(p);
  return 0;
}

// out[0] = sum of x[0 .. n) (one block)
__global__ __launch_bounds__(256) void k_sum(int n, const double *__restrict__ x, double *out) {
  __shared__ double sh[4];
  double s = 0.;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += x[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = sh[0] + sh[1] + sh[2] + sh[3];
}

// fm_heat_nusselt with its per-element scratch in the library's fixed device array: no device or host
// allocation.  Calls that overlap on one device race on that array.
extern "C" int fm_heat_nusselt_noalloc(fm_system *s, const fm_heat_params *p, int axis, const double *T_host, double *J,
                                       double *dJdT_host, double *dJdk_host) {
  if (!s || !s->initialized || s->lev.empty()) { fm_set_error("fm_heat_nusselt_noalloc: system not initialised"); return FM_ESTATE; }
  if (!p || !T_host || s->nq2 != 1 || s->nq1 != 0 || axis < 0 || axis >= s->mesh->dim || !(p->alpha0 > 0.))
    return fm_einval(__func__);
  const int top = (int)s->lev.size() - 1;
  Level &l = s->lev[top];
  if (l.nel > FM_NOALLOC_MAXEL) {
    fm_set_error("fm_heat_nusselt_noalloc: " + std::to_string(l.nel) + " top-level elements, the fixed scratch holds " +
                 std::to_string(FM_NOALLOC_MAXEL));
    return FM_EINVAL;
  }
  double *ptr = g_ptr[s->ctx->device];
  if (!ptr) { fm_set_error("fm_heat_nusselt_noalloc: scratch not resolved on this device"); return FM_ESTATE; }
  hipStream_t st = s->ctx->stream;
  const fm_mesh &m = *s->mesh;
  FM_HIP(hipSetDevice(s->ctx->device));
  FM_HIP(hipMemcpyAsync(l.t, T_host, l.n * sizeof(double), hipMemcpyHostToDevice, st));
  FM_CHECK(fmk::fill<double>(st, l.n, l.d, 0.0));
  // every element writes its own two entries: nothing in the scratch needs clearing
  double *jel = ptr, *dk = ptr + FM_NOALLOC_MAXEL, *sum = ptr + 2 * FM_NOALLOC_MAXEL;
  for (int k = 0; k < l.ncolor; k++) {
    const int n = l.color_off[k + 1] - l.color_off[k];
    FM_CHECK(fmk::heat_nusselt(st, m.dim, n, l.color_elems + l.color_off[k], l.conn, l.eldof, s->xyz, m.Ntop,
                               s->d_dof_top, s->vel, l.t, l.coef, p->alpha0, axis, s->d_tab, jel, dk, l.d));
  }
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, st, l.nel, jel, sum);
  FM_HIP(hipGetLastError());
  if (J) FM_HIP(hipMemcpyAsync(J, sum, sizeof(double), hipMemcpyDeviceToHost, st));
  if (dJdk_host) FM_HIP(hipMemcpyAsync(dJdk_host, dk, l.nel * sizeof(double), hipMemcpyDeviceToHost, st));
  if (dJdT_host) FM_HIP(hipMemcpyAsync(dJdT_host, l.d, l.n * sizeof(double), hipMemcpyDeviceToHost, st));
  FM_HIP(hipStreamSynchronize(st));
  return FM_OK;
}
