// fm_device.h -- device helpers shared by the element kernels.
#pragma once
#include <hip/hip_runtime.h>

// Block index for element loops "for (i = xcd_block(); i < n; i += gridDim.x)": the hardware hands blocks to
// the 8 XCDs round-robin (block b to XCD b mod 8), so consecutive list entries -- neighbouring elements, which
// share matrix rows and vector entries -- would land on 8 different L2 caches.  Renumbered, the blocks of one
// XCD take a contiguous run of gridDim.x / 8 entries in every round of the loop.  Grids that are not a multiple
// of 8 keep the hardware order.
__device__ __forceinline__ int xcd_block() {
  const int g = (int)gridDim.x, b = (int)blockIdx.x;
  return (g & 7) ? b : (b & 7) * (g >> 3) + (b >> 3);
}
