// stencil.hpp — device helpers shared by the forward (kernels.hip) and reverse (adjoint.hip)
// step kernels.
#pragma once

#include "engine.hpp"

#include <hip/hip_runtime.h>

namespace mx {

constexpr int BLOCK = 256;

// logical block index, dealt to XCDs in contiguous runs (gridDim.x is a multiple of 8);
// rev walks them in descending order
__device__ __forceinline__ uint32_t xcd_block(int rev = 0) {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t lb = (b & 7u) * (nb >> 3) + (b >> 3);
  return rev ? nb - 1 - lb : lb;
}

// halo point along a folded axis fed by the owned point z of a component with Yee shift ysh
// there; -1 when z is not the wrapped end
__device__ __forceinline__ int zwrap_dst(int ysh, int z, int n) {
  if (ysh == 0) return z == n ? 0 : -1;
  return z == 0 ? n : -1;
}

// coordinate along loop axis k (a select over registers, never an indexed access)
__device__ __forceinline__ int asel(int j0, int j1, int j2, int k) { return k == 0 ? j0 : (k == 1 ? j1 : j2); }

__device__ __forceinline__ bool slot_owned(const AxisSlot &s, int j0, int j1, int j2) {
  return j0 >= s.lo[0] && j0 <= s.hi[0] && j1 >= s.lo[1] && j1 <= s.hi[1] && j2 >= s.lo[2] && j2 <= s.hi[2];
}
// on_metal_boundary for an owned point (src/boundaries.cpp:191-205)
__device__ __forceinline__ bool slot_metal(const AxisSlot &s, int j0, int j1, int j2) {
  return j0 == s.mhi[0] || j0 == s.mlo[0] || j1 == s.mhi[1] || j1 == s.mlo[1] || j2 == s.mhi[2] || j2 == s.mlo[2];
}

template <class R> __device__ __forceinline__ bool pml_on(const R *sig, const R *kap, int k) {
  return sig && (sig[k] != R(0) || kap[k] != R(1));
}

static inline unsigned nblocks(size_t n) { return unsigned((n + BLOCK - 1) / BLOCK); }
static inline unsigned nblocks8(size_t n) { return (nblocks(n) + 7u) / 8u * 8u; }

} // namespace mx
