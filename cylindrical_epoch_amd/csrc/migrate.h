// Direction counts of the neighbour migration sort (sort.hip), shared by the
// push kernels, which count their own tile's particles after the push, and
// by the stand-alone counting kernel.
#pragma once
#include "launch.h"

namespace cylpic {

// direction of the move from source tile S to key d: 0..8 = (dty+1)*3 + (dtx+1),
// MIG_REMOVED for a removed particle, -1 for a farther move (x is taken
// modulo the tile columns, which covers periodic wraps)
__device__ __forceinline__ int mig_dir(uint32_t d, int S, int ntx, uint32_t ntiles) {
  if (d >= ntiles) return MIG_REMOVED;
  const int dty = (int)(d / (uint32_t)ntx) - S / ntx;
  int dtx = (int)(d % (uint32_t)ntx) - S % ntx;
  if (dtx == ntx - 1) dtx = -1;
  else if (dtx == 1 - ntx) dtx = 1;
  if (dtx < -1 || dtx > 1 || dty < -1 || dty > 1) return -1;
  return (dty + 1) * 3 + (dtx + 1);
}

// Counts the keys of the range [p0, p1) of tile S by direction into
// counts[S * MIG_DIRS + dir]; flag[0] |= 1 on a farther move, flag[1] +=
// particles kept.  All work-items of the workgroup call it; c is MIG_DIRS
// words of LDS.  The keys must be visible to the workgroup.
__device__ __forceinline__ void mig_count_tile(const uint32_t* __restrict__ key, int S, int p0, int p1, int ntx,
                                               int ntiles, uint32_t* counts, int* flag, uint32_t* c) {
  if (threadIdx.x < MIG_DIRS) c[threadIdx.x] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  bool far = false;
  for (int base = p0; base < p1; base += blockDim.x) {
    const int i = base + (int)threadIdx.x;
    const bool valid = i < p1;
    const int dir = valid ? mig_dir(key[i], S, ntx, (uint32_t)ntiles) : -2;
    far |= valid && dir < 0;
#pragma unroll
    for (int k = 0; k < MIG_DIRS; ++k) {
      const unsigned long long m = __ballot(dir == k);
      if (lane == 0 && m) atomicAdd(&c[k], (uint32_t)__popcll(m));
    }
  }
  if (__ballot(far) && lane == 0) atomicOr(flag, 1);
  __syncthreads();
  if (threadIdx.x < MIG_DIRS) counts[(long)S * MIG_DIRS + threadIdx.x] = c[threadIdx.x];
  if (threadIdx.x == 0) {
    uint32_t alive = 0;
    for (int k = 0; k < 9; ++k) alive += c[k];
    if (alive) atomicAdd(flag + 1, (int)alive);
  }
  __syncthreads();  // c may be reused by the caller
}

}  // namespace cylpic
