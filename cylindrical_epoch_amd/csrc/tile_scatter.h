// Scatter of per-particle 3 x 3 stencils, one workgroup per tile, for the
// output diagnostics.  The tile's particles (its sorted range, then its range
// of injected batches) add into an LDS box of the tile plus a 2-cell margin;
// the box is then added to the output with one atomic per nonzero word.  A
// stencil that leaves the box (a particle that drifted since the last sort)
// goes straight to the output.  Sums in FP64; the order of the additions is
// not fixed.
//
// Op supplies, for particle i: the first cell (cx0, cy0) of its 3 x 3
// stencil, the weights w[b][a] of cell (cx0 + a, cy0 + b) and NC factors
// f[k] (the value added to word k of a cell is w[b][a] * f[k]), or false to
// skip the particle; and out_index(g, cx, cy, k), the output word of word k
// of a cell.
#pragma once

#include "launch.h"

namespace cylpic {

template <int NC, class Op>
__global__ __launch_bounds__(256) void k_tile_scatter(Grid g, Tiles tl, Op op, const int* __restrict__ key,
                                                      const int* __restrict__ tile_start, ExtraRange xr,
                                                      double* __restrict__ out) {
  extern __shared__ double box[];
  const int tile = tile_of_block(blockIdx.x, tl);
  const int ttx = tile % tl.ntx, tty = tile / tl.ntx;
  const int bx0 = tl.x0(ttx) - 2, br0 = tty * tl.tr + 1 - 2;
  const int BX = tl.tx + 4, BR = tl.tr + 4, nbox = BX * BR;
  for (int e = threadIdx.x; e < NC * nbox; e += blockDim.x) box[e] = 0.0;
  __syncthreads();
  const int p0 = tile_start[tile], p1 = tile_start[tile + 1];
  const int q0 = xr.start ? xr.base + xr.start[tile] : 0, q1 = xr.start ? xr.base + xr.start[tile + 1] : 0;
  const int np_main = p1 - p0, np = np_main + (q1 - q0);
  for (int k = threadIdx.x; k < np; k += blockDim.x) {
    const int ip = k < np_main ? p0 + k : q0 + (k - np_main);
    if ((unsigned)key[ip] >= (unsigned)tl.ntiles) continue;
    int cx0, cy0;
    double w[3][3], f[NC];
    if (!op.stencil(ip, cx0, cy0, w, f)) continue;
    const bool in = cx0 >= bx0 && cx0 + 2 < bx0 + BX && cy0 >= br0 && cy0 + 2 < br0 + BR;
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          if (f[c] == 0.0) continue;
          const double v = w[b][a] * f[c];
          if (in) atomicAdd(&box[c * nbox + (cy0 + b - br0) * BX + (cx0 + a - bx0)], v);
          else unsafeAtomicAdd(out + op.out_index(g, cx0 + a, cy0 + b, c), v);
        }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < NC * nbox; e += blockDim.x) {
    const double v = box[e];
    if (v == 0.0) continue;
    const int c = e / nbox, cell = e - c * nbox;
    const int cx = bx0 + cell % BX, cy = br0 + cell / BX;
    if (cx < 1 - g.ng || cx > g.nx + g.ng || cy < 1 - g.ng || cy > g.ny + g.ng) continue;
    unsafeAtomicAdd(out + op.out_index(g, cx, cy, c), v);
  }
}

template <int NC, class Op>
inline void launch_tile_scatter(const Grid& g, const Tiles& tl, const Op& op, const int* key, const int* tile_start,
                                ExtraRange xr, double* out, hipStream_t s) {
  const size_t lds = (size_t)NC * (tl.tx + 4) * (tl.tr + 4) * sizeof(double);
  hipLaunchKernelGGL((k_tile_scatter<NC, Op>), dim3(tl.ntiles), dim3(256), lds, s, g, tl, op, key, tile_start, xr,
                     out);
}

}  // namespace cylpic
