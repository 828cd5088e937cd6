// Shared definitions of the gfx950 quasi-3D PIC library (device + host).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace cylpic {

constexpr double C_LIGHT = 2.99792458e8;                              // constants.F90:175
constexpr double EPSILON0 = 8.854187817620389850536563031710750e-12;  // constants.F90:179
constexpr double PI = 3.141592653589793238462643383279503;            // constants.F90:172

// c_bc_* codes, constants.F90:56-72
enum BcCode {
  BC_PERIODIC = 1, BC_OTHER = 2, BC_SIMPLE_LASER = 3, BC_SIMPLE_OUTFLOW = 4,
  BC_OPEN = 5, BC_ZERO_GRADIENT = 7, BC_CLAMP = 8, BC_REFLECT = 9,
  BC_CONDUCT = 10, BC_THERMAL = 11, BC_CPML_LASER = 12, BC_CPML_OUTFLOW = 13,
  BC_MIXED = 14, BC_HEAT_BATH = 15, BC_ZERO_B = 16
};
enum { XMIN = 0, XMAX = 1, YMIN = 2, YMAX = 3 };

// device error flags (bitmask, read back by the host after a step)
enum { ERR_CELL_RANGE = 1, ERR_DCELL = 2, ERR_SORT_OVERFLOW = 4 };

// FP64 push: a particle key with this bit is left for the second pass
// (k_push_deposit_f64_rest); tile counts stay below it
constexpr int KEY_DEFERRED = 1 << 30;
// ... and a tile holding such a particle has this in its removed-count slot
// of the migration counts until the second pass counts it
constexpr unsigned TILE_DEFERRED = 0xffffffffu;

template <class T>
struct alignas(2 * sizeof(T)) cplx {
  T re, im;
};

// Grid scalars, passed by value to every kernel.
struct Grid {
  int nx, ny, M, ng;
  int nx2, ny2;          // nx + 2 ng, ny + 2 ng
  long plane;            // nx2 * ny2
  double dx, dy, dt;
  double x_min, x_max, y_min, y_max;
  double x_grid_min, y_grid_min;
  int bc_field[4];
  int bc_particle[4];
  // Fortran f(ix, iy, m), lower bounds (1-ng, 1-ng, 0)
  __host__ __device__ __forceinline__ long idx(int ix, int iy, int m) const {
    return (long)m * plane + (long)(iy + ng - 1) * nx2 + (ix + ng - 1);
  }
  // boundary column s(iy, m), lower bounds (1-ng, 0)
  __host__ __device__ __forceinline__ long idx2(int iy, int m) const {
    return (long)m * ny2 + (iy + ng - 1);
  }
};

// Workgroup -> tile: the dispatcher hands workgroup b to XCD b % 8, and each
// XCD has its own L2.  Giving XCD x the contiguous tile range
// [x q + min(x, r), ...) (q = n / 8, r = n % 8) in dispatch order keeps
// neighbouring tiles, whose field and current halos overlap, on one L2.
// A bijection of [0, n).
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int q = n / NUM_XCD, r = n % NUM_XCD;
  const int x = b % NUM_XCD, k = b / NUM_XCD;
  return x * q + (x < r ? x : r) + k;
}

// Tile decomposition of the push/deposit: tiles of tile_x x tile_r cells,
// numbered row-major in x; H halo cells around each tile in the LDS current box.
// Tile column t covers cells [t tx - shift + 1, (t + 1) tx - shift]: a moving
// window raises `shift` by one per cell it moves, so the tiles move with the
// plasma and no particle changes tile because of a window shift; at shift =
// tx the columns are relabelled one down.  With non-periodic x there is one
// spare tile column for that offset.
struct Tiles {
  int tx, tr, ntx, ntr, ntiles;
  int shift;
  // Key of a (clamped) cell: its tile (cell_key = false), or the tile-major
  // cell number tile * tx * tr + row-in-tile * tx + column-in-tile.
  __host__ __device__ __forceinline__ int key(int cx, int cy, int nx, int ny, bool cell_key) const {
    cx = cx < 1 ? 1 : (cx > nx ? nx : cx);
    cy = cy < 1 ? 1 : (cy > ny ? ny : cy);
    const int lx = cx - 1 + shift, ly = cy - 1, t = (ly / tr) * ntx + lx / tx;
    return cell_key ? (t * tr + ly % tr) * tx + lx % tx : t;
  }
  // first cell (x) of tile column ttx
  __host__ __device__ __forceinline__ int x0(int ttx) const { return ttx * tx + 1 - shift; }
};
constexpr int DEPOSIT_HALO = 3;

// The tile order that the XCD ranges above cut: bands of TILE_BAND tile rows,
// each band column by column.  The ~100 workgroups one XCD holds at a time
// then cover a block of about 12 x 8 tiles instead of a strip of one tile
// row, so the vertical halos of their field and current footprints are
// shared in that XCD's L2 instead of fetched again a whole tile row later.
// A bijection of [0, ntiles).
constexpr int TILE_BAND = 8;
__device__ __forceinline__ int band_tile(int L, int ntx, int ntr) {
  const int band = L / (TILE_BAND * ntx);
  const int rows = min(TILE_BAND, ntr - band * TILE_BAND);
  const int rem = L - band * TILE_BAND * ntx;
  return (band * TILE_BAND + rem % rows) * ntx + rem / rows;
}
// the tile of workgroup b of a one-workgroup-per-tile launch
__device__ __forceinline__ int tile_of_block(int b, const Tiles& tl) {
  return band_tile(xcd_tile(b, tl.ntiles), tl.ntx, tl.ntr);
}

}  // namespace cylpic
