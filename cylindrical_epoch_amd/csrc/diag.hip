// Output diagnostics computed from the device-resident particles: the mode
// decomposition of a species' number density (calc_number_density_modes,
// src/io/calc_df.F90:588-661) and the binned distribution function
// (general_dist_fn, src/io/dist_fn.F90:118-715).  Both read the particle
// arrays once (HBM-bound) and accumulate with FP64 atomics; the sums are
// exact up to the order of the additions.  Particles are reconstructed in
// FP64 exactly as cylpic_download_particles returns them, so the cell
// indices, shape weights and bins equal those of a CPU pass over the
// downloaded particles.
#include <algorithm>

#include "cont_modes.h"
#include "cplx_ops.h"
#include "launch.h"
#include "tile_scatter.h"

namespace cylpic {

namespace {

constexpr double PI_D = 3.141592653589793238462643383279503;
constexpr double C_TINY_D = 2.2250738585072014e-308;  // TINY(1.0_num)

__device__ __forceinline__ void load_particle(const Grid&, const PartF64& p, long i, double& x, double& y,
                                              double& z, double* pm, double& w, double) {
  x = p.x[i];
  y = p.y[i];
  z = p.z[i];
  if (pm) {
    pm[0] = p.px[i];
    pm[1] = p.py[i];
    pm[2] = p.pz[i];
  }
  w = p.w[i];
}

// the FP64 values cylpic_download_particles reconstructs
__device__ __forceinline__ void load_particle(const Grid& g, const PartMixed& p, long i, double& x, double& y,
                                              double& z, double* pm, double& w, double mc) {
  x = g.x_grid_min + ((double)(p.cell[i] - 1) + (double)p.ox[i]) * g.dx;
  y = (double)p.y[i] * g.dy;
  z = (double)p.z[i] * g.dy;
  if (pm) {
    pm[0] = (double)p.ux[i] * mc;
    pm[1] = (double)p.uy[i] * mc;
    pm[2] = (double)p.uz[i] * mc;
  }
  w = (double)p.w[i];
}

// One particle's contribution: particle_to_grid.inc with triangle/gxfac.inc
// (quadratic weights summing to one; r < dy folds the inner radial weight
// into the nearest cell), the macro-particle volume 2 pi dx dy r
// (partlist.F90:999-1013) and the mode factor 1 or 2 e^{i m theta}.
template <class P>
__global__ __launch_bounds__(256) void k_number_density(Grid g, P p, long n, double mc, double* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x, y, z, w;
  load_particle(g, p, i, x, y, z, nullptr, w, mc);
  const double part_r = sqrt(y * y + z * z);
  const double cxr = (x - g.x_grid_min) / g.dx;
  const double cyr = (part_r - g.y_grid_min) / g.dy;
  int cx = (int)floor(cxr + 0.5), cy = (int)floor(cyr + 0.5);
  const double fx = (double)cx - cxr, fy = (double)cy - cyr;
  cx += 1;
  cy += 1;
  if (cx - 1 < 1 - g.ng || cx + 1 > g.nx + g.ng || cy - 1 < 1 - g.ng || cy + 1 > g.ny + g.ng) return;
  const double cx2 = fx * fx, cy2 = fy * fy;
  const double gx[3] = {0.5 * ((0.25 + cx2) + fx), 0.75 - cx2, 0.5 * ((0.25 + cx2) - fx)};
  double gy[3] = {0.5 * ((0.25 + cy2) + fy), 0.75 - cy2, 0.5 * ((0.25 + cy2) - fy)};
  if (part_r < g.dy) {
    gy[1] = gy[1] + gy[0];
    gy[0] = 0.0;
  }
  const double vol = (((2.0 * PI_D) * g.dx) * g.dy) * part_r;
  const double pnd = w / vol;
  const cplx<double> eith = mk(y / part_r, z / part_r);
  cplx<double> eim = mk(1.0, 0.0), mf = mk(1.0, 0.0);
  for (int m = 0; m < g.M; ++m) {
    if (m > 0) {
      eim = cmul(eim, eith);
      mf = mk(2.0 * eim.re, 2.0 * eim.im);
    }
#pragma unroll
    for (int iy = 0; iy < 3; ++iy)
#pragma unroll
      for (int ix = 0; ix < 3; ++ix) {
        const double v = (gx[ix] * gy[iy]) * pnd;
        double* o = out + 2 * g.idx(cx + ix - 1, cy + iy - 1, m);
        unsafeAtomicAdd(o, v * mf.re);
        if (m > 0) unsafeAtomicAdd(o + 1, v * mf.im);  // m = 0 adds v * 0
      }
  }
}

// The same contribution as a tile-scatter operation: NC = 2 M words per cell
// (the real and imaginary part of each mode)
template <class P>
struct DensityOp {
  Grid g;
  P p;
  __device__ bool stencil(long i, int& cx0, int& cy0, double w[3][3], double* f) const {
    double x, y, z, wt;
    load_particle(g, p, i, x, y, z, nullptr, wt, 0.0);
    const double part_r = sqrt(y * y + z * z);
    const double cxr = (x - g.x_grid_min) / g.dx;
    const double cyr = (part_r - g.y_grid_min) / g.dy;
    int cx = (int)floor(cxr + 0.5), cy = (int)floor(cyr + 0.5);
    const double fx = (double)cx - cxr, fy = (double)cy - cyr;
    cx += 1;
    cy += 1;
    if (cx - 1 < 1 - g.ng || cx + 1 > g.nx + g.ng || cy - 1 < 1 - g.ng || cy + 1 > g.ny + g.ng) return false;
    const double cx2 = fx * fx, cy2 = fy * fy;
    const double gx[3] = {0.5 * ((0.25 + cx2) + fx), 0.75 - cx2, 0.5 * ((0.25 + cx2) - fx)};
    double gy[3] = {0.5 * ((0.25 + cy2) + fy), 0.75 - cy2, 0.5 * ((0.25 + cy2) - fy)};
    if (part_r < g.dy) {
      gy[1] = gy[1] + gy[0];
      gy[0] = 0.0;
    }
    const double pnd = wt / ((((2.0 * PI_D) * g.dx) * g.dy) * part_r);
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a) w[b][a] = (gx[a] * gy[b]) * pnd;
    const cplx<double> eith = mk(y / part_r, z / part_r);
    cplx<double> eim = mk(1.0, 0.0);
    f[0] = 1.0;
    f[1] = 0.0;
    for (int m = 1; m < g.M; ++m) {
      eim = cmul(eim, eith);
      f[2 * m] = 2.0 * eim.re;
      f[2 * m + 1] = 2.0 * eim.im;
    }
    cx0 = cx - 1;
    cy0 = cy - 1;
    return true;
  }
  __device__ long out_index(const Grid& gg, int cx, int cy, int k) const {
    return 2 * gg.idx(cx, cy, k >> 1) + (k & 1);
  }
};

// calc_boundary_modes (calc_df.F90:35-55) = the real processor_summation_bcs
// (boundary.F90:1305-1326) on both parts: particle_reflection_bcs without a
// flip (:833-914), x part, one thread per (row, mode)
__global__ void k_dens_reflect_x(Grid g, cplx<double>* __restrict__ a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.ny2 * g.M) return;
  const int j = 1 - g.ng + t % g.ny2, m = t / g.ny2, nx = g.nx, ng = g.ng;
  if (g.bc_particle[XMIN] == BC_REFLECT)
    for (int i = 1; i <= ng - 1; ++i) {
      a[g.idx(i, j, m)] = cadd(a[g.idx(i, j, m)], a[g.idx(1 - i, j, m)]);
      a[g.idx(1 - i, j, m)] = czero<double>();
    }
  if (g.bc_particle[XMAX] == BC_REFLECT)
    for (int i = 1; i <= ng; ++i) {
      a[g.idx(nx + 1 - i, j, m)] = cadd(a[g.idx(nx + 1 - i, j, m)], a[g.idx(nx + i, j, m)]);
      a[g.idx(nx + i, j, m)] = czero<double>();
    }
}

// ... its r_max part, one thread per (column, mode)
__global__ void k_dens_reflect_rmax(Grid g, cplx<double>* __restrict__ a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.nx2 * g.M) return;
  const int ix = 1 - g.ng + t % g.nx2, m = t / g.nx2, ny = g.ny;
  for (int i = 1; i <= g.ng; ++i) {
    a[g.idx(ix, ny + 1 - i, m)] = cadd(a[g.idx(ix, ny + 1 - i, m)], a[g.idx(ix, ny + i, m)]);
    a[g.idx(ix, ny + i, m)] = czero<double>();
  }
}

// particle_periodic_bcs on one rank (boundary.F90:1019-1129): the ghost
// columns are added into the interior and keep their values
__global__ void k_dens_periodic(Grid g, cplx<double>* __restrict__ a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.ny2 * g.M) return;
  const int j = 1 - g.ng + t % g.ny2, m = t / g.ny2, nx = g.nx, ng = g.ng;
  for (int i = 1; i <= ng; ++i) a[g.idx(i, j, m)] = cadd(a[g.idx(i, j, m)], a[g.idx(nx + i, j, m)]);
  for (int i = 1; i <= ng; ++i)
    a[g.idx(nx - ng + i, j, m)] = cadd(a[g.idx(nx - ng + i, j, m)], a[g.idx(i - ng, j, m)]);
}

// field_mode_zero_gradient of a cell-centred array (boundary.F90:654-707):
// x_min and x_max per row, then y_min and y_max per column
__global__ void k_dens_zero_gradient_x(Grid g, cplx<double>* __restrict__ a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.ny2 * g.M) return;
  const int j = 1 - g.ng + t % g.ny2, m = t / g.ny2, nx = g.nx, ng = g.ng;
  if (g.bc_field[XMIN] != BC_PERIODIC)
    for (int i = 1; i <= ng; ++i) a[g.idx(i - ng, j, m)] = a[g.idx(ng + 1 - i, j, m)];
  if (g.bc_field[XMAX] != BC_PERIODIC)
    for (int i = 1; i <= ng; ++i) a[g.idx(nx + i, j, m)] = a[g.idx(nx + 1 - i, j, m)];
}

__global__ void k_dens_zero_gradient_y(Grid g, cplx<double>* __restrict__ a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.nx2 * g.M) return;
  const int ix = 1 - g.ng + t % g.nx2, m = t / g.nx2, ny = g.ny, ng = g.ng;
  if (g.bc_field[YMIN] != BC_PERIODIC)
    for (int i = 1; i <= ng; ++i) a[g.idx(ix, i - ng, m)] = a[g.idx(ix, ng + 1 - i, m)];
  if (g.bc_field[YMAX] != BC_PERIODIC)
    for (int i = 1; i <= ng; ++i) a[g.idx(ix, ny + i, m)] = a[g.idx(ix, ny + 1 - i, m)];
}

// ------------------------------------------------------------------------
// distribution functions

// particle_data(:) of a non-photon particle (dist_fn.F90:486-553); angles
// only where `need` asks; false when a needed angle has |p| < tiny (CYCLE)
template <class P>
__device__ __forceinline__ bool df_data(const Grid& g, const P& p, long i, double mc, double mc2, const DistFn& d,
                                        bool wrap, double* pd, double& w) {
  double x, y, z, pm[3];
  load_particle(g, p, i, x, y, z, pm, w, mc);
  const double ux = pm[0] / mc, uy = pm[1] / mc, uz = pm[2] / mc;
  const double u2 = (ux * ux + uy * uy) + uz * uz;
  const double gam = sqrt(u2 + 1.0);
  const double gm1 = u2 / (gam + 1.0);
  pd[DF_X] = x;
  pd[DF_Y] = y;
  pd[DF_PX] = pm[0];
  pd[DF_PY] = pm[1];
  pd[DF_PZ] = pm[2];
  pd[DF_EN] = gm1 * mc2;
  pd[DF_GAMMA_M1] = gm1;
  pd[DF_MOD_P] = sqrt((pm[0] * pm[0] + pm[1] * pm[1]) + pm[2] * pm[2]);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    pd[DF_XY_ANGLE + k] = 0.0;
    if (!d.need_angle[k]) continue;
    const double a1 = pm[k], a2 = pm[(k + 1) % 3];
    const double pp = sqrt(a1 * a1 + a2 * a2);
    if (pp < C_TINY_D) return false;
    double th = asin(a2 / pp);
    if (a1 < 0.0) th = copysign(1.0, a2) * PI_D - th;
    if (wrap && th + 2.0 * PI_D < d.angle_max[k]) th = th + 2.0 * PI_D;
    pd[DF_XY_ANGLE + k] = th;
  }
  return true;
}

__device__ __forceinline__ bool df_restricted(const DistFn& d, const double* pd) {
  for (int k = 1; k <= DF_MAXDIRS; ++k)
    if (d.use_restriction[k - 1] && (pd[k] < d.restriction[k - 1][0] || pd[k] > d.restriction[k - 1][1]))
      return true;
  return false;
}

// per-block minima and maxima of the auto-ranged directions
template <class P>
__global__ __launch_bounds__(256) void k_df_range(Grid g, P p, long n, double mc, DistFn d, double* __restrict__ part) {
  __shared__ double smin[3][256], smax[3][256];
  double lo[3] = {1.7976931348623157e308 * 0.4, 1.7976931348623157e308 * 0.4, 1.7976931348623157e308 * 0.4};
  double hi[3] = {-lo[0], -lo[1], -lo[2]};
  const double mc2 = mc * C_LIGHT;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double pd[DF_MAXDIRS + 1], w;
    if (!df_data(g, p, i, mc, mc2, d, false, pd, w)) continue;
    if (df_restricted(d, pd)) continue;
    for (int k = 0; k < d.ndims; ++k)
      if (d.calc_range[k]) {
        const double v = pd[d.direction[k]];
        lo[k] = v < lo[k] ? v : lo[k];
        hi[k] = v > hi[k] ? v : hi[k];
      }
  }
  for (int k = 0; k < 3; ++k) {
    smin[k][threadIdx.x] = lo[k];
    smax[k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 3; ++k) {
        smin[k][threadIdx.x] = fmin(smin[k][threadIdx.x], smin[k][threadIdx.x + s]);
        smax[k][threadIdx.x] = fmax(smax[k][threadIdx.x], smax[k][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) {
      part[(long)blockIdx.x * 6 + 2 * k] = smin[k][0];
      part[(long)blockIdx.x * 6 + 2 * k + 1] = smax[k][0];
    }
}

// the binning pass (dist_fn.F90:482-578): a histogram of up to DF_LDS_BINS
// bins is accumulated per workgroup in LDS and added to `out` at the end;
// larger ones go straight to `out`
template <class P, bool LDS>
__global__ __launch_bounds__(256) void k_df_bin(Grid g, P p, long n, double mc, DistFn d, double* __restrict__ out) {
  __shared__ double hist[LDS ? DF_LDS_BINS : 1];
  const long nbins = (long)d.resolution[0] * d.resolution[1] * d.resolution[2];
  if (LDS) {
    for (int k = threadIdx.x; k < nbins; k += blockDim.x) hist[k] = 0.0;
    __syncthreads();
  }
  const double mc2 = mc * C_LIGHT;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double pd[DF_MAXDIRS + 1], w;
    if (!df_data(g, p, i, mc, mc2, d, true, pd, w)) continue;
    if (df_restricted(d, pd)) continue;
    long cell[3] = {0, 0, 0};
    bool ok = true;
    for (int k = 0; k < d.ndims; ++k) {
      const double c = floor((pd[d.direction[k]] - d.range[k][0]) / d.dgrid[k]);
      if (!(c >= 0.0 && c < (double)d.resolution[k])) {
        ok = false;
        break;
      }
      cell[k] = (long)c;
    }
    if (!ok) continue;
    const long b = (cell[2] * d.resolution[1] + cell[1]) * d.resolution[0] + cell[0];
    if (LDS) atomicAdd(&hist[b], w);
    else unsafeAtomicAdd(&out[b], w);
  }
  if (LDS) {
    __syncthreads();
    for (int k = threadIdx.x; k < nbins; k += blockDim.x)
      if (hist[k] != 0.0) unsafeAtomicAdd(&out[k], hist[k]);
  }
}

// ------------------------------------------------------------------------
// continuity check of the FP64 deposit: the m = 0 charge of a particle at
// x + v dt/2 (the half push of particles.F90:296-312) on the staggered grid
// with quadratic-spline fractions summing to one, the grid on which the
// charge-conserving deposit (particles.F90:508-665) keeps
// Q(t+3dt/2) - Q(t+dt/2) + dt (net current out of the cell) = 0.  The mixed
// path's charge (its own FP32 half step) is in push_mixed.hip.
// The charge at x + vs v dt/2: vs = 1 the half step ahead, -1 the half step
// behind, 0 at x itself (a species that does not move).
template <class P>
__global__ __launch_bounds__(256) void k_charge_half(Grid g, P p, long n, int ntiles, double mc, double q, double vs,
                                                     double* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (unsigned)p.key[i] >= (unsigned)ntiles) return;
  double x, y, z, pm[3], w;
  load_particle(g, p, i, x, y, z, pm, w, mc);
  const double ux = pm[0] / mc, uy = pm[1] / mc, uz = pm[2] / mc;
  const double root = vs * ((C_LIGHT * (g.dt / 2.0)) / sqrt(((ux * ux + uy * uy) + uz * uz) + 1.0));
  const double xh = x + ux * root, yh = y + uy * root, zh = z + uz * root;
  const double xr = (xh - g.x_grid_min) / g.dx;
  const double rr = (sqrt(yh * yh + zh * zh) - g.y_grid_min) / g.dy;
  const double fcx = floor(xr), fcy = floor(rr);
  const int cx = (int)fcx, cy = (int)fcy;  // first of the three cells (Fortran index)
  if (cx < 1 - g.ng || cx + 2 > g.nx + g.ng || cy < 1 - g.ng || cy + 2 > g.ny + g.ng) return;
  const double fx = fcx - xr + 0.5, fy = fcy - rr + 0.5;
  const double wx[3] = {0.5 * ((0.25 + fx * fx) + fx), 0.5 * (1.5 - 2.0 * fx * fx), 0.5 * ((0.25 + fx * fx) - fx)};
  const double wy[3] = {0.5 * ((0.25 + fy * fy) + fy), 0.5 * (1.5 - 2.0 * fy * fy), 0.5 * ((0.25 + fy * fy) - fy)};
  const double qw = q * w;
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k) unsafeAtomicAdd(out + g.idx(cx + k, cy + j, 0), qw * (wx[k] * wy[j]));
}

// the same charge as a tile-scatter operation (one real word per cell)
struct ChargeOp64 {
  Grid g;
  PartF64 p;
  double mc, q, vs;
  __device__ bool stencil(long i, int& cx0, int& cy0, double w[3][3], double* f) const {
    double x, y, z, pm[3], wt;
    load_particle(g, p, i, x, y, z, pm, wt, mc);
    const double ux = pm[0] / mc, uy = pm[1] / mc, uz = pm[2] / mc;
    const double root = vs * ((C_LIGHT * (g.dt / 2.0)) / sqrt(((ux * ux + uy * uy) + uz * uz) + 1.0));
    const double xh = x + ux * root, yh = y + uy * root, zh = z + uz * root;
    const double xr = (xh - g.x_grid_min) / g.dx;
    const double rr = (sqrt(yh * yh + zh * zh) - g.y_grid_min) / g.dy;
    const double fcx = floor(xr), fcy = floor(rr);
    cx0 = (int)fcx;
    cy0 = (int)fcy;
    if (cx0 < 1 - g.ng || cx0 + 2 > g.nx + g.ng || cy0 < 1 - g.ng || cy0 + 2 > g.ny + g.ng) return false;
    const double fx = fcx - xr + 0.5, fy = fcy - rr + 0.5;
    const double wx[3] = {0.5 * ((0.25 + fx * fx) + fx), 0.5 * (1.5 - 2.0 * fx * fx), 0.5 * ((0.25 + fx * fx) - fx)};
    const double wy[3] = {0.5 * ((0.25 + fy * fy) + fy), 0.5 * (1.5 - 2.0 * fy * fy), 0.5 * ((0.25 + fy * fy) - fy)};
    const double qw = q * wt;
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a) w[b][a] = qw * (wx[a] * wy[b]);
    f[0] = 1.0;
    return true;
  }
  __device__ long out_index(const Grid& gg, int cx, int cy, int) const { return gg.idx(cx, cy, 0); }
};

// max |Q1 - Q0 + dt (fluxes of Jx, Jr)|, max |Q1 - Q0| and max |Q1| of the
// m = 0 cells at least ng from every edge; out[0..2] hold the bit patterns
// of the non-negative maxima (ordered as unsigned integers)
template <class T>
__global__ __launch_bounds__(256) void k_continuity_m0(Grid g, const double* __restrict__ q0,
                                                       const double* __restrict__ q1, const cplx<T>* __restrict__ jx,
                                                       const cplx<T>* __restrict__ jr, Geo geo,
                                                       unsigned long long* __restrict__ out) {
  __shared__ double red[3][4];
  const int nxi = g.nx - 2 * g.ng, nyi = g.ny - 2 * g.ng;
  const long n = (long)nxi * nyi;
  double res = 0.0, dq = 0.0, qq = 0.0;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int cx = g.ng + 1 + (int)(e % nxi), cy = g.ng + 1 + (int)(e / nxi);
    const int jng = g.ng;
    const double flux = ((double)jx[g.idx(cx + 1, cy, 0)].re - (double)jx[g.idx(cx, cy, 0)].re) /
                            geo.inv_area_rt[cy + jng] +
                        (double)jr[g.idx(cx, cy + 1, 0)].re / geo.inv_area_xt[cy + jng] -
                        (double)jr[g.idx(cx, cy, 0)].re / geo.inv_area_xt[cy - 1 + jng];
    const long k = g.idx(cx, cy, 0);
    const double d = q1[k] - q0[k];
    res = fmax(res, fabs(d + g.dt * flux));
    dq = fmax(dq, fabs(d));
    qq = fmax(qq, fabs(q1[k]));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    res = fmax(res, __shfl_xor(res, o));
    dq = fmax(dq, __shfl_xor(dq, o));
    qq = fmax(qq, __shfl_xor(qq, o));
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][wv] = res;
    red[1][wv] = dq;
    red[2][wv] = qq;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double v = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v = fmax(v, red[threadIdx.x][w]);
    atomicMax(out + threadIdx.x, (unsigned long long)__double_as_longlong(v));
  }
}

// Gauss's law for m = 0 on the cells of the continuity check.  A step
// changes E by -dt (J_prev + J) / (2 eps0) plus a curl whose net flux
// vanishes, J_prev and J carrying the charge between the half-step positions
// of two successive pushes, so eps0 (net flux of Ex, Er out of the cell) - Q
// keeps its value from step to step, Q the mean of the charges at
// x - v dt/2 (qa) and x + v dt/2 (qb) between steps.  out[0..2]: max |that
// difference|, max |Q|, max |eps0 net flux| (bit patterns, as above).
template <class T>
__global__ __launch_bounds__(256) void k_gauss_m0(Grid g, const double* __restrict__ qa, const double* __restrict__ qb,
                                                  const cplx<T>* __restrict__ ex,
                                                  const cplx<T>* __restrict__ er, Geo geo,
                                                  unsigned long long* __restrict__ out) {
  __shared__ double red[3][4];
  const int nxi = g.nx - 2 * g.ng, nyi = g.ny - 2 * g.ng;
  const long n = (long)nxi * nyi;
  double res = 0.0, qq = 0.0, ff = 0.0;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int cx = g.ng + 1 + (int)(e % nxi), cy = g.ng + 1 + (int)(e / nxi);
    const int jng = g.ng;
    const double flux = EPSILON0 * (((double)ex[g.idx(cx + 1, cy, 0)].re - (double)ex[g.idx(cx, cy, 0)].re) /
                                        geo.inv_area_rt[cy + jng] +
                                    (double)er[g.idx(cx, cy + 1, 0)].re / geo.inv_area_xt[cy + jng] -
                                    (double)er[g.idx(cx, cy, 0)].re / geo.inv_area_xt[cy - 1 + jng]);
    const long k = g.idx(cx, cy, 0);
    const double qc = 0.5 * (qa[k] + qb[k]);
    res = fmax(res, fabs(flux - qc));
    qq = fmax(qq, fabs(qc));
    ff = fmax(ff, fabs(flux));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    res = fmax(res, __shfl_xor(res, o));
    qq = fmax(qq, __shfl_xor(qq, o));
    ff = fmax(ff, __shfl_xor(ff, o));
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][wv] = res;
    red[1][wv] = qq;
    red[2][wv] = ff;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double v = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v = fmax(v, red[threadIdx.x][w]);
    atomicMax(out + threadIdx.x, (unsigned long long)__double_as_longlong(v));
  }
}

// Continuity of every mode, FP64 path: the charge modes at the push's half
// step (the arithmetic of k_charge_half) before the push ...
struct CmBeginF64 {
  PartF64 p;
  double mc, q;
  template <class To>
  __device__ void operator()(const Grid& g, long i, const To& to) const {
    const double ux = p.px[i] / mc, uy = p.py[i] / mc, uz = p.pz[i] / mc;
    const double root = (C_LIGHT * (g.dt / 2.0)) / sqrt(((ux * ux + uy * uy) + uz * uz) + 1.0);
    const double xh = p.x[i] + ux * root, yh = p.y[i] + uy * root, zh = p.z[i] + uz * root;
    const double rh = sqrt(yh * yh + zh * zh);
    const ContPos a{(xh - g.x_grid_min) / g.dx, (rh - g.y_grid_min) / g.dy, rh > 0.0 ? yh / rh : 1.0,
                    rh > 0.0 ? zh / rh : 0.0, false};
    cm_charge(g, q * p.w[i], a, to, 0);
  }
};

// ... and after it: the charge modes at t + 3dt/2 = x(t+dt) + v dt/2 and the
// azimuthal terms of the move from x(t+dt) - v dt/2 to there
struct CmEndF64 {
  PartF64 p;
  double mc, q;
  template <class To>
  __device__ void operator()(const Grid& g, long i, const To& to) const {
    const double ux = p.px[i] / mc, uy = p.py[i] / mc, uz = p.pz[i] / mc;
    const double gam = sqrt(((ux * ux + uy * uy) + uz * uz) + 1.0);
    const double root = (C_LIGHT * (g.dt / 2.0)) / gam;
    const double x = p.x[i], y = p.y[i], zz = p.z[i];
    ContPos pos[2];
    for (int k = 0; k < 2; ++k) {
      const double sg = k ? 1.0 : -1.0;
      const double xs = x + sg * (ux * root), ys = y + sg * (uy * root), zs = zz + sg * (uz * root);
      const double rs = sqrt(ys * ys + zs * zs);
      pos[k] = ContPos{(xs - g.x_grid_min) / g.dx, (rs - g.y_grid_min) / g.dy, rs > 0.0 ? ys / rs : 1.0,
                       rs > 0.0 ? zs / rs : 0.0, false};
    }
    const double qw = q * p.w[i];
    cm_charge(g, qw, pos[1], to, 0);
    const double r1 = sqrt(y * y + zz * zz);
    const double vtdt_r = r1 > 0.0 ? ((-(uy * zz) + uz * y) / r1 / gam) * (C_LIGHT * g.dt) / r1 : 0.0;
    cm_azimuthal(g, qw, pos[0], pos[1], vtdt_r, to);
  }
};

// one thread per slot (species without tile ranges)
template <class Part>
__global__ __launch_bounds__(256) void k_cont_modes_flat(Grid g, long n, int ntiles, const int* __restrict__ key,
                                                         const int* __restrict__ ts, ExtraRange xr, Part part,
                                                         CmGlobal out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !cm_live(i, ts, xr.start, xr.base, ntiles) || (unsigned)key[i] >= (unsigned)ntiles) return;
  part(g, i, out);
}

template <class T>
__global__ __launch_bounds__(256) void k_continuity_modes(Grid g, const double* __restrict__ q0,
                                                          const double* __restrict__ q1, const double* __restrict__ z,
                                                          const double* __restrict__ zp, const cplx<T>* __restrict__ jx,
                                                          const cplx<T>* __restrict__ jr, Geo geo,
                                                          unsigned long long* __restrict__ out) {
  __shared__ double red[4][4];
  const int m = blockIdx.y;
  const int nxi = g.nx - 2 * g.ng, nyi = g.ny - 2 * g.ng;
  const long n = (long)nxi * nyi;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int cx = g.ng + 1 + (int)(e % nxi), cy = g.ng + 1 + (int)(e / nxi);
    const int jng = g.ng;
    const cplx<T> a = jx[g.idx(cx + 1, cy, m)], b = jx[g.idx(cx, cy, m)];
    const cplx<T> c = jr[g.idx(cx, cy + 1, m)], d = jr[g.idx(cx, cy, m)];
    const double irt = geo.inv_area_rt[cy + jng], ixt1 = geo.inv_area_xt[cy + jng], ixt0 = geo.inv_area_xt[cy - 1 + jng];
    const double fr = ((double)a.re - (double)b.re) / irt + (double)c.re / ixt1 - (double)d.re / ixt0;
    const double fi = ((double)a.im - (double)b.im) / irt + (double)c.im / ixt1 - (double)d.im / ixt0;
    const long k = 2 * g.idx(cx, cy, m);
    const double dr = q1[k] - q0[k], di = q1[k + 1] - q0[k + 1];
    const double rr = dr + g.dt * fr - z[k], ri = di + g.dt * fi - z[k + 1];
    v[0] = fmax(v[0], sqrt(rr * rr + ri * ri));
    v[1] = fmax(v[1], sqrt(dr * dr + di * di));
    const double er = z[k] - zp[k], ei = z[k + 1] - zp[k + 1];
    v[2] = fmax(v[2], sqrt(er * er + ei * ei));
    v[3] = fmax(v[3], sqrt(z[k] * z[k] + z[k + 1] * z[k + 1]));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int t = 0; t < 4; ++t) v[t] = fmax(v[t], __shfl_xor(v[t], o));
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int t = 0; t < 4; ++t) red[t][wv] = v[t];
  __syncthreads();
  if (threadIdx.x < 4) {
    double r = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r = fmax(r, red[threadIdx.x][w]);
    atomicMax(out + 4 * m + threadIdx.x, (unsigned long long)__double_as_longlong(r));
  }
}

inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

template <class P>
void number_density(const Grid& g, const P& p, long n, double mc, cplx<double>* out, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_number_density<P>, dim3(nblk(n)), dim3(256), 0, s, g, p, n, mc, (double*)out);
}

template <class P>
void df_range(const Grid& g, const P& p, long n, double mc, const DistFn& d, double* part, int nblocks,
              hipStream_t s) {
  hipLaunchKernelGGL(k_df_range<P>, dim3(nblocks), dim3(256), 0, s, g, p, n, mc, d, part);
}

template <class P>
void df_bin(const Grid& g, const P& p, long n, double mc, const DistFn& d, double* out, int nblocks,
            hipStream_t s) {
  const long nbins = (long)d.resolution[0] * d.resolution[1] * d.resolution[2];
  if (nbins <= DF_LDS_BINS)
    hipLaunchKernelGGL((k_df_bin<P, true>), dim3(nblocks), dim3(256), 0, s, g, p, n, mc, d, out);
  else
    hipLaunchKernelGGL((k_df_bin<P, false>), dim3(nblocks), dim3(256), 0, s, g, p, n, mc, d, out);
}

}  // namespace

template <class P>
static void number_density_tiled(const Grid& g, const Tiles& tl, const P& p, const int* tile_start, ExtraRange xr,
                                 long n, cplx<double>* out, hipStream_t s) {
  const DensityOp<P> op{g, p};
  const size_t lds = (size_t)2 * g.M * (tl.tx + 4) * (tl.tr + 4) * sizeof(double);
  if (n <= 0) return;
  if (!tile_start || lds > 65536) {  // no tile ranges, or a box beyond the default LDS: one atomic per word
    number_density(g, p, n, 0.0, out, s);
    return;
  }
  switch (g.M) {
    case 1: launch_tile_scatter<2>(g, tl, op, p.key, tile_start, xr, (double*)out, s); break;
    case 2: launch_tile_scatter<4>(g, tl, op, p.key, tile_start, xr, (double*)out, s); break;
    case 3: launch_tile_scatter<6>(g, tl, op, p.key, tile_start, xr, (double*)out, s); break;
    default: number_density(g, p, n, 0.0, out, s); break;
  }
}
void launch_number_density_f64(const Grid& g, const Tiles& tl, const PartF64& p, const int* tile_start,
                               ExtraRange xr, long n, cplx<double>* out, hipStream_t s) {
  number_density_tiled(g, tl, p, tile_start, xr, n, out, s);
}
void launch_number_density_mixed(const Grid& g, const Tiles& tl, const PartMixed& p, const int* tile_start,
                                 ExtraRange xr, long n, cplx<double>* out, hipStream_t s) {
  number_density_tiled(g, tl, p, tile_start, xr, n, out, s);
}

void launch_charge_half_f64(const Grid& g, const Tiles& tl, const PartF64& p, const int* tile_start, ExtraRange xr,
                            long n, double mass, double q, double vs, double* out, hipStream_t s) {
  if (n <= 0) return;
  const size_t lds = (size_t)(tl.tx + 4) * (tl.tr + 4) * sizeof(double);
  if (tile_start && lds <= 65536)
    launch_tile_scatter<1>(g, tl, ChargeOp64{g, p, mass * C_LIGHT, q, vs}, p.key, tile_start, xr, out, s);
  else
    hipLaunchKernelGGL(k_charge_half<PartF64>, dim3(nblk(n)), dim3(256), 0, s, g, p, n, tl.ntiles, mass * C_LIGHT, q,
                       vs, out);
}
template <class T>
void launch_continuity_m0(const Grid& g, const double* q0, const double* q1, const cplx<T>* jx, const cplx<T>* jr,
                          const Geo& geo, unsigned long long* out, hipStream_t s) {
  const long n = (long)std::max(0, g.nx - 2 * g.ng) * std::max(0, g.ny - 2 * g.ng);
  if (n > 0)
    hipLaunchKernelGGL(k_continuity_m0<T>, dim3(std::min(nblk(n), 1024u)), dim3(256), 0, s, g, q0, q1, jx, jr, geo,
                       out);
}
// the charge set (nsets 1) or the charge, Z and Zp sets (3) of a check
template <class Part>
static void launch_cont_modes(const Grid& g, const Tiles& tl, long n, const int* key, const int* ts, ExtraRange xr,
                              const Part& part, int nsets, double* planes, long stride, hipStream_t s) {
  if (n <= 0) return;
  const CmGlobal out{planes, stride};
  const size_t lds = ts ? cm_tiles_lds(g, tl, nsets) : 0;
  if (lds)
    hipLaunchKernelGGL(k_cont_modes_tiles<Part>, dim3(tl.ntiles), dim3(256), lds, s, g, tl, key, ts, xr, nsets, part,
                       out);
  else
    hipLaunchKernelGGL(k_cont_modes_flat<Part>, dim3(nblk(n)), dim3(256), 0, s, g, n, tl.ntiles, key, ts, xr, part,
                       out);
}
void launch_cont_modes_begin_f64(const Grid& g, const Tiles& tl, const PartF64& p, long n, const int* ts,
                                 ExtraRange xr, double mass, double q, double* q0, hipStream_t s) {
  launch_cont_modes(g, tl, n, p.key, ts, xr, CmBeginF64{p, mass * C_LIGHT, q}, 1, q0, 0, s);
}
void launch_cont_modes_end_f64(const Grid& g, const Tiles& tl, const PartF64& p, long n, const int* ts,
                               ExtraRange xr, double mass, double q, double* planes, long stride, hipStream_t s) {
  launch_cont_modes(g, tl, n, p.key, ts, xr, CmEndF64{p, mass * C_LIGHT, q}, 3, planes, stride, s);
}
template <class T>
void launch_continuity_modes(const Grid& g, const double* q0, const double* q1, const double* z, const double* zp,
                             const cplx<T>* jx, const cplx<T>* jr, const Geo& geo, unsigned long long* out,
                             hipStream_t s) {
  const long n = (long)std::max(0, g.nx - 2 * g.ng) * std::max(0, g.ny - 2 * g.ng);
  if (n > 0)
    hipLaunchKernelGGL(k_continuity_modes<T>, dim3(std::min(nblk(n), 512u), g.M), dim3(256), 0, s, g, q0, q1, z, zp,
                       jx, jr, geo, out);
}
template void launch_continuity_modes<float>(const Grid&, const double*, const double*, const double*, const double*,
                                             const cplx<float>*, const cplx<float>*, const Geo&, unsigned long long*,
                                             hipStream_t);
template void launch_continuity_modes<double>(const Grid&, const double*, const double*, const double*,
                                              const double*, const cplx<double>*, const cplx<double>*, const Geo&,
                                              unsigned long long*, hipStream_t);

__global__ void k_zero_f64(double* __restrict__ p, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0.0;
}

void launch_zero_f64(double* p, long n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_f64, dim3(std::min(nblk(n), 8192u)), dim3(256), 0, s, p, n);
}

template <class T>
void launch_gauss_m0(const Grid& g, const double* qa, const double* qb, const cplx<T>* ex, const cplx<T>* er,
                     const Geo& geo, unsigned long long* out, hipStream_t s) {
  const long n = (long)std::max(0, g.nx - 2 * g.ng) * std::max(0, g.ny - 2 * g.ng);
  if (n > 0)
    hipLaunchKernelGGL(k_gauss_m0<T>, dim3(std::min(nblk(n), 1024u)), dim3(256), 0, s, g, qa, qb, ex, er, geo, out);
}
template void launch_gauss_m0<float>(const Grid&, const double*, const double*, const cplx<float>*,
                                     const cplx<float>*, const Geo&, unsigned long long*, hipStream_t);
template void launch_gauss_m0<double>(const Grid&, const double*, const double*, const cplx<double>*,
                                      const cplx<double>*, const Geo&, unsigned long long*, hipStream_t);
template void launch_continuity_m0<float>(const Grid&, const double*, const double*, const cplx<float>*,
                                          const cplx<float>*, const Geo&, unsigned long long*, hipStream_t);
template void launch_continuity_m0<double>(const Grid&, const double*, const double*, const cplx<double>*,
                                           const cplx<double>*, const Geo&, unsigned long long*, hipStream_t);

void launch_density_bcs(const Grid& g, cplx<double>* a, hipStream_t s) {
  const bool xref = g.bc_particle[XMIN] == BC_REFLECT || g.bc_particle[XMAX] == BC_REFLECT;
  const bool per = g.bc_particle[XMIN] == BC_PERIODIC && g.bc_particle[XMAX] == BC_PERIODIC;
  const unsigned brow = nblk((long)g.ny2 * g.M), bcol = nblk((long)g.nx2 * g.M);
  if (xref) hipLaunchKernelGGL(k_dens_reflect_x, dim3(brow), dim3(256), 0, s, g, a);
  if (g.bc_particle[YMAX] == BC_REFLECT) hipLaunchKernelGGL(k_dens_reflect_rmax, dim3(bcol), dim3(256), 0, s, g, a);
  if (per) hipLaunchKernelGGL(k_dens_periodic, dim3(brow), dim3(256), 0, s, g, a);
  hipLaunchKernelGGL(k_dens_zero_gradient_x, dim3(brow), dim3(256), 0, s, g, a);
  hipLaunchKernelGGL(k_dens_zero_gradient_y, dim3(bcol), dim3(256), 0, s, g, a);
}

void launch_df_range_f64(const Grid& g, const PartF64& p, long n, double mc, const DistFn& d, double* part,
                         int nblocks, hipStream_t s) {
  df_range(g, p, n, mc, d, part, nblocks, s);
}
void launch_df_range_mixed(const Grid& g, const PartMixed& p, long n, double mc, const DistFn& d, double* part,
                           int nblocks, hipStream_t s) {
  df_range(g, p, n, mc, d, part, nblocks, s);
}
void launch_df_bin_f64(const Grid& g, const PartF64& p, long n, double mc, const DistFn& d, double* out,
                       int nblocks, hipStream_t s) {
  df_bin(g, p, n, mc, d, out, nblocks, s);
}
void launch_df_bin_mixed(const Grid& g, const PartMixed& p, long n, double mc, const DistFn& d, double* out,
                         int nblocks, hipStream_t s) {
  df_bin(g, p, n, mc, d, out, nblocks, s);
}

}  // namespace cylpic
