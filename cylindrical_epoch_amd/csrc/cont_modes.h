// Continuity of every azimuthal mode across one push (no counterpart in the
// reference, which checks nothing; the identity below is the one its deposit,
// particles.F90:508-665, satisfies).
//
// A particle moves from its half-step position a (t + dt/2) to b
// (t + 3dt/2) along s in [0, 1] with its shape weights S(s) = S_a + s H
// (H = S_b - S_a per dimension) and its angle theta_a + s dtheta (dtheta the
// difference of the two ATAN2 values, as the deposit takes it).  Its charge
// in mode m is q w S_x S_y c_m e^{i m theta} (c_0 = 1, c_m = 2), and
//   Q_m(b) - Q_m(a) = [the x-r part, carried by the deposited Jx, Jr]
//                   + Z_m,  Z_m = i m dtheta q w c_m e^{i m theta_a}
//                                 (f2 A + f3 B + f4 C),
// with A = S_ax S_ay, B = S_ax H_y + H_x S_ay, C = H_x H_y per cell and
// f_k(x) = int_0^1 s^(k-2) e^{i x s} ds at x = m dtheta (the reference's
// m-factors m_fac_k = 2 e^{i m theta_a} f_k).  So per cell and mode
//   Q_m(b) - Q_m(a) + dt (net x-r flux of J_m) = Z_m
// holds to round-off for the sum over particles; m = 0 has Z_0 = 0.
//
// The deposited J_theta carries v_theta(t + dt) where the charge's azimuthal
// motion needs r dtheta / dt: Zp_m, Z_m with dtheta replaced by
// v_theta dt / r (both at t + dt), is what the deposited J_theta divergence
// i m J_theta / r amounts to particle by particle, and |Z_m - Zp_m| measures
// the difference.  Per particle (Z - Zp) / Zp = r dtheta / (v_theta dt) - 1,
// the mean of r^2 / r(s)^2 - 1 over the straight step, which is
// (dt / r)^2 (3 v_r^2 - v_theta^2) / 12 to leading order: at most
// (v_perp dt / r)^2 / 4.
#pragma once

#include "launch.h"

namespace cylpic {

// one particle at one half step, in grid units: xr = (x - x_grid_min) / dx,
// rr = (r - y_grid_min) / dy, and its direction (cos, sin).  f32: the
// weights' polynomial in FP32 as the mixed push evaluates it (its positions
// make the fraction below exact in either precision).
struct ContPos {
  double xr, rr, c, s;
  bool f32;
};

// the check's quadratic weights (summing to one) of the three cells from
// floor(u): the staggered charge of the m = 0 continuity check
__device__ __forceinline__ int cm_weights(double u, double w[3], bool f32) {
  const double fc = floor(u);
  const double f = fc - u + 0.5;
  if (f32) {
    const float ff = (float)f;
    w[0] = 0.5 * (double)((0.25f + ff * ff) + ff);
    w[1] = 0.5 * (double)(1.5f - 2.0f * ff * ff);
    w[2] = 0.5 * (double)((0.25f + ff * ff) - ff);
  } else {
    w[0] = 0.5 * ((0.25 + f * f) + f);
    w[1] = 0.5 * (1.5 - 2.0 * f * f);
    w[2] = 0.5 * ((0.25 + f * f) - f);
  }
  return (int)fc;
}

// f2, f3, f4 at x: the series sum_n (i x)^n / (n! (n + k - 1)) below
// |x| = 0.5 (16 terms: error below 1e-18), else f2 = (e^{ix} - 1) / (ix),
// f3 = (e^{ix} - f2) / (ix), f4 = (e^{ix} - 2 f3) / (ix)
__device__ inline void cm_factors(double x, cplx<double>& f2, cplx<double>& f3, cplx<double>& f4) {
  if (fabs(x) < 0.5) {
    double tr = 1.0, ti = 0.0;  // (i x)^n / n!
    f2 = f3 = f4 = cplx<double>{0.0, 0.0};
    for (int n = 0; n < 16; ++n) {
      f2.re += tr / (n + 1);
      f2.im += ti / (n + 1);
      f3.re += tr / (n + 2);
      f3.im += ti / (n + 2);
      f4.re += tr / (n + 3);
      f4.im += ti / (n + 3);
      const double nr = -ti * x / (n + 1), ni = tr * x / (n + 1);
      tr = nr;
      ti = ni;
    }
    return;
  }
  double sn, cs;
  sincos(x, &sn, &cs);
  // 1 / (i x) = -i / x
  auto div_ix = [&](double re, double im) { return cplx<double>{im / x, -re / x}; };
  f2 = div_ix(cs - 1.0, sn);
  f3 = div_ix(cs - f2.re, sn - f2.im);
  f4 = div_ix(cs - 2.0 * f3.re, sn - 2.0 * f3.im);
}

// Is slot i a particle?  With tile ranges (ts non-null) the particles are
// the main ranges [0, ts[ntiles]) and the appended ranges [xbase + xs[0],
// xbase + xs[ntiles]) (xs null: none); the slots between hold stale copies
// that a sort left behind.  Without ranges every slot below n is one.
__device__ __forceinline__ bool cm_live(long i, const int* ts, const int* xs, int xbase, int ntiles) {
  if (!ts || i < ts[ntiles]) return true;
  return xs && i >= (long)xbase + xs[0] && i < (long)xbase + xs[ntiles];
}

__device__ __forceinline__ bool cm_inside(const Grid& g, int cx, int cy, int wx, int wy) {
  return cx >= 1 - g.ng && cx + wx - 1 <= g.nx + g.ng && cy >= 1 - g.ng && cy + wy - 1 <= g.ny + g.ng;
}

// Where a word of the check goes.  The check's planes form sets of M
// complex planes, `stride` words apart (set 0: the charge; at the end also
// set 1: Z, set 2: Zp); word part (0 real, 1 imaginary) of cell (cx, cy) of
// mode m is base + set * stride + 2 idx(cx, cy, m) + part.
struct CmGlobal {
  double* base;
  long stride;
  __device__ __forceinline__ void add(const Grid& g, int set, int cx, int cy, int m, int part, double v) const {
    unsafeAtomicAdd(base + set * stride + 2 * g.idx(cx, cy, m) + part, v);
  }
};
// ... or a workgroup's LDS box of its tile plus CM_HALO cells, words
// [set][m][part][BR][BX]; a cell outside the box goes to the planes
constexpr int CM_HALO = 3;
extern __shared__ double cm_box[];  // the dynamic LDS of k_cont_modes_tiles
struct CmBox {
  int bx0, br0, BX, BR, M;
  CmGlobal out;
  __device__ __forceinline__ void add(const Grid& g, int set, int cx, int cy, int m, int part, double v) const {
    const int lx = cx - bx0, lr = cy - br0;
    if ((unsigned)lx < (unsigned)BX && (unsigned)lr < (unsigned)BR)
      atomicAdd(&cm_box[((set * M + m) * 2 + part) * BR * BX + lr * BX + lx], v);
    else
      out.add(g, set, cx, cy, m, part, v);
  }
};

// charge modes of one particle at p into plane set `set`
template <class To>
__device__ inline void cm_charge(const Grid& g, double qw, const ContPos& p, const To& to, int set) {
  double wx[3], wy[3];
  const int cx = cm_weights(p.xr, wx, p.f32), cy = cm_weights(p.rr, wy, p.f32);
  if (!cm_inside(g, cx, cy, 3, 3)) return;
  double er = 1.0, ei = 0.0;  // e^{i m theta}
  for (int m = 0; m < g.M; ++m) {
    const double cm = m == 0 ? 1.0 : 2.0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double v = cm * qw * (wx[i] * wy[j]);
        to.add(g, set, cx + i, cy + j, m, 0, v * er);
        to.add(g, set, cx + i, cy + j, m, 1, v * ei);
      }
    const double nr = er * p.c - ei * p.s, ni = er * p.s + ei * p.c;
    er = nr;
    ei = ni;
  }
}

// Z_m (set 1) and Zp_m (set 2) of one particle moving from a to b;
// vtdt_r = v_theta dt / r at t + dt.  Returns false (nothing added) if the
// move spans more than one cell in x or r or leaves the arrays.
template <class To>
__device__ inline bool cm_azimuthal(const Grid& g, double qw, const ContPos& a, const ContPos& b, double vtdt_r,
                                    const To& to) {
  double wa[3], wb[3], ra[3], rb[3];
  const int ax = cm_weights(a.xr, wa, a.f32), bx = cm_weights(b.xr, wb, b.f32);
  const int ay = cm_weights(a.rr, ra, a.f32), by = cm_weights(b.rr, rb, b.f32);
  if (abs(ax - bx) > 1 || abs(ay - by) > 1) return false;
  const int x0 = min(ax, bx), y0 = min(ay, by);
  if (!cm_inside(g, x0, y0, 4, 4)) return false;
  double Sx[4] = {0, 0, 0, 0}, Hx[4] = {0, 0, 0, 0}, Sy[4] = {0, 0, 0, 0}, Hy[4] = {0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    Sx[ax - x0 + k] += wa[k];
    Hx[bx - x0 + k] += wb[k];
    Sy[ay - y0 + k] += ra[k];
    Hy[by - y0 + k] += rb[k];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    Hx[k] -= Sx[k];
    Hy[k] -= Sy[k];
  }
  const double dth = atan2(b.s, b.c) - atan2(a.s, a.c);
  double er = a.c, ei = a.s;  // e^{i m theta_a}, m = 1
  for (int m = 1; m < g.M; ++m) {
    cplx<double> f2, f3, f4;
    cm_factors((double)m * dth, f2, f3, f4);
    // i m (angle) q w 2 e^{i m theta_a}, for angle dtheta (Z) and v_theta dt / r (Zp)
    const double kz = 2.0 * (double)m * dth * qw, kp = 2.0 * (double)m * vtdt_r * qw;
    const double pr = -ei, pi = er;  // i e^{i m theta_a}
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const double A = Sx[i] * Sy[j], B = Sx[i] * Hy[j] + Hx[i] * Sy[j], C = Hx[i] * Hy[j];
        const double vr = f2.re * A + f3.re * B + f4.re * C, vi = f2.im * A + f3.im * B + f4.im * C;
        if (vr == 0.0 && vi == 0.0) continue;
        const double tr = pr * vr - pi * vi, ti = pr * vi + pi * vr;
        to.add(g, 1, x0 + i, y0 + j, m, 0, kz * tr);
        to.add(g, 1, x0 + i, y0 + j, m, 1, kz * ti);
        to.add(g, 2, x0 + i, y0 + j, m, 0, kp * tr);
        to.add(g, 2, x0 + i, y0 + j, m, 1, kp * ti);
      }
    const double nr = er * a.c - ei * a.s, ni = er * a.s + ei * a.c;
    er = nr;
    ei = ni;
  }
  return true;
}

// One workgroup per tile: the particles of the tile's ranges (its sorted
// range, then its appended range) add through `part(g, i, box)` into the
// LDS box of `nsets` plane sets, which is then added to the planes with
// one atomic per nonzero word.  The box holds 2 nsets M (tx + 6)(tr + 6)
// doubles.
template <class Part>
__global__ __launch_bounds__(256) void k_cont_modes_tiles(Grid g, Tiles tl, const int* __restrict__ key,
                                                          const int* __restrict__ tile_start, ExtraRange xr,
                                                          int nsets, Part part, CmGlobal out) {
  double* cmbox = cm_box;
  const int tile = tile_of_block(blockIdx.x, tl);
  const int ttx = tile % tl.ntx, tty = tile / tl.ntx;
  const CmBox b{tl.x0(ttx) - CM_HALO, tty * tl.tr + 1 - CM_HALO, tl.tx + 2 * CM_HALO, tl.tr + 2 * CM_HALO, g.M, out};
  const int ncell = b.BX * b.BR;
  const int nw = nsets * g.M * 2 * ncell;
  for (int e = threadIdx.x; e < nw; e += blockDim.x) cmbox[e] = 0.0;
  __syncthreads();
  const int p0 = tile_start[tile], p1 = tile_start[tile + 1];
  const int q0 = xr.start ? xr.base + xr.start[tile] : 0, q1 = xr.start ? xr.base + xr.start[tile + 1] : 0;
  const int np_main = p1 - p0, np = np_main + (q1 - q0);
  for (int k = threadIdx.x; k < np; k += blockDim.x) {
    const int i = k < np_main ? p0 + k : q0 + (k - np_main);
    if ((unsigned)key[i] >= (unsigned)tl.ntiles) continue;
    part(g, i, b);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nw; e += blockDim.x) {
    const double v = cmbox[e];
    if (v == 0.0) continue;
    int rem = e;
    const int lx = rem % b.BX;
    rem /= b.BX;
    const int lr = rem % b.BR;
    rem /= b.BR;
    const int w = rem % 2;
    rem /= 2;
    const int m = rem % g.M, set = rem / g.M;
    const int cx = b.bx0 + lx, cy = b.br0 + lr;
    if (cx < 1 - g.ng || cx > g.nx + g.ng || cy < 1 - g.ng || cy > g.ny + g.ng) continue;
    out.add(g, set, cx, cy, m, w, v);
  }
}

// the LDS the tiled check needs, or 0 when it exceeds the default 64 KB
// (the per-particle kernels run instead)
inline size_t cm_tiles_lds(const Grid& g, const Tiles& tl, int nsets) {
  const size_t b = (size_t)nsets * g.M * 2 * (tl.tx + 2 * CM_HALO) * (tl.tr + 2 * CM_HALO) * sizeof(double);
  return b <= 65536 ? b : 0;
}

}  // namespace cylpic
