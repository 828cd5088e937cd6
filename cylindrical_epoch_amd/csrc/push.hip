// Particle push + charge-conserving mode deposit + particle boundaries on
// gfx950 (push_particles, src/particles.F90:28-734; particle_bcs,
// src/boundary.F90:1541-1889).
//
// Work decomposition: one workgroup per tile of tile_x x tile_r cells.
// Particles are kept sorted by tile (stable), so a workgroup streams one
// contiguous slice of the SoA arrays (coalesced).  The current of the tile
// plus a DEPOSIT_HALO-cell border (the reach of a triangle-shape particle
// that started in the tile: 2 cells of shape + 1 cell of motion) is
// accumulated in LDS with LDS atomics and added to global J once per tile.
// A particle that drifted further (sorting less often than every step)
// deposits straight into global J with global atomics.  Fields are gathered
// from global memory: a tile's 3x3-stencil footprint is a few KB per mode
// and stays in L1/L2 across the tile's particles.
#include "cplx_ops.h"
#include "launch.h"
#include "migrate.h"

namespace cylpic {

// triangle/gx.inc: centred quadratic-spline weights (sum 2)
template <class T>
__device__ __forceinline__ void weights3(T frac, T& wm, T& w0, T& wp) {
  T cf2 = frac * frac;
  wm = ((T)0.25 + cf2) + frac;
  w0 = (T)1.5 - (T)2 * cf2;
  wp = ((T)0.25 + cf2) - frac;
}

// One 3x3 patch weighted by wx (x) and wy (r), e_part.inc order: F[k0] is
// the centre, rows `stride` apart (global mode plane or LDS field region).
template <class T, class S>
__device__ __forceinline__ cplx<T> patch(const cplx<S>* __restrict__ F, long k0, long stride, const T* wx,
                                         const T* wy) {
  cplx<T> acc;
#pragma unroll
  for (int jy = -1; jy <= 1; ++jy) {
    const long k = k0 + jy * stride;
    cplx<T> a = conv<T>(F[k - 1]), b = conv<T>(F[k]), c = conv<T>(F[k + 1]);
    cplx<T> row = cadd(cadd(rmul(wx[0], a), rmul(wx[1], b)), rmul(wx[2], c));
    acc = (jy == -1) ? rmul(wy[0], row) : cadd(acc, rmul(wy[jy + 1], row));
  }
  return acc;
}

// a[k + 2] for a run-time k in -2..2, without indexing a register array
template <class T>
__device__ __forceinline__ T sel5(const T* a, int k) {
  return k == -2 ? a[0] : (k == -1 ? a[1] : (k == 0 ? a[2] : (k == 1 ? a[3] : a[4])));
}


// ------------------------------------------------------------------------
// particle_bcs (one rank, boundary.F90:1541-1889) for particle ip at its new
// position (x, y, z), already stored; returns the particle's tile key, or
// ntiles if it left through an open boundary.
// The tile key is the particle's cell (src/housekeeping/split_particle.F90:
// 61-62, clamped into 1..nx, 1..ny) mapped to its tile.
// ------------------------------------------------------------------------
__device__ __forceinline__ int key_f64(const Grid& g, const Tiles& tl, double x, double y, double z,
                                       bool cell_key = false) {
  const double r = sqrt(y * y + z * z);
  const double fx = floor((x - g.x_grid_min) / g.dx + 1.5);
  const double fy = floor((r - g.y_grid_min) / g.dy + 1.5);
  const int cx = fx < -1.0e9 ? -1000000000 : (fx > 1.0e9 ? 1000000000 : (int)fx);
  const int cy = fy < -1.0e9 ? -1000000000 : (fy > 1.0e9 ? 1000000000 : (int)fy);
  return tl.key(cx, cy, g.nx, g.ny, cell_key);
}

__device__ __forceinline__ int bcs_f64(const Grid& g, const Tiles& tl, const PartF64& sp, long ip,
                                       double x, double y, double z) {
  const int png = 3;
  const double bshift_x = g.dx * (double)((1 + png + 0) / 2);
  const double x_min_outer = g.x_min - bshift_x, x_max_outer = g.x_max + bshift_x;
  const double x_shift = g.x_max - g.x_min;
  const double y_max_outer = g.y_max + g.dy * (double)((1 + png + 0) / 2);
  bool out_of_bounds = false;
  int bc = 0;
  const double part_pos = x;
  if (part_pos < g.x_min) {
    bc = g.bc_particle[XMIN];
    if (bc == BC_REFLECT) {
      x = 2.0 * g.x_min - part_pos;
      sp.x[ip] = x;
      sp.px[ip] = -sp.px[ip];
    } else if (bc == BC_PERIODIC) {
      x = part_pos - (-1.0) * x_shift;
      sp.x[ip] = x;
    }
    if (part_pos < x_min_outer && bc != BC_PERIODIC) out_of_bounds = true;
  }
  if (part_pos >= g.x_max) {
    bc = g.bc_particle[XMAX];
    if (bc == BC_REFLECT) {
      x = 2.0 * g.x_max - part_pos;
      sp.x[ip] = x;
      sp.px[ip] = -sp.px[ip];
    } else if (bc == BC_PERIODIC) {
      x = part_pos - 1.0 * x_shift;
      sp.x[ip] = x;
    }
    if (part_pos >= x_max_outer && bc != BC_PERIODIC) out_of_bounds = true;
  }
  const double r = sqrt(y * y + z * z);
  if (r >= g.y_max) {
    bc = g.bc_particle[YMAX];
    if (bc == BC_REFLECT) {
      const double radial_reduction = 2.0 * g.y_max / r - 1.0;
      y = y * radial_reduction;
      z = z * radial_reduction;
      sp.y[ip] = y;
      sp.z[ip] = z;
      const double inv_final_r = 1.0 / sqrt(y * y + z * z);
      const double cos_theta = y * inv_final_r, sin_theta = z * inv_final_r;
      const double py = sp.py[ip], pz = sp.pz[ip];
      const double part_pr = py * cos_theta + pz * sin_theta;
      const double part_pt = -(py * sin_theta) + pz * cos_theta;
      sp.py[ip] = -(part_pr * cos_theta) - part_pt * sin_theta;
      sp.pz[ip] = -(part_pr * sin_theta) + part_pt * cos_theta;
    }
    if (r >= y_max_outer && bc != BC_PERIODIC) out_of_bounds = true;
  }
  return out_of_bounds ? tl.ntiles : key_f64(g, tl, x, y, z);
}

// ------------------------------------------------------------------------
// FP64 push + deposit: the arithmetic of particles.F90:268-665 in the same
// operation order (this file is compiled without FP contraction).
//
// The deposit runs each mode's real and imaginary parts as two real passes.
// Every product and sum of the reference's complex deposit is componentwise
// (complex m-factors times real shape weights and real geometry), so the
// two passes evaluate exactly the operations of the complex loop, one
// component at a time.  That halves the running sums held in registers and
// lets the kernel run 3 wavefronts per SIMD (168 VGPRs) without spilling.
// ------------------------------------------------------------------------

// where a contribution goes: the workgroup's LDS box (planes [comp][m][part],
// part 0 the real words, 1 the imaginary words) or global J
struct LdsDep {
  double* lds;
  int x0, r0, LX, LR, M;
  __device__ __forceinline__ void add(int comp, int m, int part, int cx, int cy, double v) const {
    atomicAdd(lds + ((2 * (comp * M + m) + part) * LR + (cy - r0)) * LX + (cx - x0), v);
  }
};
struct GlobalDep {
  double* j[3];
  long plane;
  int nx2, off;  // off = (ng - 1) (nx2 + 1): Fortran (1 - ng, 1 - ng) at element 0
  __device__ __forceinline__ void add(int comp, int m, int part, int cx, int cy, double v) const {
    unsafeAtomicAdd(j[comp] + 2 * ((long)m * plane + (long)cy * nx2 + (cx + off)) + part, v);
  }
};

// per-particle inputs of the deposit loops
struct DepIn {
  int cell_x2, cell_y2, xmin, xmax, ymin, ymax, dcelly;
  double dtheta, fcx, fcy, fcz;
  cplx<double> ei05, eidth;
};

// particles.F90:586-665 for one particle: gx, gy the shape at t + dt/2 on
// slots -2..2 around (cell_x2, cell_y2), hx the change of the x shape, wyn
// the r weights at t + 3dt/2 (slots -1..1 around cell_y2 + dcelly; the r
// change is formed per row as the reference forms hy)
template <class Dep>
__device__ __forceinline__ void deposit_f64(const Dep& dep, const DepIn& d, const double* gx, const double* hx,
                                            const double* gy, const double* wyn, const Geo& geo, int jng, int M) {
  const double third = 1.0 / 3.0, sixth = 0.5 * third;
  // column slot j is ix = xmin + j: xmin is -2 or -1, so four slots hold
  // every particle's 3 or 4 columns in ascending ix
  double gxs[4], hxs[4];
  const bool from_m2 = d.xmin == -2;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    gxs[j] = from_m2 ? gx[j] : gx[j + 1];
    hxs[j] = from_m2 ? hx[j] : hx[j + 1];
  }
  cplx<double> exp_imtheta0 = mk(1.0, 0.0), exp_imdtheta = mk(1.0, 0.0);
  // m = 0: the reference's real factors 1, 1/2, 1/3 (x 1.0 is exact)
  cplx<double> m_fac_2 = mk(1.0, 0.0), m_fac_3 = mk(0.5, 0.0), m_fac_4 = mk(third, 0.0);
  double inv_mdth = 0.0, inv_m2dth2 = 0.0;
  for (int im = 0; im < M; ++im) {
    const double mdth = (double)im * d.dtheta;
    const double m2dth2 = mdth * mdth;
    const bool small_angle_change = fabs(mdth) < 1.0e-4;
    if (!small_angle_change && im > 0) {
      inv_mdth = 1.0 / mdth;
      inv_m2dth2 = inv_mdth * inv_mdth;
    }
    if (im > 0) {
      exp_imtheta0 = cmul(exp_imtheta0, d.ei05);
      exp_imdtheta = cmul(exp_imdtheta, d.eidth);
      cplx<double> m_fac_1;
      if (small_angle_change) {
        m_fac_1 = rmul(2.0, exp_imtheta0);
        m_fac_2 = cmul(m_fac_1, mk(1.0 - sixth * m2dth2, 0.5 * mdth));
        m_fac_3 = cmul(m_fac_1, mk(0.5 - 0.125 * m2dth2, third * mdth));
        m_fac_4 = cmul(m_fac_1, mk(third - 0.1 * m2dth2, 0.25 * mdth));
      } else {
        m_fac_1 = rmul(2.0 * inv_mdth, exp_imtheta0);
        m_fac_2 = cmul(m_fac_1, mk(exp_imdtheta.im, -(exp_imdtheta.re - 1.0)));
        const cplx<double> t = cmul(exp_imdtheta, mk(1.0, -mdth));
        m_fac_3 = cmul(m_fac_1, rmul(inv_mdth, mk(t.re - 1.0, t.im)));
        cplx<double> u = cmul(exp_imdtheta, mk(-m2dth2 + 2.0, -(2.0 * mdth)));
        u.re = u.re - 2.0;
        m_fac_4 = cmul(m_fac_1, imul(inv_m2dth2, u));
      }
    }
    // the m = 0 current has no imaginary part
    for (int part = 0; part < (im ? 2 : 1); ++part) {
      const double f2 = part ? m_fac_2.im : m_fac_2.re;
      const double f3 = part ? m_fac_3.im : m_fac_3.re;
      const double f4 = part ? m_fac_4.im : m_fac_4.re;
      // DO iy / DO ix of particles.F90:628-664, the column slots unrolled
      // so the running sums stay in registers
      double jyh[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) jyh[k] = 0.0;
      for (int iy = d.ymin; iy <= d.ymax; ++iy) {
        const int cy = d.cell_y2 + iy;
        const double gyv = sel5(gy, iy);
        const int oy = iy - d.dcelly;
        const double hyv = (oy == -1 ? wyn[0] : (oy == 0 ? wyn[1] : (oy == 1 ? wyn[2] : 0.0))) - gyv;
        const double w_rt = f2 * gyv + f3 * hyv;
        const double ym_fac_1 = f3 * gyv + f4 * hyv;
        const double fjx = d.fcx * geo.inv_area_rt[cy + jng];
        const double fjy = (d.fcy * hyv) * geo.inv_area_xt[cy + jng];
        const double fjz = d.fcz * geo.inv_volume[cy + jng];
        const double rat = geo.ratio_area_xt[cy + jng];
        double jxh = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (d.xmin + j > d.xmax) continue;
          const int cx = d.cell_x2 + d.xmin + j;
          const double gxv = gxs[j], hxv = hxs[j];
          const double w_xt = f2 * gxv + f3 * hxv;
          const double w_xr = gxv * w_rt + hxv * ym_fac_1;
          jxh = jxh - (fjx * hxv) * w_rt;
          jyh[j] = jyh[j] * rat - fjy * w_xt;
          dep.add(0, im, part, cx + 1, cy, jxh);
          dep.add(1, im, part, cx, cy + 1, jyh[j]);
          dep.add(2, im, part, cx, cy, fjz * w_xr);
        }
      }
    }
  }
}

// the tile's field region staged in LDS ([6][M][FR][FX], LDSF = 1)
struct FieldRegion {
  const cplx<double>* fl;
  int fx0, fr0, FX, FR;
  long fplane;
};
struct FieldsF64 {
  const cplx<double>* __restrict__ f[6];  // ex, er, et, bx, br, bt
};

// One particle: push, deposit into dep, particle boundaries and key.  With
// box_only, a particle whose stencil can leave the box (cells cell_x2 - 2
// .. cell_x2 + 3, the same in r: a particle in its tile never does) is left
// untouched and false returned.
template <int LDSF, class Dep>
__device__ __forceinline__ bool push_one_f64(const Grid& g, const Tiles& tl, const PushConst& pc, const PartF64& sp,
                                             int ip, const FieldsF64& F, const FieldRegion& fr, const Geo& geo,
                                             const Dep& dep, const LdsDep& box, bool box_only, int* err) {
  const int jng = g.ng;
  const double fac = 0.25;
  const int M = g.M;
  const double part_weight = sp.w[ip];
  double part_x = sp.x[ip], part_y = sp.y[ip], part_z = sp.z[ip];
  double part_ux = sp.px[ip] * pc.ipart_mc;
  double part_uy = sp.py[ip] * pc.ipart_mc;
  double part_uz = sp.pz[ip] * pc.ipart_mc;

  double gamma_rel = sqrt(((part_ux * part_ux + part_uy * part_uy) + part_uz * part_uz) + 1.0);
  double root = pc.dtco2 / gamma_rel;
  part_x = part_x + part_ux * root;
  part_y = part_y + part_uy * root;
  part_z = part_z + part_uz * root;

  const double part_x_local = part_x - g.x_grid_min;
  double part_r = sqrt(part_y * part_y + part_z * part_z);
  const double part_r_local = part_r - g.y_grid_min;
  double cell_x_r = part_x_local * pc.idx;
  double cell_y_r = part_r_local * pc.idy;
  int cell_x1 = (int)floor(cell_x_r + 0.5);
  double frac = (double)cell_x1 - cell_x_r;
  cell_x1 += 1;
  double gx[5], gy[5], hx[5], hy[5];
  gx[0] = gx[4] = gy[0] = gy[4] = 0.0;
  weights3(frac, gx[1], gx[2], gx[3]);
  int cell_y1 = (int)floor(cell_y_r + 0.5);
  frac = (double)cell_y1 - cell_y_r;
  cell_y1 += 1;
  weights3(frac, gy[1], gy[2], gy[3]);
  int cell_x2 = (int)floor(cell_x_r);
  double frac_x2 = ((double)cell_x2 - cell_x_r) + 0.5;
  cell_x2 += 1;
  int cell_y2 = (int)floor(cell_y_r);
  double frac_y2 = ((double)cell_y2 - cell_y_r) + 0.5;
  cell_y2 += 1;

  // every index the gather and deposit touch must stay inside the arrays
  if (cell_x1 < 2 - jng || cell_x1 > g.nx + jng - 1 || cell_y1 < 2 - jng || cell_y1 > g.ny + jng - 1 ||
      cell_x2 < -2 || cell_x2 > g.nx + 2 || cell_y2 < -2 || cell_y2 > g.ny + 2) {
    atomicOr(err, ERR_CELL_RANGE);
    return true;
  }
  if (box_only && !(cell_x2 - 2 >= box.x0 && cell_x2 + 3 < box.x0 + box.LX && cell_y2 - 2 >= box.r0 &&
                    cell_y2 + 3 < box.r0 + box.LR))
    return false;

  const cplx<double> emi = mk(part_y / part_r, (-part_z) / part_r);  // exp(-i theta) at t+dt/2
  const double theta_05 = atan2(part_z, part_y);
  // 1 / exp(-i theta) by Smith's algorithm
  cplx<double> ei05;
  if (fabs(emi.re) >= fabs(emi.im)) {
    double r = emi.im / emi.re, den = emi.re + emi.im * r;
    ei05 = mk(1.0 / den, -r / den);
  } else {
    double r = emi.re / emi.im, den = emi.re * r + emi.im;
    ei05 = mk(r / den, -1.0 / den);
  }
  hx[0] = hx[4] = hy[0] = hy[4] = 0.0;
  weights3(frac_x2, hx[1], hx[2], hx[3]);
  weights3(frac_y2, hy[1], hy[2], hy[3]);

  // e_part.inc / b_part.inc
  double ex_part = 0.0, er_part = 0.0, et_part = 0.0;
  double bx_part = 0.0, br_part = 0.0, bt_part = 0.0;
  {
    const bool in_lds = LDSF && min(cell_x1, cell_x2) - 1 >= fr.fx0 && max(cell_x1, cell_x2) + 1 < fr.fx0 + fr.FX &&
                        min(cell_y1, cell_y2) - 1 >= fr.fr0 && max(cell_y1, cell_y2) + 1 < fr.fr0 + fr.FR;
    // patch of component c (ex, er, et, bx, br, bt) of mode im centred on (cx, cy)
    auto P = [&](int c, int cx, int cy, int im, const double* wx, const double* wy) {
      if (in_lds)
        return patch<double>(fr.fl + (long)(c * M + im) * fr.fplane, (long)(cy - fr.fr0) * fr.FX + (cx - fr.fx0),
                             fr.FX, wx, wy);
      return patch<double>(F.f[c], g.idx(cx, cy, im), g.nx2, wx, wy);
    };
    cplx<double> e_m = mk(1.0, 0.0);
    for (int im = 0; im < M; ++im) {
      ex_part = ex_part + re_mul(e_m, P(0, cell_x1, cell_y2, im, gx + 1, hy + 1));
      er_part = er_part + re_mul(e_m, P(1, cell_x2, cell_y1, im, hx + 1, gy + 1));
      et_part = et_part + re_mul(e_m, P(2, cell_x2, cell_y2, im, hx + 1, hy + 1));
      e_m = cmul(e_m, emi);
    }
    e_m = mk(1.0, 0.0);
    for (int im = 0; im < M; ++im) {
      bx_part = bx_part + re_mul(e_m, P(3, cell_x2, cell_y1, im, hx + 1, gy + 1));
      br_part = br_part + re_mul(e_m, P(4, cell_x1, cell_y2, im, gx + 1, hy + 1));
      bt_part = bt_part + re_mul(e_m, P(5, cell_x1, cell_y1, im, gx + 1, gy + 1));
      e_m = cmul(e_m, emi);
    }
  }
  const double ey_part = er_part * emi.re + et_part * emi.im;
  const double ez_part = -(er_part * emi.im) + et_part * emi.re;
  const double by_part = br_part * emi.re + bt_part * emi.im;
  const double bz_part = -(br_part * emi.im) + bt_part * emi.re;

  // Boris rotation, particles.F90:404-451
  const double cmratio = pc.cmratio;
  const double uxm = part_ux + cmratio * ex_part;
  const double uym = part_uy + cmratio * ey_part;
  const double uzm = part_uz + cmratio * ez_part;
  gamma_rel = sqrt(((uxm * uxm + uym * uym) + uzm * uzm) + 1.0);
  root = pc.ccmratio / gamma_rel;
  const double taux = bx_part * root, tauy = by_part * root, tauz = bz_part * root;
  const double taux2 = taux * taux, tauy2 = tauy * tauy, tauz2 = tauz * tauz;
  const double tau = 1.0 / (((1.0 + taux2) + tauy2) + tauz2);
  const double uxp = ((((1.0 + taux2) - tauy2) - tauz2) * uxm +
                      2.0 * ((taux * tauy + tauz) * uym + (taux * tauz - tauy) * uzm)) * tau;
  const double uyp = ((((1.0 - taux2) + tauy2) - tauz2) * uym +
                      2.0 * ((tauy * tauz + taux) * uzm + (tauy * taux - tauz) * uxm)) * tau;
  const double uzp = ((((1.0 - taux2) - tauy2) + tauz2) * uzm +
                      2.0 * ((tauz * taux + tauy) * uxm + (tauz * tauy - taux) * uym)) * tau;
  part_ux = uxp + cmratio * ex_part;
  part_uy = uyp + cmratio * ey_part;
  part_uz = uzp + cmratio * ez_part;

  const double part_u2 = (part_ux * part_ux + part_uy * part_uy) + part_uz * part_uz;
  gamma_rel = sqrt(part_u2 + 1.0);
  const double igamma = 1.0 / gamma_rel;
  root = pc.dtco2 * igamma;
  const double delta_x = part_ux * root, delta_y = part_uy * root, delta_z = part_uz * root;
  part_x = part_x + delta_x;
  part_y = part_y + delta_y;
  part_z = part_z + delta_z;
  sp.x[ip] = part_x;
  sp.y[ip] = part_y;
  sp.z[ip] = part_z;
  sp.px[ip] = pc.part_mc * part_ux;
  sp.py[ip] = pc.part_mc * part_uy;
  sp.pz[ip] = pc.part_mc * part_uz;

  const double part_vy = (part_uy * C_LIGHT) * igamma;
  const double part_vz = (part_uz * C_LIGHT) * igamma;
  part_r = sqrt(part_y * part_y + part_z * part_z);
  const double part_vt = -(part_vy * (part_z / part_r)) + part_vz * (part_y / part_r);

  // current from the motion t+dt/2 -> t+3dt/2, particles.F90:508-584
  const double px15 = (part_x + delta_x) - g.x_grid_min;
  const double py15 = part_y + delta_y, pz15 = part_z + delta_z;
  const double r15 = sqrt(py15 * py15 + pz15 * pz15);
  const cplx<double> ei15 = mk(py15 / r15, pz15 / r15);
  const double theta_15 = atan2(pz15, py15);

  for (int k = 0; k < 5; ++k) {
    gx[k] = hx[k];
    gy[k] = hy[k];
  }
  cell_x_r = px15 * pc.idx;
  cell_y_r = (r15 - g.y_grid_min) * pc.idy;
  int cell_x3 = (int)floor(cell_x_r);
  const double frac_x3 = ((double)cell_x3 - cell_x_r) + 0.5;
  cell_x3 += 1;
  int cell_y3 = (int)floor(cell_y_r);
  const double frac_y3 = ((double)cell_y3 - cell_y_r) + 0.5;
  cell_y3 += 1;
  const int dcellx = cell_x3 - cell_x2, dcelly = cell_y3 - cell_y2;
  if (dcellx < -1 || dcellx > 1 || dcelly < -1 || dcelly > 1) {
    atomicOr(err, ERR_DCELL);
    return true;
  }
  double wyn[3];
  {
    double wm, w0, wp;
    weights3(frac_x3, wm, w0, wp);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const int o = k - 2 - dcellx;  // offset of slot k from the new centre
      const double h = o == -1 ? wm : (o == 0 ? w0 : (o == 1 ? wp : 0.0));
      hx[k] = h - gx[k];
    }
    weights3(frac_y3, wyn[0], wyn[1], wyn[2]);
  }

  DepIn d;
  d.cell_x2 = cell_x2;
  d.cell_y2 = cell_y2;
  d.xmin = -1 + (dcellx - 1) / 2;
  d.xmax = 1 + (dcellx + 1) / 2;
  d.ymin = -1 + (dcelly - 1) / 2;
  d.ymax = 1 + (dcelly + 1) / 2;
  d.dcelly = dcelly;
  d.dtheta = theta_15 - theta_05;
  const double q_weight_fac = (pc.part_q * part_weight) * fac;
  d.fcx = q_weight_fac * pc.idt;
  d.fcy = q_weight_fac * pc.idt;
  d.fcz = q_weight_fac * part_vt;
  d.ei05 = ei05;
  d.eidth = cmul(ei15, emi);
  deposit_f64(dep, d, gx, hx, gy, wyn, geo, jng, M);

  // particle_bcs (boundary.F90:1541-1889) on the new position, read back
  // from the arrays (fewer registers live across the deposit); it does not
  // read the currents, so it runs here rather than after the axis fold
  sp.key[ip] = bcs_f64(g, tl, sp, ip, sp.x[ip], sp.y[ip], sp.z[ip]);
  return true;
}

// LDSF: gather from the tile's fields staged in LDS (1) or from L1/L2 (0);
// the arithmetic is the same either way.  3 wavefronts per SIMD need at
// most 53 KB of LDS per workgroup (the box alone is 31.7 KB for 16 x 4
// tiles and 3 modes); with the staged fields (77 KB) LDS allows 2, and the
// register limit follows.  A particle whose stencil can leave the box is
// marked and left to k_push_deposit_f64_rest, so this kernel's deposit
// needs no per-add test and no 64-bit addresses.
constexpr int FIELD_HALO_F64 = 2;
template <int LDSF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LDSF ? 2 : 3))) void k_push_deposit_f64(
    Grid g, Tiles tl, PushConst pc, PartF64 sp, const int* __restrict__ tile_start, ExtraRange xr,
    FieldsF64 F, cplx<double>* __restrict__ jx, cplx<double>* __restrict__ jr, cplx<double>* __restrict__ jt,
    Geo geo, int* __restrict__ err, MigCount migc) {
  extern __shared__ double lds[];
  __shared__ uint32_t migs[MIG_DIRS];
  __shared__ int deferred;
  if (threadIdx.x == 0) deferred = 0;
  const int tile = tile_of_block(blockIdx.x, tl);
  const int ttx = tile % tl.ntx, tty = tile / tl.ntx;
  LdsDep box;
  box.lds = lds;
  box.x0 = tl.x0(ttx) - DEPOSIT_HALO;
  box.r0 = tty * tl.tr + 1 - DEPOSIT_HALO;
  box.LX = tl.tx + 2 * DEPOSIT_HALO;
  box.LR = tl.tr + 2 * DEPOSIT_HALO;
  box.M = g.M;
  const int nl = 3 * g.M * box.LR * box.LX;
  for (int e = threadIdx.x; e < 2 * nl; e += blockDim.x) lds[e] = 0.0;
  // field region [6][M][FR][FX] after the current box: the tile plus the
  // 2-cell reach of the gather stencils
  FieldRegion fr;
  fr.fl = (const cplx<double>*)(lds + 2 * nl);
  fr.fx0 = tl.x0(ttx) - FIELD_HALO_F64;
  fr.fr0 = tty * tl.tr + 1 - FIELD_HALO_F64;
  fr.FX = tl.tx + 2 * FIELD_HALO_F64;
  fr.FR = tl.tr + 2 * FIELD_HALO_F64;
  fr.fplane = (long)fr.FX * fr.FR;
  if (LDSF) {
    cplx<double>* fl = (cplx<double>*)(lds + 2 * nl);
    const int nf = 6 * g.M * fr.FR * fr.FX;
    for (int e = threadIdx.x; e < nf; e += blockDim.x) {
      int rem = e;
      const int fx = rem % fr.FX;
      rem /= fr.FX;
      const int fy = rem % fr.FR;
      rem /= fr.FR;
      const int m = rem % g.M;
      const int c = rem / g.M;
      const int cx = fr.fx0 + fx, cy = fr.fr0 + fy;
      cplx<double> v = mk(0.0, 0.0);
      if (cx >= 1 - g.ng && cx <= g.nx + g.ng && cy >= 1 - g.ng && cy <= g.ny + g.ng) v = F.f[c][g.idx(cx, cy, m)];
      fl[e] = v;
    }
  }
  __syncthreads();

  // the tile's particles: its main range, then its extra range
  const int p0 = tile_start[tile], p1 = tile_start[tile + 1];
  const int q0 = xr.start ? xr.base + xr.start[tile] : 0, q1 = xr.start ? xr.base + xr.start[tile + 1] : 0;
  const int np_main = p1 - p0, np = np_main + (q1 - q0);
  for (int k = threadIdx.x; k < np; k += blockDim.x) {
    const int ip = k < np_main ? p0 + k : q0 + (k - np_main);
    if (sp.key[ip] >= tl.ntiles) continue;  // removed since the last sort
    // a particle that moved away from its tile since the last sort is left
    // to k_push_deposit_f64_rest
    if (!push_one_f64<LDSF>(g, tl, pc, sp, ip, F, fr, geo, box, box, true, err)) {
      sp.key[ip] |= KEY_DEFERRED;
      deferred = 1;
    }
  }
  __syncthreads();
  // one pass of global atomics per tile
  const int ncell = box.LR * box.LX;
  const int M = g.M;
  for (int e = threadIdx.x; e < nl; e += blockDim.x) {
    const int wr = e + (e / ncell) * ncell;  // real word of complex cell e
    const double re = lds[wr], imv = lds[wr + ncell];
    if (re == 0.0 && imv == 0.0) continue;
    int rem = e;
    const int lx = rem % box.LX;
    rem /= box.LX;
    const int lr = rem % box.LR;
    rem /= box.LR;
    const int m = rem % M;
    const int comp = rem / M;
    const int cx = box.x0 + lx, cy = box.r0 + lr;
    if (cx < 1 - g.ng || cx > g.nx + g.ng || cy < 1 - g.ng || cy > g.ny + g.ng) continue;
    cplx<double>* J = comp == 0 ? jx : (comp == 1 ? jr : jt);
    double* p = (double*)(J + g.idx(cx, cy, m));
    if (re != 0.0) unsafeAtomicAdd(p, re);
    if (imv != 0.0) unsafeAtomicAdd(p + 1, imv);
  }
  // direction counts of the tile's new keys for the next migration sort,
  // here if every particle is done; otherwise the second pass counts, and
  // the removed-count slot marks the tile for it
  if (migc.counts) {
    __syncthreads();
    if (!deferred) {
      mig_count_tile((const uint32_t*)sp.key, tile, p0, p1, tl.ntx, tl.ntiles, migc.counts, migc.flag, migs);
      if (xr.start)
        mig_count_tile((const uint32_t*)sp.key, tile, q0, q1, tl.ntx, tl.ntiles, migc.counts_x, migc.flag, migs);
    } else if (threadIdx.x == 0) {
      migc.counts[(long)tile * MIG_DIRS + MIG_REMOVED] = TILE_DEFERRED;
    }
  }
}

// The particles the box kernel left (key marked KEY_DEFERRED), with their
// whole deposit into global J; then the direction counts of the tile's new
// keys for the next migration sort.  One workgroup per tile, the same tile
// order.  With migration counts, a tile the box kernel finished and counted
// is skipped after one read; without them every tile's keys are scanned.
__global__ __launch_bounds__(256) void k_push_deposit_f64_rest(
    Grid g, Tiles tl, PushConst pc, PartF64 sp, const int* __restrict__ tile_start, ExtraRange xr, FieldsF64 F,
    cplx<double>* __restrict__ jx, cplx<double>* __restrict__ jr, cplx<double>* __restrict__ jt, Geo geo,
    int* __restrict__ err, MigCount migc) {
  __shared__ uint32_t migs[MIG_DIRS];
  __shared__ uint32_t mark;
  const int tile = tile_of_block(blockIdx.x, tl);
  if (migc.counts) {
    if (threadIdx.x == 0) mark = migc.counts[(long)tile * MIG_DIRS + MIG_REMOVED];
    __syncthreads();
    if (mark != TILE_DEFERRED) return;
  }
  const int ttx = tile % tl.ntx, tty = tile / tl.ntx;
  LdsDep box;  // only its extent: the predicate is not evaluated here
  box.lds = nullptr;
  box.x0 = tl.x0(ttx) - DEPOSIT_HALO;
  box.r0 = tty * tl.tr + 1 - DEPOSIT_HALO;
  box.LX = tl.tx + 2 * DEPOSIT_HALO;
  box.LR = tl.tr + 2 * DEPOSIT_HALO;
  box.M = g.M;
  FieldRegion fr{};
  GlobalDep gdep;
  gdep.j[0] = (double*)jx;
  gdep.j[1] = (double*)jr;
  gdep.j[2] = (double*)jt;
  gdep.plane = g.plane;
  gdep.nx2 = g.nx2;
  gdep.off = (g.ng - 1) * (g.nx2 + 1);
  const int p0 = tile_start[tile], p1 = tile_start[tile + 1];
  const int q0 = xr.start ? xr.base + xr.start[tile] : 0, q1 = xr.start ? xr.base + xr.start[tile + 1] : 0;
  const int np_main = p1 - p0, np = np_main + (q1 - q0);
  for (int k = threadIdx.x; k < np; k += blockDim.x) {
    const int ip = k < np_main ? p0 + k : q0 + (k - np_main);
    const int key = sp.key[ip];
    if (!(key & KEY_DEFERRED)) continue;
    sp.key[ip] = key & ~KEY_DEFERRED;
    push_one_f64<0>(g, tl, pc, sp, ip, F, fr, geo, gdep, box, false, err);
  }
  if (migc.counts) {
    __syncthreads();
    mig_count_tile((const uint32_t*)sp.key, tile, p0, p1, tl.ntx, tl.ntiles, migc.counts, migc.flag, migs);
    if (xr.start)
      mig_count_tile((const uint32_t*)sp.key, tile, q0, q1, tl.ntx, tl.ntiles, migc.counts_x, migc.flag, migs);
  }
}

// moving window (remove_particles, window.F90:330-350): particles left of the
// moved x_min are removed
__global__ void k_window_f64(Grid g, PartF64 sp, long n, int ntiles) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (sp.x[i] < g.x_min) sp.key[i] = ntiles;
}

void launch_window_f64(const Grid& g, const PartF64& sp, long n, int ntiles, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_window_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, sp, n, ntiles);
}

// With cell_key: 4 x the tile-major cell key + the particle's half-cell
// quadrant (x, r: in the lower or upper half of its cell, which decides
// whether its deposit stencil starts at its cell or the one below)
__global__ void k_keys_f64(Grid g, Tiles tl, PartF64 sp, long n, int cell_key) {
  const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= n) return;
  // 2: re-key after pushes; removed particles keep sorting last
  if (cell_key == 2 && sp.key[ip] >= tl.ntiles) {
    sp.key[ip] = tl.ntiles * tl.tx * tl.tr * 4;
    return;
  }
  const double x = sp.x[ip], y = sp.y[ip], z = sp.z[ip];
  int k = key_f64(g, tl, x, y, z, cell_key != 0);
  if (cell_key) {
    const double fx = (x - g.x_grid_min) / g.dx + 1.5;
    const double fy = (sqrt(y * y + z * z) - g.y_grid_min) / g.dy + 1.5;
    k = 4 * k + 2 * (fy - floor(fy) < 0.5) + (fx - floor(fx) < 0.5);
  }
  sp.key[ip] = k;
}

// ------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------
size_t push_lds_bytes(const Grid& g, const Tiles& tl, int ldsf) {
  const size_t cells = (size_t)(tl.tx + 2 * DEPOSIT_HALO) * (size_t)(tl.tr + 2 * DEPOSIT_HALO);
  const size_t fcells = ldsf ? (size_t)(tl.tx + 2 * FIELD_HALO_F64) * (tl.tr + 2 * FIELD_HALO_F64) : 0;
  return (3 * cells + 6 * fcells) * (size_t)g.M * sizeof(cplx<double>);
}

void launch_push_deposit_f64(const Grid& g, const Tiles& tl, const PushConst& pc, const PartF64& sp,
                             const int* tile_start, ExtraRange xr, FieldSet<double>& F, const Geo& geo, int ldsf,
                             int* err, MigCount migc, hipStream_t s) {
  const size_t lds = push_lds_bytes(g, tl, ldsf);
  const FieldsF64 fs{{F.ex(), F.er(), F.et(), F.bx(), F.br(), F.bt()}};
  if (ldsf) {
    if (lds > 65536)
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_push_deposit_f64<1>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_push_deposit_f64<1>, dim3(tl.ntiles), dim3(256), lds, s, g, tl, pc, sp, tile_start, xr, fs,
                       F.jx(), F.jr(), F.jt(), geo, err, migc);
  } else {
    hipLaunchKernelGGL(k_push_deposit_f64<0>, dim3(tl.ntiles), dim3(256), lds, s, g, tl, pc, sp, tile_start, xr, fs,
                       F.jx(), F.jr(), F.jt(), geo, err, migc);
  }
  hipLaunchKernelGGL(k_push_deposit_f64_rest, dim3(tl.ntiles), dim3(256), 0, s, g, tl, pc, sp, tile_start, xr, fs,
                     F.jx(), F.jr(), F.jt(), geo, err, migc);
}

void launch_keys_f64(const Grid& g, const Tiles& tl, const PartF64& sp, long n, int cell_key, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_keys_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, tl, sp, n, cell_key);
}

}  // namespace cylpic
