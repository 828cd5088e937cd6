// Mixed-precision push + deposit on gfx950: the physics of push_particles
// (src/particles.F90:268-665) with
//   - particle x as an int32 cell (1..nx, nearest cell) plus an FP32 offset
//     from the cell centre in cells, y and z in FP32 units of dy, momenta as
//     FP32 u = p / (m c), FP32 weight;
//   - FP32 fields gathered from L1/L2, or staged per tile in LDS (6
//     components x modes, tile plus a 2-cell border: the reach of the 3x3
//     gather stencils);
//   - the current deposited into a per-tile LDS box (tile plus a 3-cell
//     border) of int64 fixed-point words by default, i.e. exact,
//     order-independent per-tile sums, flushed once per tile to the FP32
//     global J with global atomics;
//   - the deposit written as mode-independent real bases times per-mode
//     complex factors (the same algebra as the reference's loops, fewer
//     operations; the operation order differs, so results agree with the
//     FP64 path to a tolerance, not bit for bit);
//   - the m > 0 factors (particles.F90:589-626) evaluated by their power
//     series to x^8 for |m dtheta| < 0.7 (the exact expressions cancel
//     catastrophically in FP32 below that), by the exact expressions above;
//   - dtheta = arg(exp(i theta_15) exp(-i theta_05)) plus the reference's
//     +-2 pi jump when the two angles straddle the theta = pi cut (the
//     reference subtracts two ATAN2 values without unwrapping,
//     particles.F90:321-322, 522-524).
// A particle whose stencils leave the tile's LDS regions (it drifted while
// the sort ran less often than every step) gathers from and deposits to
// global memory directly.
#include <cstring>

#include "cont_modes.h"
#include "cplx_ops.h"
#include "launch.h"
#include "migrate.h"
#include "tile_scatter.h"

namespace cylpic {

// LDS regions around a tile.  A particle whose nearest cell c is in the tile
// gathers from cells c-2 .. c+2 and deposits into c-3 .. c+2 (the half-step
// cell is c-1 or c, the footprint spans -2..+2 around it, and the always-zero
// last Jx column and Jr row are not added).  In x the regions reach one cell
// further down, for the cell a moving-window shift moves every particle by
// between two sorts.
constexpr int FIELD_LO_X = 3, FIELD_LO_R = 2, FIELD_HI = 2;
constexpr int BOX_LO_X = 4, BOX_LO_R = 3, BOX_HI = 2;


struct MixConst {
  float kx, ky;           // c dt / (2 dx), c dt / (2 dy): half-step move in cells per unit u/gamma
  float cmratio, ccmratio;
  float q_fac_idt;        // q * fac / dt
  float q_fac_c;          // q * fac * c
  float yoff;             // y_grid_min / dy
  float nyf;              // y_max / dy
  int png_shift;          // (1 + png) / 2 outer-boundary margin in cells
};

struct GeoF {
  const float* rt;
  const float* xt;
  const float* vol;
  const float* rat;
};

__device__ __forceinline__ float fsel5(const float* a, int k) {
  return k == -2 ? a[0] : (k == -1 ? a[1] : (k == 0 ? a[2] : (k == 1 ? a[3] : a[4])));
}

__device__ __forceinline__ void w3f(float frac, float& wm, float& w0, float& wp) {
  const float cf2 = frac * frac;
  wm = (0.25f + cf2) + frac;
  w0 = 1.5f - 2.0f * cf2;
  wp = (0.25f + cf2) - frac;
}

// f2, f3, f4 of the m-factors: m_fac_k = 2 exp(i m theta0) f_k(m dtheta).
// f2 = (e^{ix}-1)/(ix), f3 = (e^{ix}(1-ix)-1)/x^2, f4 = i(e^{ix}(2-x^2-2ix)-2)/x^3
// The half push to t + dt/2 (particles.F90:296-312) in cell units, shared by
// the push and the continuity check so that both work with the same FP32
// values: the position, 1/r, r, the radial cell coordinate and the floors
// that place the staggered weights (hx_dcell.inc).
struct HalfStep {
  float oxh, yh, zh, irh, rh, cyr, fl2x, fl2y;
};
__device__ __forceinline__ HalfStep half_step(const MixConst& mc, float ox, float y, float z, float ux, float uy,
                                              float uz) {
  HalfStep h;
  const float ig = __builtin_amdgcn_rsqf(1.0f + (ux * ux + uy * uy + uz * uz));
  h.oxh = ox + ux * (mc.kx * ig);
  h.yh = y + uy * (mc.ky * ig);
  h.zh = z + uz * (mc.ky * ig);
  const float rh2 = h.yh * h.yh + h.zh * h.zh;
  h.irh = __builtin_amdgcn_rsqf(rh2);
  h.rh = rh2 * h.irh;
  h.cyr = h.rh - mc.yoff;
  h.fl2x = floorf(h.oxh);
  h.fl2y = floorf(h.cyr);
  return h;
}

__device__ __forceinline__ void mfactors(float x, cplx<float>& f2, cplx<float>& f3, cplx<float>& f4) {
  const float ax = fabsf(x);
  if (ax < 0.7f) {
    const float s = x * x;
    f2 = mk(1.0f + s * (-1.6666667e-1f + s * (8.3333333e-3f + s * (-1.9841270e-4f + s * 2.7557319e-6f))),
            x * (0.5f + s * (-4.1666667e-2f + s * (1.3888889e-3f + s * -2.4801587e-5f))));
    f3 = mk(0.5f + s * (-0.125f + s * (6.9444444e-3f + s * (-1.7361111e-4f + s * 2.4801587e-6f))),
            x * (3.3333333e-1f + s * (-3.3333333e-2f + s * (1.1904762e-3f + s * -2.2045855e-5f))));
    f4 = mk(3.3333333e-1f + s * (-0.1f + s * (5.9523810e-3f + s * (-1.5432099e-4f + s * 2.2546898e-6f))),
            x * (0.25f + s * (-2.7777778e-2f + s * (1.0416667e-3f + s * -1.9841270e-5f))));
  } else {
    float sn, cs;
    __sincosf(x, &sn, &cs);
    const float ix = 1.0f / x, ix2 = ix * ix;
    // e^{ix} - 1 = (cs - 1, sn) ; divide by i x
    f2 = mk(sn * ix, -(cs - 1.0f) * ix);
    // e^{ix}(1 - ix) - 1
    const cplx<float> t3 = mk(cs + sn * x - 1.0f, sn - cs * x);
    f3 = mk(t3.re * ix2, t3.im * ix2);
    // i (e^{ix}(2 - x^2 - 2ix) - 2) / x^3
    const float a = 2.0f - x * x, b = -2.0f * x;
    const cplx<float> u = mk(cs * a - sn * b - 2.0f, cs * b + sn * a);
    const float ix3 = ix2 * ix;
    f4 = mk(-u.im * ix3, u.re * ix3);
  }
}

template <int M>
struct Boxes {
  int fx0, fr0, FX, FR;  // field region origin (Fortran cell) and size
  int FP;                // field region row pitch (field_pitch)
  int jx0, jr0, LX, LR;  // current box origin and size
};

// Per-tile current accumulators (ACC):
//   0  FP32 LDS atomics;
//   1  FP32 with a 2Sum-compensated second word (the rounding error of every
//      add goes to a low word);
//   2  FP64 LDS atomics;
//   3  int64 fixed point per real word (ds_add_u64): every contribution is
//      truncated once to an integer multiple of 1/S and summed exactly, so the
//      tile sums do not depend on the order of the adds;
//   4  one int64 per complex cell holding two 32-bit fixed-point halves
//      (re * 2^32 + im, summed exactly as integers): half the atomics for
//      m > 0, 32-bit sums (S is then sized for the tile's particle count).
// The geometric factor of each current component depends only on the row it
// lands in (Jx: inv_area_rt(cy), Jr: inv_area_xt(cy - 1), Jt: inv_volume(cy);
// the reference's recursion jrh(iy) = jrh(iy-1) ratio_area_xt(iy) + ... is
// inv_area_xt(iy) times a running sum because ratio_area_xt(iy) =
// inv_area_xt(iy) / inv_area_xt(iy-1), particles.F90:209-216, 636-647), so it
// is applied once per box cell when the box is flushed.
// LDSF: gather the fields from the LDS-staged tile region (1), from L1/L2
// (0), or E from LDS and B from L1/L2 (2: half the LDS reads and region).
// Row pitch of the current box: the box's width, except that the int64 box
// (ACC 3) of 16-wide tiles uses 24 words.  A wavefront's lanes sit on the
// tile's 4 x 16 cells, so its 64-bit adds at one stencil offset touch four
// rows of 16 consecutive words; at a pitch of 24 words (48 banks) those rows
// start at banks 0, 48, 32 and 16, and every bank serves exactly two lanes.
__host__ __device__ inline int box_pitch(int acc, int tx) {
  const int w = tx + BOX_LO_X + BOX_HI;
  return (acc == 3 && tx == 16) ? 24 : w;
}
// int64 box planes per component (ACC 3): the m = 0 current has no
// imaginary part, so mode 0 keeps only its real plane
__host__ __device__ inline int acc3_planes(int M) { return 2 * M - 1; }

// Field region rows (LDS gather): a wavefront reads a stencil tap of 4 rows
// of 16 lanes (8 rows of 8 on 8-wide tiles), 8 bytes per lane.  A row pitch
// of P complex values with P = 8 (mod 16) starts the rows' 32-bank blocks at
// banks 0, 48, 32, 16: every bank serves two lanes, the minimum.  The gather
// uses only the real part of mode 0 (the reference's REAL(e^{-i0 theta} F_0)
// is Re F_0 for finite values), so mode 0 is staged as a real plane with the
// same pitch (in floats); per component the region holds that plane, then
// the complex planes of modes 1..M-1.
__host__ __device__ inline int field_pitch(int fx) { return fx + ((8 - fx % 16) + 16) % 16; }
__host__ __device__ inline long field_comp_floats(int M, int fr, int pitch) {
  return (long)fr * pitch * (2 * M - 1);
}

template <int ACC>
struct AccTraits {
  static constexpr bool fixed = ACC >= 3;
  // 32-bit LDS words per complex cell of the current box
  static constexpr int words = ACC == 0 ? 2 : (ACC == 4 ? 2 : 4);
};

// scale S (a power of two) with S * bound <= 2^30
__device__ __forceinline__ float fixed_scale(float bound) {
  if (!(bound > 0.0f)) return 1.0f;
  int e = ilogbf(1073741824.0f / bound);
  e = e < -126 ? -126 : (e > 126 ? 126 : e);
  return ldexpf(1.0f, e);
}

// TX, TR > 0: the tile shape fixed at compile time, so that the current box's
// pitch and plane strides are constants and every LDS atomic addresses one
// per-row base register plus an immediate offset; 0: the shape of `tl`.
// The push and deposit of one tile by the whole workgroup: the body of the
// per-run kernel (one workgroup per tile) and of the batched kernel (a
// persistent workgroup loops over (run, tile) items).  lds_raw is the
// dynamic LDS, wred and migs small static LDS arrays of the kernel.
template <int M, int ACC, int LDSF, int TX, int TR>
__device__ __forceinline__ void push_tile_mixed(
    const Grid& g, const Tiles& tl, const MixConst& mc, const PartMixed& sp, const int* __restrict__ tile_start,
    const ExtraRange& xr, const cplx<float>* __restrict__ ex, const cplx<float>* __restrict__ er,
    const cplx<float>* __restrict__ et, const cplx<float>* __restrict__ bx, const cplx<float>* __restrict__ br,
    const cplx<float>* __restrict__ bt, cplx<float>* __restrict__ jx, cplx<float>* __restrict__ jr,
    cplx<float>* __restrict__ jt, const GeoF& geo, int* __restrict__ err, const MigCount& migc, const int tile,
    double* lds_raw, float* wred, uint32_t* migs) {
  const int ttx = tile % tl.ntx, tty = tile / tl.ntx;
  Boxes<M> b;
  const int tx = TX > 0 ? TX : tl.tx, tr = TR > 0 ? TR : tl.tr;
  b.fx0 = ttx * tx + 1 - tl.shift - FIELD_LO_X;
  b.fr0 = tty * tr + 1 - FIELD_LO_R;
  b.FX = tx + FIELD_LO_X + FIELD_HI;
  b.FR = tr + FIELD_LO_R + FIELD_HI;
  b.FP = field_pitch(b.FX);
  b.jx0 = ttx * tx + 1 - tl.shift - BOX_LO_X;
  b.jr0 = tty * tr + 1 - BOX_LO_R;
  b.LX = box_pitch(ACC, tx);
  b.LR = tr + BOX_LO_R + BOX_HI;
  const int nj = 3 * M * b.LR * b.LX;      // complex cells of the current box
  constexpr int NFC = LDSF == 2 ? 3 : 6;              // components staged in LDS
  const int nf = LDSF ? NFC * M * b.FR * b.FX : 0;  // field values staged (complex, mode 0 real)
  const long fcomp = field_comp_floats(M, b.FR, b.FP);  // floats per staged component
  // current box: planes [3][M] of LR x LX cells; each plane holds the real
  // words then the imaginary words (FP64 for ACC 2, int64 for ACC 3, FP32 for
  // ACC 0, 1, where ACC 1 adds a second FP32 box for the low words), or one
  // packed int64 per cell (ACC 4).  ACC 3 stores mode 0's real words only.
  // Consecutive cells are consecutive words, so lanes at consecutive cells
  // hit distinct LDS banks.
  const int ncell = b.LR * b.LX;
  const int jmode = ACC == 4 ? ncell : 2 * ncell;                       // words per mode plane (m > 0)
  const int jcomp = ACC == 3 ? acc3_planes(M) * ncell : M * jmode;      // words per component
  // first word of mode m's real plane within a component
  auto mode_off = [&](int m) { return ACC == 3 ? (m == 0 ? 0 : (2 * m - 1) * ncell) : m * jmode; };
  double* jl = lds_raw;
  float* jf = (float*)lds_raw;
  float* jlo = jf + 2 * nj;
  unsigned long long* jq = (unsigned long long*)lds_raw;
  const int jwords = ACC == 3 ? 2 * 3 * jcomp : AccTraits<ACC>::words * nj;  // 32-bit words
  // field region: per component [FR][FP] floats (mode 0) then [M-1][FR][FP] complex
  float* fl = ((float*)lds_raw) + ((jwords + 1) & ~1);

  for (int e = threadIdx.x; e < jwords; e += blockDim.x) jf[e] = 0.0f;
  // the tile's particles: its main range, then its extra range
  const int p0 = tile_start[tile], p1 = tile_start[tile + 1];
  const int q0 = xr.start ? xr.base + xr.start[tile] : 0, q1 = xr.start ? xr.base + xr.start[tile + 1] : 0;
  const int np_main = p1 - p0, np = np_main + (q1 - q0);
  // fixed-point scales: |Jx|, |Jr| words of one particle stay below
  // 9 |q fac / dt| w and |Jt| words below 10.5 |q fac c| w (|Hx|, |sum hy| <= 2,
  // |m2| <= 2, |m3| <= 1, |m4| <= 2/3, shape weights <= 1.5, |v_theta| <= c)
  float sxr = 1.0f, st = 1.0f;
  if (AccTraits<ACC>::fixed) {
    float wm = 0.0f;
    for (int k = threadIdx.x; k < np; k += blockDim.x) wm = fmaxf(wm, fabsf(sp.w[k < np_main ? p0 + k : q0 + (k - np_main)]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) wm = fmaxf(wm, __shfl_xor(wm, o));
    if ((threadIdx.x & 63) == 0) wred[threadIdx.x >> 6] = wm;
    __syncthreads();
    wm = fmaxf(fmaxf(wred[0], wred[1]), fmaxf(wred[2], wred[3]));
    // ACC 4 sums in 32 bits: one contribution per particle per word
    const float cnt = ACC == 4 ? (float)np : 1.0f;
    sxr = fixed_scale(9.0f * fabsf(mc.q_fac_idt) * wm * cnt);
    st = fixed_scale(10.5f * fabsf(mc.q_fac_c) * wm * cnt);
  }
  const float isxr = 1.0f / sxr, ist = 1.0f / st;  // exact: powers of two
  if (LDSF) {
    for (int e = threadIdx.x; e < nf; e += blockDim.x) {
      int rem = e;
      const int fx = rem % b.FX;
      rem /= b.FX;
      const int fr = rem % b.FR;
      rem /= b.FR;
      const int m = rem % M;
      const int c = rem / M;
      const int cx = b.fx0 + fx, cy = b.fr0 + fr;
      cplx<float> v = mk(0.0f, 0.0f);
      const cplx<float>* src = c == 0 ? ex : (c == 1 ? er : (c == 2 ? et : (c == 3 ? bx : (c == 4 ? br : bt))));
      if (cx >= 1 - g.ng && cx <= g.nx + g.ng && cy >= 1 - g.ng && cy <= g.ny + g.ng) v = src[g.idx(cx, cy, m)];
      float* plane = fl + c * fcomp + (m == 0 ? 0 : (long)(2 * m - 1) * b.FR * b.FP);
      const int o = fr * b.FP + fx;
      if (m == 0) plane[o] = v.re;
      else ((cplx<float>*)plane)[o] = v;
    }
  }
  __syncthreads();

  const int jng = g.ng;

  for (int k = threadIdx.x; k < np; k += blockDim.x) {
    const int ip = k < np_main ? p0 + k : q0 + (k - np_main);
    if (sp.key[ip] >= tl.ntiles) continue;
    int cell = sp.cell[ip];
    const float ox = sp.ox[ip], y = sp.y[ip], z = sp.z[ip];
    float ux = sp.ux[ip], uy = sp.uy[ip], uz = sp.uz[ip];
    const float wgt = sp.w[ip];

    // half push to t + dt/2 (particles.F90:296-312), in cell units
    const HalfStep hs = half_step(mc, ox, y, z, ux, uy, uz);
    const float oxh = hs.oxh, yh = hs.yh, zh = hs.zh, irh = hs.irh, rh = hs.rh;
    const cplx<float> emi = mk(yh * irh, -(zh * irh));  // exp(-i theta_05)

    // weights at t + dt/2 (gx.inc, hx_dcell.inc) relative to the cell
    const float fl1x = floorf(oxh + 0.5f), fl2x = hs.fl2x;
    const int cx1 = cell + (int)fl1x, cx2 = cell + (int)fl2x;
    const float cyr = hs.cyr;
    const float fl1y = floorf(cyr + 0.5f), fl2y = hs.fl2y;
    const int cy1 = (int)fl1y + 1, cy2 = (int)fl2y + 1;
    float gx[5], gy[5], hx[5], hy[5];
    gx[0] = gx[4] = gy[0] = gy[4] = hx[0] = hx[4] = hy[0] = hy[4] = 0.0f;
    w3f(fl1x - oxh, gx[1], gx[2], gx[3]);
    w3f(fl1y - cyr, gy[1], gy[2], gy[3]);
    w3f((fl2x - oxh) + 0.5f, hx[1], hx[2], hx[3]);
    w3f((fl2y - cyr) + 0.5f, hy[1], hy[2], hy[3]);

    if (cx1 < 2 - jng || cx1 > g.nx + jng - 1 || cy1 < 2 - jng || cy1 > g.ny + jng - 1 || cx2 < -2 ||
        cx2 > g.nx + 2 || cy2 < -2 || cy2 > g.ny + 2) {
      atomicOr(err, ERR_CELL_RANGE);
      continue;
    }

    // gather (e_part.inc, b_part.inc): E and B as sum_m Re[e^{-im theta} S_m]
    float ef[6];
    {
      // 2D weights of the four stencils
      float Wgh[9], Whg[9], Whh[9], Wgg[9];
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          Wgh[3 * j + i] = gx[1 + i] * hy[1 + j];
          Whg[3 * j + i] = hx[1 + i] * gy[1 + j];
          Whh[3 * j + i] = hx[1 + i] * hy[1 + j];
          Wgg[3 * j + i] = gx[1 + i] * gy[1 + j];
        }
      const int lo_x = min(cx1, cx2) - 1, hi_x = max(cx1, cx2) + 1;
      const int lo_y = min(cy1, cy2) - 1, hi_y = max(cy1, cy2) + 1;
      const bool in_lds = LDSF && lo_x >= b.fx0 && hi_x < b.fx0 + b.FX && lo_y >= b.fr0 && hi_y < b.fr0 + b.FR;
      // stencil centre per component: ex (x1,y2,gh) er (x2,y1,hg) et (x2,y2,hh)
      //                               bx (x2,y1,hg) br (x1,y2,gh) bt (x1,y1,gg)
      const int ccx[6] = {cx1, cx2, cx2, cx2, cx1, cx1};
      const int ccy[6] = {cy2, cy1, cy2, cy1, cy2, cy1};
      const float* Ws[6] = {Wgh, Whg, Whh, Whg, Wgh, Wgg};
      const cplx<float>* gsrc[6] = {ex, er, et, bx, br, bt};
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const float* W = Ws[c];
        float acc = 0.0f;
        cplx<float> e_m = mk(1.0f, 0.0f);
#pragma unroll
        for (int m = 0; m < M; ++m) {
          cplx<float> S = mk(0.0f, 0.0f);
          if (in_lds && c < NFC) {
            const long o = (long)(ccy[c] - 1 - b.fr0) * b.FP + (ccx[c] - 1 - b.fx0);
            if (m == 0) {
              const float* base = fl + c * fcomp + o;
#pragma unroll
              for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int i = 0; i < 3; ++i) S.re = fmaf(W[3 * j + i], base[j * b.FP + i], S.re);
            } else {
              const cplx<float>* base = (const cplx<float>*)(fl + c * fcomp + (long)(2 * m - 1) * b.FR * b.FP) + o;
#pragma unroll
              for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                  const cplx<float> f = base[j * b.FP + i];
                  S.re = fmaf(W[3 * j + i], f.re, S.re);
                  S.im = fmaf(W[3 * j + i], f.im, S.im);
                }
            }
          } else {
            const cplx<float>* base = gsrc[c] + g.idx(ccx[c] - 1, ccy[c] - 1, m);
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
              for (int i = 0; i < 3; ++i) {
                const cplx<float> f = base[j * g.nx2 + i];
                S.re = fmaf(W[3 * j + i], f.re, S.re);
                if (m > 0) S.im = fmaf(W[3 * j + i], f.im, S.im);
              }
          }
          // mode 0: e_m = 1 and only Re S_0 enters
          acc = m == 0 ? acc + S.re : fmaf(e_m.re, S.re, fmaf(-e_m.im, S.im, acc));
          if (m + 1 < M) e_m = cmul(e_m, emi);
        }
        ef[c] = acc;
      }
    }
    const float ex_p = ef[0];
    const float ey_p = ef[1] * emi.re + ef[2] * emi.im;
    const float ez_p = -(ef[1] * emi.im) + ef[2] * emi.re;
    const float bx_p = ef[3];
    const float by_p = ef[4] * emi.re + ef[5] * emi.im;
    const float bz_p = -(ef[4] * emi.im) + ef[5] * emi.re;

    // Boris rotation (particles.F90:404-451)
    const float uxm = fmaf(mc.cmratio, ex_p, ux), uym = fmaf(mc.cmratio, ey_p, uy), uzm = fmaf(mc.cmratio, ez_p, uz);
    const float igm = __builtin_amdgcn_rsqf(1.0f + (uxm * uxm + uym * uym + uzm * uzm));
    const float root = mc.ccmratio * igm;
    const float tx = bx_p * root, ty = by_p * root, tz = bz_p * root;
    const float tx2 = tx * tx, ty2 = ty * ty, tz2 = tz * tz;
    const float tau = __builtin_amdgcn_rcpf(1.0f + tx2 + ty2 + tz2);
    const float uxp = ((1.0f + tx2 - ty2 - tz2) * uxm + 2.0f * ((tx * ty + tz) * uym + (tx * tz - ty) * uzm)) * tau;
    const float uyp = ((1.0f - tx2 + ty2 - tz2) * uym + 2.0f * ((ty * tz + tx) * uzm + (ty * tx - tz) * uxm)) * tau;
    const float uzp = ((1.0f - tx2 - ty2 + tz2) * uzm + 2.0f * ((tz * tx + ty) * uxm + (tz * ty - tx) * uym)) * tau;
    ux = fmaf(mc.cmratio, ex_p, uxp);
    uy = fmaf(mc.cmratio, ey_p, uyp);
    uz = fmaf(mc.cmratio, ez_p, uzp);
    const float ig = __builtin_amdgcn_rsqf(1.0f + (ux * ux + uy * uy + uz * uz));
    const float dxn = ux * (mc.kx * ig), dyn = uy * (mc.ky * ig), dzn = uz * (mc.ky * ig);
    float oxn = oxh + dxn;
    float yn = yh + dyn, zn = zh + dzn;

    // v_theta at t+dt (particles.F90:474-479): (-uy sin + uz cos) c / gamma
    const float irn = __builtin_amdgcn_rsqf(yn * yn + zn * zn);
    const float ut = (-(uy * zn) + uz * yn) * irn;
    // position at t + 3dt/2, relative to the cell
    const float ox15 = oxn + dxn;
    const float y15 = yn + dyn, z15 = zn + dzn;
    const float r152 = y15 * y15 + z15 * z15;
    const float ir15 = __builtin_amdgcn_rsqf(r152);
    const float r15 = r152 * ir15;
    const float c15 = y15 * ir15, s15 = z15 * ir15;
    const float c05 = emi.re, s05 = -emi.im;
    float dth = atan2f(s15 * c05 - c15 * s05, c15 * c05 + s15 * s05);
    {
      // theta_15 - theta_05 of two ATAN2 values in (-pi, pi]: the short arc
      // plus -+2 pi when it crosses the cut at theta = pi
      const bool h05 = !signbit(s05), h15 = !signbit(s15);  // theta >= 0
      const bool lt = (h15 != h05) ? !h15 : (c15 * s05 - s15 * c05 > 0.0f);
      const bool gt = (h15 != h05) ? h15 : (c05 * s15 - s05 * c15 > 0.0f);
      if (dth > 0.0f && lt) dth -= 6.2831853f;
      else if (dth < 0.0f && gt) dth += 6.2831853f;
    }

    // weight changes t+dt/2 -> t+3dt/2 (hx_dcell.inc)
    const float fl3x = floorf(ox15);
    const int dcx = (cell + (int)fl3x) - cx2;
    const float cyr15 = r15 - mc.yoff;
    const float fl3y = floorf(cyr15);
    const int dcy = ((int)fl3y + 1) - cy2;
    if (dcx < -1 || dcx > 1 || dcy < -1 || dcy > 1) {
      atomicOr(err, ERR_DCELL);
      continue;
    }
    {
      float wm, w0, wp;
      w3f((fl3x - ox15) + 0.5f, wm, w0, wp);
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const int o = k - 2 - dcx;
        const float h = o == -1 ? wm : (o == 0 ? w0 : (o == 1 ? wp : 0.0f));
        const float gold = hx[k];
        gx[k] = gold;
        hx[k] = h - gold;
      }
      w3f((fl3y - cyr15) + 0.5f, wm, w0, wp);
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const int o = k - 2 - dcy;
        const float h = o == -1 ? wm : (o == 0 ? w0 : (o == 1 ? wp : 0.0f));
        const float gold = hy[k];
        gy[k] = gold;
        hy[k] = h - gold;
      }
    }
    const int xmin = -1 + (dcx - 1) / 2, xmax = 1 + (dcx + 1) / 2;
    const int ymin = -1 + (dcy - 1) / 2, ymax = 1 + (dcy + 1) / 2;

    // per-mode factors m2, m3, m4 (m = 0: 1, 1/2, 1/3)
    cplx<float> m2[M], m3[M], m4[M];
    {
      m2[0] = mk(1.0f, 0.0f);
      m3[0] = mk(0.5f, 0.0f);
      m4[0] = mk(0.33333333f, 0.0f);
      const cplx<float> e05 = mk(c05, s05);
      cplx<float> e_m = mk(1.0f, 0.0f);
#pragma unroll
      for (int m = 1; m < M; ++m) {
        e_m = cmul(e_m, e05);
        cplx<float> f2, f3, f4;
        mfactors((float)m * dth, f2, f3, f4);
        const cplx<float> two_e = mk(2.0f * e_m.re, 2.0f * e_m.im);
        m2[m] = cmul(two_e, f2);
        m3[m] = cmul(two_e, f3);
        m4[m] = cmul(two_e, f4);
      }
    }

    // prefactors in box units (times S for the fixed-point boxes)
    const float fcx = (mc.q_fac_idt * sxr) * wgt;
    const float fcz = ((mc.q_fac_c * st) * wgt) * (ut * ig);
    // cumulative x sums of hx (jxh of the reference) are mode independent
    float Hx[5];
    {
      float run = 0.0f;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const bool on = (k - 2) >= xmin && (k - 2) <= xmax;
        run += on ? hx[k] : 0.0f;
        Hx[k] = run;
      }
    }
    // is the whole footprint inside the LDS current box?
    const bool jin = (cx2 - 2) >= b.jx0 && (cx2 + 2) < b.jx0 + b.LX && (cy2 - 2) >= b.jr0 && (cy2 + 2) < b.jr0 + b.LR;
    const int jb = (cy2 - b.jr0) * b.LX + (cx2 - b.jx0);      // box cell of (cx2, cy2), comp 0, mode 0
    // real (part 0) or imaginary (part 1) word of the cell at word index c (ACC 0..3)
    auto lds_add = [&](int c, int part, float v) {
      const int off = c + part * ncell;
      if (ACC == 3) {
        // truncation toward zero: the bias is below 2^-30 of the bound per add
        atomicAdd(jq + off, (unsigned long long)(long long)(int)v);
      } else if (ACC == 2) {
        atomicAdd(jl + off, (double)v);
      } else if (ACC == 0) {
        atomicAdd(jf + off, v);
      } else {
        // Knuth's TwoSum of the stored value and v; the error goes to the low word
        const float old = atomicAdd(jf + off, v);
        const float sum = old + v;
        const float bp = sum - old;
        const float err = (old - (sum - bp)) + (v - bp);
        if (err != 0.0f) atomicAdd(jlo + off, err);
      }
    };
    auto lds_add_c = [&](int c, float re, float im, bool has_im) {
      if (ACC == 4) {
        // round to nearest: the 32-bit unit is coarse (sized for the tile's
        // particle count), and rounding keeps the per-add errors unbiased
        const int qr = (int)rintf(re), qi = has_im ? (int)rintf(im) : 0;
        atomicAdd(jq + c, ((unsigned long long)(unsigned)qr << 32) + (unsigned long long)(long long)qi);
      } else {
        lds_add(c, 0, re);
        if (has_im) lds_add(c, 1, im);
      }
    };
    // Jx(cx+1, cy) += -fcx w_rt Hx ; Jr(cx, cy+1) += -w_xt R ; Jt(cx, cy) += fcz w_xr,
    // each still to be multiplied by its row's geometric factor.  The last
    // column's Jx and the last row's Jr are sums of all the weight changes,
    // zero by charge conservation (both shapes sum to 2): skipped, as FP32
    // would only add their round-off.
    if (jin) {
      float R = 0.0f;
      for (int iy = ymin; iy <= ymax; ++iy) {
        const float gyv = fsel5(gy, iy), hyv = fsel5(hy, iy);
        R = fmaf(fcx, hyv, R);
        // row factors: w_rt = m2 gy + m3 hy per mode; fcz gy, fcz hy
        cplx<float> wrt[M];
#pragma unroll
        for (int m = 0; m < M; ++m)
          wrt[m] = mk(m2[m].re * gyv + m3[m].re * hyv, m2[m].im * gyv + m3[m].im * hyv);
        const float ay = fcz * gyv, by = fcz * hyv;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          const int ix = k - 2;
          if (ix < xmin || ix > xmax) continue;
          const float gxv = gx[k], hxv = hx[k];
          const float A = ay * gxv, B = fmaf(by, gxv, ay * hxv), C = by * hxv;
          const float PX = -fcx * Hx[k];
          const int c = jb + iy * b.LX + ix;
#pragma unroll
          for (int m = 0; m < M; ++m) {
            const cplx<float> w_xt = mk(m2[m].re * gxv + m3[m].re * hxv, m2[m].im * gxv + m3[m].im * hxv);
            const int cm = c + mode_off(m);
            if (ix != xmax) lds_add_c(cm + 1, PX * wrt[m].re, PX * wrt[m].im, m != 0);      // (cx+1, cy)
            if (iy != ymax) lds_add_c(cm + jcomp + b.LX, -R * w_xt.re, -R * w_xt.im, m != 0);  // (cx, cy+1)
            lds_add_c(cm + 2 * jcomp, m2[m].re * A + m3[m].re * B + m4[m].re * C,            // (cx, cy)
                      m2[m].im * A + m3[m].im * B + m4[m].im * C, m != 0);
          }
        }
      }
    } else {
      // the footprint leaves the box (the particle drifted while sorts were
      // skipped): the same sums straight into the global FP32 J
      float R = 0.0f;
      for (int iy = ymin; iy <= ymax; ++iy) {
        const int cy = cy2 + iy;
        const float gyv = fsel5(gy, iy), hyv = fsel5(hy, iy);
        R = fmaf(fcx, hyv, R);
        const float gxf = geo.rt[cy + jng] * isxr, grf = geo.xt[cy + jng] * isxr, gtf = geo.vol[cy + jng] * ist;
        float hrun = 0.0f;  // Hx at column ix: running sum of hx over the footprint
#pragma unroll 1
        for (int ix = xmin; ix <= xmax; ++ix) {
          const float gxv = fsel5(gx, ix), hxv = fsel5(hx, ix);
          hrun += hxv;
          const float A = fcz * (gxv * gyv), B = fcz * (gxv * hyv + hxv * gyv), C = fcz * (hxv * hyv);
          const float PX = -fcx * hrun;
          const int cx = cx2 + ix;
#pragma unroll
          for (int m = 0; m < M; ++m) {
            const cplx<float> w_rt = mk(m2[m].re * gyv + m3[m].re * hyv, m2[m].im * gyv + m3[m].im * hyv);
            const cplx<float> w_xt = mk(m2[m].re * gxv + m3[m].re * hxv, m2[m].im * gxv + m3[m].im * hxv);
            const cplx<float> vt = mk(m2[m].re * A + m3[m].re * B + m4[m].re * C,
                                      m2[m].im * A + m3[m].im * B + m4[m].im * C);
            float* qx = (float*)(jx + g.idx(cx + 1, cy, m));
            float* qr = (float*)(jr + g.idx(cx, cy + 1, m));
            float* qt = (float*)(jt + g.idx(cx, cy, m));
            if (ix != xmax) unsafeAtomicAdd(qx, PX * w_rt.re * gxf);
            if (iy != ymax) unsafeAtomicAdd(qr, -R * w_xt.re * grf);
            unsafeAtomicAdd(qt, vt.re * gtf);
            if (m) {
              if (ix != xmax) unsafeAtomicAdd(qx + 1, PX * w_rt.im * gxf);
              if (iy != ymax) unsafeAtomicAdd(qr + 1, -R * w_xt.im * grf);
              unsafeAtomicAdd(qt + 1, vt.im * gtf);
            }
          }
        }
      }
    }

    // new position relative to its nearest cell; particle_bcs (boundary.F90:
    // 1541-1889) in cell units: periodic wraps are exact integer shifts
    const float sh = floorf(oxn + 0.5f);
    cell += (int)sh;
    oxn -= sh;
    bool out = false;
    if (cell < 1) {
      const int bc = g.bc_particle[XMIN];
      const float xr = (float)(cell - 1) + oxn;  // (x - x_grid_min) / dx
      if (bc == BC_PERIODIC) {
        cell += g.nx;
      } else {
        if (bc == BC_REFLECT) {
          ux = -ux;
          const float xn = -1.0f - xr;            // 2 x_min - x
          const float s2 = floorf(xn + 0.5f);
          cell = (int)s2 + 1;
          oxn = xn - s2;
        }
        if (xr < -0.5f - (float)mc.png_shift) out = true;
      }
    } else if (cell > g.nx) {
      const int bc = g.bc_particle[XMAX];
      const float xr = (float)(cell - 1) + oxn;
      if (bc == BC_PERIODIC) {
        cell -= g.nx;
      } else {
        if (bc == BC_REFLECT) {
          ux = -ux;
          const float xn = 2.0f * ((float)g.nx - 0.5f) - xr;
          const float s2 = floorf(xn + 0.5f);
          cell = (int)s2 + 1;
          oxn = xn - s2;
        }
        if (xr >= (float)g.nx - 0.5f + (float)mc.png_shift) out = true;
      }
    }
    const float rn = (yn * yn + zn * zn) * irn;
    if (rn >= mc.nyf) {
      const int bc = g.bc_particle[YMAX];
      if (bc == BC_REFLECT) {
        const float rr = 2.0f * mc.nyf / rn - 1.0f;
        yn *= rr;
        zn *= rr;
        const float ir2 = __builtin_amdgcn_rsqf(yn * yn + zn * zn);
        const float cs = yn * ir2, sn = zn * ir2;
        const float pr = uy * cs + uz * sn, pt = -(uy * sn) + uz * cs;
        uy = -(pr * cs) - pt * sn;
        uz = -(pr * sn) + pt * cs;
      }
      if (rn >= mc.nyf + (float)mc.png_shift) out = true;
    }
    sp.cell[ip] = cell;
    sp.ox[ip] = oxn;
    sp.y[ip] = yn;
    sp.z[ip] = zn;
    sp.ux[ip] = ux;
    sp.uy[ip] = uy;
    sp.uz[ip] = uz;
    int key = tl.ntiles;
    if (!out) {
      const float rk = sqrtf(yn * yn + zn * zn);
      const int cyk = (int)floorf(rk - mc.yoff + 1.5f);
      key = tl.key(cell, cyk, g.nx, g.ny, false);
    }
    sp.key[ip] = key;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nj; e += blockDim.x) {
    // complex cell e = plane * ncell + cell; real and imaginary words
    const int wr = ACC == 4 ? e : e + (e / ncell) * ncell, wi = wr + ncell;
    double re, imv;
    if (ACC == 4) {
      const unsigned long long q = jq[e];
      const int lo = (int)(unsigned)(q & 0xffffffffull);
      re = (double)((long long)(q - (unsigned long long)(long long)lo) >> 32);
      imv = (double)lo;
    } else if (ACC == 3) {
      const int plane = e / ncell, cell = e - plane * ncell;
      const int comp3 = plane / M, mm = plane - comp3 * M;
      const int w3 = comp3 * jcomp + mode_off(mm) + cell;
      re = (double)(long long)jq[w3];
      imv = mm == 0 ? 0.0 : (double)(long long)jq[w3 + ncell];
    } else if (ACC == 2) {
      re = jl[wr];
      imv = jl[wi];
    } else if (ACC == 0) {
      re = jf[wr];
      imv = jf[wi];
    } else {
      re = (double)jf[wr] + (double)jlo[wr];
      imv = (double)jf[wi] + (double)jlo[wi];
    }
    if (re == 0.0 && imv == 0.0) continue;
    int rem = e;
    const int lx = rem % b.LX;
    rem /= b.LX;
    const int lr = rem % b.LR;
    rem /= b.LR;
    const int m = rem % M;
    const int comp = rem / M;
    const int cx = b.jx0 + lx, cy = b.jr0 + lr;
    if (cx < 1 - g.ng || cx > g.nx + g.ng || cy < 1 - g.ng || cy > g.ny + g.ng) continue;
    // row factor: Jx inv_area_rt(cy), Jr inv_area_xt(cy - 1), Jt inv_volume(cy); undo S
    double fac;
    if (comp == 0) fac = (double)geo.rt[cy + jng] / (double)sxr;
    else if (comp == 1) fac = (double)geo.xt[cy - 1 + jng] / (double)sxr;
    else fac = (double)geo.vol[cy + jng] / (double)st;
    cplx<float>* J = comp == 0 ? jx : (comp == 1 ? jr : jt);
    float* p = (float*)(J + g.idx(cx, cy, m));
    if (re != 0.0) unsafeAtomicAdd(p, (float)(re * fac));
    if (imv != 0.0) unsafeAtomicAdd(p + 1, (float)(imv * fac));
  }
  // direction counts of the tile's new keys for the next migration sort
  if (migc.counts) {
    mig_count_tile((const uint32_t*)sp.key, tile, p0, p1, tl.ntx, tl.ntiles, migc.counts, migc.flag, migs);
    if (xr.start)
      mig_count_tile((const uint32_t*)sp.key, tile, q0, q1, tl.ntx, tl.ntiles, migc.counts_x, migc.flag, migs);
  }
}

// One run: one workgroup per tile, tiles mapped XCD-contiguously.
template <int M, int ACC, int LDSF, int TX = 0, int TR = 0>
__global__ __launch_bounds__(256) void k_push_deposit_mixed(
    Grid g, Tiles tl, MixConst mc, PartMixed sp, const int* __restrict__ tile_start, ExtraRange xr,
    const cplx<float>* __restrict__ ex, const cplx<float>* __restrict__ er, const cplx<float>* __restrict__ et,
    const cplx<float>* __restrict__ bx, const cplx<float>* __restrict__ br, const cplx<float>* __restrict__ bt,
    cplx<float>* __restrict__ jx, cplx<float>* __restrict__ jr, cplx<float>* __restrict__ jt, GeoF geo,
    int* __restrict__ err, MigCount migc) {
  extern __shared__ double lds_raw[];
  __shared__ float wred[8];
  __shared__ uint32_t migs[MIG_DIRS];
  push_tile_mixed<M, ACC, LDSF, TX, TR>(g, tl, mc, sp, tile_start, xr, ex, er, et, bx, br, bt, jx, jr, jt, geo, err,
                                        migc, tile_of_block(blockIdx.x, tl), lds_raw, wred, migs);
}

// The arguments of one (run, species) push for the batched kernel.
struct PushArgsMixed {
  Grid g;
  Tiles tl;
  MixConst mc;
  PartMixed sp;
  const int* tile_start;
  ExtraRange xr;
  const cplx<float>* f[6];  // ex er et bx br bt
  cplx<float>* j[3];        // jx jr jt
  GeoF geo;
  int* err;
  MigCount migc;
};

// Many runs (or species) that share the mode count and the tile shape, in one
// persistent launch: the work items are their tiles, descriptor d owning
// items [item_start[d], item_start[d + 1]).  The item range is cut into one
// contiguous chunk per XCD (workgroup b runs on XCD b mod 8), whose
// workgroups take items in order from the XCD's counter: neighbouring tiles
// share an L2.  A workgroup whose chunk is exhausted moves on to the next
// XCD's chunk (x + 1, x + 2, ... mod 8), so that uneven chunks (runs of
// different loads) do not leave XCDs idle; it exits once all eight counters
// have passed their ends.  counters[0..7] must be zero at launch.
template <int M, int ACC, int LDSF, int TX, int TR>
__global__ __launch_bounds__(256) void k_push_deposit_mixed_batch(const PushArgsMixed* __restrict__ descs,
                                                                  const int* __restrict__ item_start, int ndesc,
                                                                  int* __restrict__ counters) {
  extern __shared__ double lds_raw[];
  __shared__ float wred[8];
  __shared__ uint32_t migs[MIG_DIRS];
  __shared__ int s_item;
  const int total = item_start[ndesc];
  const int x0 = (int)(blockIdx.x % NUM_XCD);
  int tried = 0;  // chunks found exhausted (thread 0's count)
  for (;;) {
    __syncthreads();  // every work-item is done with the previous tile's LDS
    if (threadIdx.x == 0) {
      int got = -1;
      while (tried < NUM_XCD) {
        const int x = (x0 + tried) % NUM_XCD;
        const int hi = (int)((long)total * (x + 1) / NUM_XCD);
        const int it = (int)((long)total * x / NUM_XCD) + atomicAdd(counters + x, 1);
        if (it < hi) {
          got = it;
          break;
        }
        ++tried;
      }
      s_item = got;
    }
    __syncthreads();
    const int it = s_item;
    if (it < 0) break;
    int d = 0;
    while (d + 1 < ndesc && item_start[d + 1] <= it) ++d;
    d = __builtin_amdgcn_readfirstlane(d);
    const PushArgsMixed& a = descs[d];
    push_tile_mixed<M, ACC, LDSF, TX, TR>(a.g, a.tl, a.mc, a.sp, a.tile_start, a.xr, a.f[0], a.f[1], a.f[2], a.f[3],
                                          a.f[4], a.f[5], a.j[0], a.j[1], a.j[2], a.geo, a.err, a.migc,
                                          it - item_start[d], lds_raw, wred, migs);
  }
}

// ------------------------------------------------------------------------
// continuity check: the m = 0 charge q w hx hy / 4 of each particle at its
// half-step position, with the push's own FP32 half step and staggered
// weights (the charge whose motion the deposit's currents carry), summed
// in FP64 into one plane
__global__ __launch_bounds__(256) void k_charge_half_mixed(Grid g, MixConst mc, PartMixed sp, long n, int ntiles,
                                                           double qfac, float vs, double* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (unsigned)sp.key[i] >= (unsigned)ntiles) return;
  const HalfStep hs = half_step(mc, sp.ox[i], sp.y[i], sp.z[i], vs * sp.ux[i], vs * sp.uy[i], vs * sp.uz[i]);
  const int cx2 = sp.cell[i] + (int)hs.fl2x, cy2 = (int)hs.fl2y + 1;
  if (cx2 - 1 < 1 - g.ng || cx2 + 1 > g.nx + g.ng || cy2 - 1 < 1 - g.ng || cy2 + 1 > g.ny + g.ng) return;
  float hx[3], hy[3];
  w3f((hs.fl2x - hs.oxh) + 0.5f, hx[0], hx[1], hx[2]);
  w3f((hs.fl2y - hs.cyr) + 0.5f, hy[0], hy[1], hy[2]);
  const double qw = qfac * (double)sp.w[i];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      unsafeAtomicAdd(out + g.idx(cx2 + k - 1, cy2 + j - 1, 0), qw * ((double)hx[k] * (double)hy[j]));
}

// Continuity of every mode, mixed path: the charge modes at the push's own
// FP32 half step before the push ...
struct CmBeginMixed {
  MixConst mc;
  PartMixed sp;
  double q;
  template <class To>
  __device__ void operator()(const Grid& g, long i, const To& to) const {
    const HalfStep hs = half_step(mc, sp.ox[i], sp.y[i], sp.z[i], sp.ux[i], sp.uy[i], sp.uz[i]);
    const ContPos a{(double)(sp.cell[i] - 1) + (double)hs.oxh, (double)hs.cyr, (double)(hs.yh * hs.irh),
                    (double)(hs.zh * hs.irh), true};
    cm_charge(g, q * (double)sp.w[i], a, to, 0);
  }
};

// ... and after it, from the stored state t + dt and the push's FP32 half
// move (y, z in units of dy, x as cell plus offset)
struct CmEndMixed {
  MixConst mc;
  PartMixed sp;
  double q;
  template <class To>
  __device__ void operator()(const Grid& g, long i, const To& to) const {
    const float ux = sp.ux[i], uy = sp.uy[i], uz = sp.uz[i];
    const float ig = __builtin_amdgcn_rsqf(1.0f + (ux * ux + uy * uy + uz * uz));
    const float dxn = ux * (mc.kx * ig), dyn = uy * (mc.ky * ig), dzn = uz * (mc.ky * ig);
    const float ox = sp.ox[i], y = sp.y[i], zz = sp.z[i];
    // the radius and direction with the push's FP32 operations
    // (t + 3dt/2: exactly its values; t + dt/2: within one rounding)
    ContPos pos[2];
    for (int k = 0; k < 2; ++k) {
      const float oxs = k ? ox + dxn : ox - dxn, ys = k ? y + dyn : y - dyn, zs = k ? zz + dzn : zz - dzn;
      const float rs2 = ys * ys + zs * zs;
      const float irs = __builtin_amdgcn_rsqf(rs2);
      const float rs = rs2 * irs;
      pos[k] = ContPos{(double)(sp.cell[i] - 1) + (double)oxs, (double)(rs - mc.yoff), (double)(ys * irs),
                       (double)(zs * irs), true};
    }
    const double qw = q * (double)sp.w[i];
    cm_charge(g, qw, pos[1], to, 0);
    const double yd = y, zd = zz, r1 = sqrt(yd * yd + zd * zd);
    const double ud[3] = {ux, uy, uz};
    const double gam = sqrt(1.0 + (ud[0] * ud[0] + ud[1] * ud[1] + ud[2] * ud[2]));
    // v_theta dt / r in cells: u_theta / gamma (c dt / dy) / r, c dt / dy = 2 ky
    const double vtdt_r = r1 > 0.0 ? ((-(ud[1] * zd) + ud[2] * yd) / r1 / gam) * (2.0 * (double)mc.ky) / r1 : 0.0;
    cm_azimuthal(g, qw, pos[0], pos[1], vtdt_r, to);
  }
};

// one thread per slot (species without tile ranges)
template <class Part>
__global__ __launch_bounds__(256) void k_cont_modes_flat_mixed(Grid g, long n, int ntiles, const int* __restrict__ key,
                                                               const int* __restrict__ ts, ExtraRange xr, Part part,
                                                               CmGlobal out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !cm_live(i, ts, xr.start, xr.base, ntiles) || (unsigned)key[i] >= (unsigned)ntiles) return;
  part(g, i, out);
}

// ------------------------------------------------------------------------
// key of an uploaded particle (no push); with cell_key, 4 x the tile-major
// cell key + the half-cell quadrant (x offset below the centre, r in the
// lower half of its cell: the deposit stencil then starts one cell lower)
__global__ void k_keys_mixed(Grid g, Tiles tl, MixConst mc, PartMixed sp, long n, int cell_key) {
  const long ip = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= n) return;
  // 2: re-key after pushes; removed particles keep sorting last
  if (cell_key == 2 && sp.key[ip] >= tl.ntiles) {
    sp.key[ip] = tl.ntiles * tl.tx * tl.tr * 4;
    return;
  }
  const float y = sp.y[ip], z = sp.z[ip];
  const float fy = sqrtf(y * y + z * z) - mc.yoff + 1.5f;
  const int cyk = (int)floorf(fy);
  int k = tl.key(sp.cell[ip], cyk, g.nx, g.ny, cell_key != 0);
  if (cell_key) k = 4 * k + 2 * (fy - (float)cyk < 0.5f) + (sp.ox[ip] < 0.0f);
  sp.key[ip] = k;
}

// moving window (remove_particles, window.F90:330-350): the grid moved
// ncells cells towards +x, so the cell numbers drop by ncells; particles now
// left of x_min (cell < 1: x - x_grid_min = (cell - 1 + ox) dx with ox in
// [-1/2, 1/2)) are removed
__global__ void k_window_mixed(PartMixed sp, long n, int ncells, int ntiles) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = sp.cell[i] - ncells;
  sp.cell[i] = c;
  if (c < 1) sp.key[i] = ntiles;
}

void launch_window_mixed(const PartMixed& sp, long n, int ncells, int ntiles, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_window_mixed, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sp, n, ncells, ntiles);
}

// FP64 SI particles -> mixed representation
__global__ void k_to_mixed(Grid g, double mass, const double* x, const double* y, const double* z,
                           const double* px, const double* py, const double* pz, const double* w,
                           PartMixed sp, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xr = (x[i] - g.x_grid_min) / g.dx;
  double c = floor(xr + 0.5);
  c = c < -1.0e6 ? -1.0e6 : (c > 1.0e6 ? 1.0e6 : c);
  sp.cell[i] = (int)c + 1;
  sp.ox[i] = (float)(xr - c);
  sp.y[i] = (float)(y[i] / g.dy);
  sp.z[i] = (float)(z[i] / g.dy);
  const double imc = 1.0 / (mass * C_LIGHT);
  sp.ux[i] = (float)(px[i] * imc);
  sp.uy[i] = (float)(py[i] * imc);
  sp.uz[i] = (float)(pz[i] * imc);
  sp.w[i] = (float)w[i];
}

__global__ void k_from_mixed(Grid g, double mass, PartMixed sp, double* x, double* y, double* z, double* px,
                             double* py, double* pz, double* w, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = g.x_grid_min + ((double)(sp.cell[i] - 1) + (double)sp.ox[i]) * g.dx;
  y[i] = (double)sp.y[i] * g.dy;
  z[i] = (double)sp.z[i] * g.dy;
  const double mc = mass * C_LIGHT;
  px[i] = (double)sp.ux[i] * mc;
  py[i] = (double)sp.uy[i] * mc;
  pz[i] = (double)sp.uz[i] * mc;
  w[i] = (double)sp.w[i];
}

// ------------------------------------------------------------------------
size_t push_lds_bytes_mixed(const Grid& g, const Tiles& tl, int acc, int ldsf) {
  const size_t ncell = (size_t)(tl.tr + BOX_LO_R + BOX_HI) * box_pitch(acc, tl.tx);
  const size_t nj = 3 * (size_t)g.M * ncell;
  // staged field floats: per component a real mode-0 plane and M-1 complex planes
  const size_t nf = ldsf ? (ldsf == 2 ? 3 : 6) * (size_t)field_comp_floats(g.M, tl.tr + FIELD_LO_R + FIELD_HI,
                                                                             field_pitch(tl.tx + FIELD_LO_X + FIELD_HI))
                         : 0;
  const size_t jwords = acc == 3 ? 2 * 3 * acc3_planes(g.M) * ncell : ((acc == 0 || acc == 4) ? 2 * nj : 4 * nj);
  return ((jwords + 1) & ~(size_t)1) * sizeof(float) + nf * sizeof(float);
}

MixConst mix_const(const Grid& g, const PushConst& pc, double charge) {
  MixConst mc;
  const double dt = g.dt;
  mc.kx = (float)(C_LIGHT * dt / 2.0 / g.dx);
  mc.ky = (float)(C_LIGHT * dt / 2.0 / g.dy);
  mc.cmratio = (float)pc.cmratio;
  mc.ccmratio = (float)pc.ccmratio;
  mc.q_fac_idt = (float)(charge * 0.25 / dt);
  mc.q_fac_c = (float)(charge * 0.25 * C_LIGHT);
  mc.yoff = (float)(g.y_grid_min / g.dy);
  mc.nyf = (float)(g.y_max / g.dy);
  mc.png_shift = (1 + 3) / 2;
  return mc;
}

// the same charge as a tile-scatter operation (one real word per cell)
struct ChargeOpMixed {
  Grid g;
  MixConst mc;
  PartMixed sp;
  double qfac;
  float vs;  // the charge at x + vs v dt/2 (0: at x, a species that does not move)
  __device__ bool stencil(long i, int& cx0, int& cy0, double w[3][3], double* f) const {
    const HalfStep hs = half_step(mc, sp.ox[i], sp.y[i], sp.z[i], vs * sp.ux[i], vs * sp.uy[i], vs * sp.uz[i]);
    const int cx2 = sp.cell[i] + (int)hs.fl2x, cy2 = (int)hs.fl2y + 1;
    if (cx2 - 1 < 1 - g.ng || cx2 + 1 > g.nx + g.ng || cy2 - 1 < 1 - g.ng || cy2 + 1 > g.ny + g.ng) return false;
    float hx[3], hy[3];
    w3f((hs.fl2x - hs.oxh) + 0.5f, hx[0], hx[1], hx[2]);
    w3f((hs.fl2y - hs.cyr) + 0.5f, hy[0], hy[1], hy[2]);
    const double qw = qfac * (double)sp.w[i];
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a) w[b][a] = qw * ((double)hx[a] * (double)hy[b]);
    f[0] = 1.0;
    cx0 = cx2 - 1;
    cy0 = cy2 - 1;
    return true;
  }
  __device__ long out_index(const Grid& gg, int cx, int cy, int) const { return gg.idx(cx, cy, 0); }
};

template <class Part>
static void launch_cont_modes_mixed(const Grid& g, const Tiles& tl, long n, const int* key, const int* ts,
                                    ExtraRange xr, const Part& part, int nsets, double* planes, long stride,
                                    hipStream_t s) {
  if (n <= 0) return;
  const CmGlobal out{planes, stride};
  const size_t lds = ts ? cm_tiles_lds(g, tl, nsets) : 0;
  if (lds)
    hipLaunchKernelGGL(k_cont_modes_tiles<Part>, dim3(tl.ntiles), dim3(256), lds, s, g, tl, key, ts, xr, nsets, part,
                       out);
  else
    hipLaunchKernelGGL(k_cont_modes_flat_mixed<Part>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, n,
                       tl.ntiles, key, ts, xr, part, out);
}
void launch_cont_modes_begin_mixed(const Grid& g, const Tiles& tl, const PartMixed& sp, long n, const int* ts,
                                   ExtraRange xr, double q, double* q0, hipStream_t s) {
  launch_cont_modes_mixed(g, tl, n, sp.key, ts, xr, CmBeginMixed{mix_const(g, PushConst{}, q), sp, q}, 1, q0, 0, s);
}
void launch_cont_modes_end_mixed(const Grid& g, const Tiles& tl, const PartMixed& sp, long n, const int* ts,
                                 ExtraRange xr, double q, double* planes, long stride, hipStream_t s) {
  launch_cont_modes_mixed(g, tl, n, sp.key, ts, xr, CmEndMixed{mix_const(g, PushConst{}, q), sp, q}, 3, planes,
                          stride, s);
}

void launch_charge_half_mixed(const Grid& g, const Tiles& tl, const PartMixed& sp, const int* tile_start,
                              ExtraRange xr, long n, double charge, double vsd, double* out, hipStream_t s) {
  if (n <= 0) return;
  // only the half-step constants of mix_const are used
  const MixConst mc = mix_const(g, PushConst{}, charge);
  const float vs = (float)vsd;
  const size_t lds = (size_t)(tl.tx + 4) * (tl.tr + 4) * sizeof(double);
  if (tile_start && lds <= 65536)
    launch_tile_scatter<1>(g, tl, ChargeOpMixed{g, mc, sp, charge * 0.25, vs}, sp.key, tile_start, xr, out, s);
  else
    hipLaunchKernelGGL(k_charge_half_mixed, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, mc, sp, n,
                       tl.ntiles, charge * 0.25, vs, out);
}

int launch_push_deposit_mixed(const Grid& g, const Tiles& tl, const PushConst& pc, double charge,
                              const PartMixed& sp, const int* tile_start, ExtraRange xr, FieldSet<float>& F,
                              const float* geo, int ntab, int acc, int ldsf, int* err, MigCount migc,
                              hipStream_t s) {
  size_t lds = push_lds_bytes_mixed(g, tl, acc, ldsf);
  MixConst mc = mix_const(g, pc, charge);
  GeoF gf{geo, geo + ntab, geo + 2 * ntab, geo + 3 * ntab};
#define LAUNCH_T(MM, AA, LL, XX, RR)                                                                    \
  do {                                                                                                  \
    if (lds > 65536)                                                                                    \
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_push_deposit_mixed<MM, AA, LL, XX, RR>),  \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                  \
    hipLaunchKernelGGL((k_push_deposit_mixed<MM, AA, LL, XX, RR>), dim3(tl.ntiles), dim3(256), lds, s, g, tl, \
                       mc, sp, tile_start, xr, F.ex(), F.er(), F.et(), F.bx(), F.br(), F.bt(), F.jx(), F.jr(), F.jt(), \
                       gf, err, migc);                                                                  \
  } while (0)
#define LAUNCH(MM, AA, LL) LAUNCH_T(MM, AA, LL, 0, 0)
#define BY_MODE(AA, LL)            \
  switch (g.M) {                   \
    case 1: LAUNCH(1, AA, LL); break; \
    case 2: LAUNCH(2, AA, LL); break; \
    case 3: LAUNCH(3, AA, LL); break; \
    case 4: LAUNCH(4, AA, LL); break; \
    default: return -1;            \
  }
#define BY_MODE_T(AA, LL, XX, RR)                 \
  switch (g.M) {                                  \
    case 1: LAUNCH_T(1, AA, LL, XX, RR); break;   \
    case 2: LAUNCH_T(2, AA, LL, XX, RR); break;   \
    case 3: LAUNCH_T(3, AA, LL, XX, RR); break;   \
    case 4: LAUNCH_T(4, AA, LL, XX, RR); break;   \
    default: return -1;                           \
  }
#define BY_SHAPE_T(AA, LL)                                                   \
  if (tl.tx == 32 && tl.tr == 4) { BY_MODE_T(AA, LL, 32, 4); return 0; }     \
  if (tl.tx == 16 && tl.tr == 8) { BY_MODE_T(AA, LL, 16, 8); return 0; }     \
  if (tl.tx == 16 && tl.tr == 4) { BY_MODE_T(AA, LL, 16, 4); return 0; }     \
  if (tl.tx == 8 && tl.tr == 8) { BY_MODE_T(AA, LL, 8, 8); return 0; }       \
  if (tl.tx == 8 && tl.tr == 4) { BY_MODE_T(AA, LL, 8, 4); return 0; }
  // compile-time tile shapes for the fixed-point accumulators
  if (acc == 3 && ldsf == 0) { BY_SHAPE_T(3, 0) }
  if (acc == 3 && ldsf == 1) { BY_SHAPE_T(3, 1) }
  if (acc == 3 && ldsf == 2) { BY_SHAPE_T(3, 2) }
  if (acc == 4 && !ldsf) { BY_SHAPE_T(4, 0) }
  if (acc == 4 && ldsf) { BY_SHAPE_T(4, 1) }
  // the split gather (2) exists for the compile-time shapes only: stage
  // everything elsewhere
  if (ldsf == 2) lds = push_lds_bytes_mixed(g, tl, acc, 1);
  const int sel = acc * 2 + (ldsf ? 1 : 0);
  switch (sel) {
    case 0: BY_MODE(0, 0); break;
    case 1: BY_MODE(0, 1); break;
    case 2: BY_MODE(1, 0); break;
    case 3: BY_MODE(1, 1); break;
    case 4: BY_MODE(2, 0); break;
    case 5: BY_MODE(2, 1); break;
    case 6: BY_MODE(3, 0); break;
    case 7: BY_MODE(3, 1); break;
    case 8: BY_MODE(4, 0); break;
    case 9: BY_MODE(4, 1); break;
    default: return -1;
  }
#undef BY_SHAPE_T
#undef BY_MODE_T
#undef BY_MODE
#undef LAUNCH
#undef LAUNCH_T
  return 0;
}

// ---------------------------------------------------------------- batched push
size_t push_desc_mixed_bytes() { return sizeof(PushArgsMixed); }

void fill_push_desc_mixed(void* out, const Grid& g, const Tiles& tl, const PushConst& pc, double charge,
                          const PartMixed& sp, const int* tile_start, ExtraRange xr, FieldSet<float>& F,
                          const float* geo, int ntab, int* err, MigCount migc) {
  PushArgsMixed a;
  a.g = g;
  a.tl = tl;
  a.mc = mix_const(g, pc, charge);
  a.sp = sp;
  a.tile_start = tile_start;
  a.xr = xr;
  for (int k = 0; k < 6; ++k) a.f[k] = F.f[k];
  for (int k = 0; k < 3; ++k) a.j[k] = F.f[6 + k];
  a.geo = GeoF{geo, geo + ntab, geo + 2 * ntab, geo + 3 * ntab};
  a.err = err;
  a.migc = migc;
  std::memcpy(out, &a, sizeof(a));
}

bool batch_push_supported(const Tiles& tl, int M, int acc, int ldsf) {
  return tl.tx == 16 && tl.tr == 4 && acc == 3 && ldsf == 1 && M >= 1 && M <= 4;
}

template <int M>
static const void* batch_kernel() {
  return reinterpret_cast<const void*>(&k_push_deposit_mixed_batch<M, 3, 1, 16, 4>);
}

int push_batch_blocks(int M, size_t lds) {
  const void* k = M == 1 ? batch_kernel<1>() : (M == 2 ? batch_kernel<2>() : (M == 3 ? batch_kernel<3>() : batch_kernel<4>()));
  if (lds > 65536) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, lds) != hipSuccess || per_cu < 1) per_cu = 1;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
  return per_cu * cus;
}

void launch_push_deposit_mixed_batch(const void* descs, const int* item_start, int ndesc, int M, size_t lds,
                                     int* counters, int nblocks, hipStream_t s) {
  launch_zero_words(counters, NUM_XCD, s);
  const PushArgsMixed* d = (const PushArgsMixed*)descs;
#define LB(MM)                                                                                          \
  hipLaunchKernelGGL((k_push_deposit_mixed_batch<MM, 3, 1, 16, 4>), dim3(nblocks), dim3(256), lds, s, d, \
                     item_start, ndesc, counters)
  switch (M) {
    case 1: LB(1); break;
    case 2: LB(2); break;
    case 3: LB(3); break;
    default: LB(4); break;
  }
#undef LB
}

void launch_keys_mixed(const Grid& g, const Tiles& tl, const PartMixed& sp, long n, int cell_key, hipStream_t s) {
  if (n <= 0) return;
  PushConst pc{};
  MixConst mc = mix_const(g, pc, 0.0);
  hipLaunchKernelGGL(k_keys_mixed, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, tl, mc, sp, n,
                     cell_key);
}

void launch_to_mixed(const Grid& g, double mass, const double* const* src, const PartMixed& sp, long n,
                     hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_to_mixed, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, mass, src[0], src[1],
                     src[2], src[3], src[4], src[5], src[6], sp, n);
}

void launch_from_mixed(const Grid& g, double mass, const PartMixed& sp, double* const* dst, long n,
                       hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_from_mixed, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, mass, sp, dst[0], dst[1],
                     dst[2], dst[3], dst[4], dst[5], dst[6], n);
}

}  // namespace cylpic
