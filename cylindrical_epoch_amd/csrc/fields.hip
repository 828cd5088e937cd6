// Per-mode FDTD field solver, field boundaries, current boundaries and the
// mode-energy reduction on gfx950.  T = double (FP64 path) or float (mixed
// path: FP32 storage, the r=0 axis rows computed in FP64).
//
// All field arrays have the Fortran layout f(1-ng:nx+ng, 1-ng:ny+ng,
// 0:n_mode-1), x fastest: a block of 256 consecutive ix threads reads and
// writes 256 consecutive complex values (4 KB FP64, 2 KB FP32) per row.
#include "cplx_ops.h"
#include "launch.h"

namespace cylpic {

// ------------------------------------------------------------------------
// update_e_field interior, fields.f90:62-113.  One thread per (ix, ir, m),
// ix in 0..nx, ir in 1..ny (the r=0 axis is local to the single rank).
// ------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_update_e(Grid g, cplx<T>* __restrict__ ex,
                                                  cplx<T>* __restrict__ er, cplx<T>* __restrict__ et,
                                                  const cplx<T>* __restrict__ bx,
                                                  const cplx<T>* __restrict__ br,
                                                  const cplx<T>* __restrict__ bt,
                                                  const cplx<T>* __restrict__ jx,
                                                  const cplx<T>* __restrict__ jr,
                                                  const cplx<T>* __restrict__ jt) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int ir = blockIdx.y + 1;
  const int m = blockIdx.z;
  if (ix > g.nx) return;
  const double c2d = C_LIGHT * C_LIGHT;
  const double r_d = fabs((double)(ir - 1) * g.dy + g.y_grid_min);
  const double r_p = r_d + 0.5 * g.dy;
  const double fac_xd = c2d / r_p;
  const T fac_x = (T)fac_xd;
  const T a_x = (T)((double)m * fac_xd);
  const T a_r = (T)(((double)m * c2d) / r_d);
  const T c2 = (T)c2d, dx = (T)g.dx, dy = (T)g.dy, dt = (T)g.dt, eps0 = (T)EPSILON0;
  const T half = (T)0.5;

  const long k = g.idx(ix, ir, m);
  const long kr1 = k + g.nx2;
  const long kx1 = k + 1;
  const cplx<T> bt0 = bt[k], bt_r1 = bt[kr1], bt_x1 = bt[kx1];
  const cplx<T> br0 = br[k], br_x1 = br[kx1];
  const cplx<T> bx0 = bx[k], bx_r1 = bx[kr1];
  // (11) exm
  {
    cplx<T> t1 = rmul(fac_x * half, cadd(bt_r1, bt0));
    cplx<T> t2 = imul(a_x, br0);
    cplx<T> t3 = cdivr(rmul(c2, csub(bt_r1, bt0)), dy);
    cplx<T> t4 = cdivr(jx[k], eps0);
    cplx<T> s = csub(cadd(cadd(t1, t2), t3), t4);
    ex[k] = cadd(ex[k], cmulr(cmulr(s, half), dt));
  }
  // (9) erm
  {
    cplx<T> t1 = cneg(imul(a_r, bx0));
    cplx<T> t2 = cdivr(rmul(c2, csub(bt_x1, bt0)), dx);
    cplx<T> t3 = cdivr(jr[k], eps0);
    cplx<T> s = csub(csub(t1, t2), t3);
    er[k] = cadd(er[k], cmulr(cmulr(s, half), dt));
  }
  // (10) etm
  {
    cplx<T> t1 = cdivr(rmul(c2, csub(br_x1, br0)), dx);
    cplx<T> t2 = cdivr(rmul(c2, csub(bx_r1, bx0)), dy);
    cplx<T> t3 = cdivr(jt[k], eps0);
    cplx<T> s = csub(csub(t1, t2), t3);
    et[k] = cadd(et[k], cmulr(cmulr(s, half), dt));
  }
}

// r=0 rows and the reflected ghost rows of update_e_field, fields.f90:116-180.
// One thread per column ix in 1-ng..nx+ng, all modes (sequential in m).
// Arithmetic in FP64 whatever the storage type.
template <class T>
__global__ __launch_bounds__(256) void k_e_axis(Grid g, cplx<T>* __restrict__ ex,
                                                cplx<T>* __restrict__ er, cplx<T>* __restrict__ et,
                                                const cplx<T>* __restrict__ bt,
                                                const cplx<T>* __restrict__ jx) {
  const int ix = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1 - g.ng;
  if (ix > g.nx + g.ng) return;
  const int ng = g.ng;
  const double c2 = C_LIGHT * C_LIGHT;
  const double dt = g.dt, dy = g.dy;
  {  // m = 0
    const int m = 0;
    long k0 = g.idx(ix, 0, m), k1 = g.idx(ix, 1, m);
    cplx<double> s = csub(rmul((4.0 * c2) / dy, conv<double>(bt[k1])), cdivr(conv<double>(jx[k0]), EPSILON0));
    ex[k0] = conv<T>(cadd(conv<double>(ex[k0]), cmulr(cmulr(s, 0.5), dt)));
    et[k0] = czero<T>();
    er[k0] = cneg(er[k1]);
    for (int ir = 1 - ng; ir <= -1; ++ir) {
      et[g.idx(ix, ir, m)] = cneg(et[g.idx(ix, -ir, m)]);
      er[g.idx(ix, ir, m)] = cneg(er[g.idx(ix, -ir + 1, m)]);
      ex[g.idx(ix, ir, m)] = ex[g.idx(ix, -ir, m)];
    }
  }
  if (g.M > 1) {
    const int m = 1;
    long k0 = g.idx(ix, 0, m), k1 = g.idx(ix, 1, m), k2 = g.idx(ix, 2, m);
    ex[k0] = czero<T>();
    cplx<double> et0 = conv<double>(et[k0]);
    cplx<double> er1 = conv<double>(er[k1]), er2 = conv<double>(er[k2]);
    er[k0] = conv<T>(csub(imul(2.0, et0), er1));
    cplx<double> w = csub(rmul(9.0, er1), er2);
    et[k0] = conv<T>(mk(0.125 * w.im, -(0.125 * w.re)));
    for (int ir = 1 - ng; ir <= -1; ++ir) {
      et[g.idx(ix, ir, m)] = et[g.idx(ix, -ir, m)];
      er[g.idx(ix, ir, m)] = er[g.idx(ix, -ir + 1, m)];
      ex[g.idx(ix, ir, m)] = cneg(ex[g.idx(ix, -ir, m)]);
    }
  }
  double mode_sign = 1.0;
  for (int m = 2; m < g.M; ++m) {
    long k0 = g.idx(ix, 0, m), k1 = g.idx(ix, 1, m), k2 = g.idx(ix, 2, m);
    ex[k0] = czero<T>();
    et[k0] = czero<T>();
    er[k0] = cneg(er[k1]);
    er[k1] = conv<T>(cdivr(conv<double>(er[k2]), 9.0));
    for (int ir = 1 - ng; ir <= -1; ++ir) {
      et[g.idx(ix, ir, m)] = conv<T>(cneg(rmul(mode_sign, conv<double>(et[g.idx(ix, -ir, m)]))));
      er[g.idx(ix, ir, m)] = conv<T>(cneg(rmul(mode_sign, conv<double>(er[g.idx(ix, -ir + 1, m)]))));
      ex[g.idx(ix, ir, m)] = conv<T>(rmul(mode_sign, conv<double>(ex[g.idx(ix, -ir, m)])));
    }
    mode_sign = -mode_sign;
  }
}

// ------------------------------------------------------------------------
// update_b_field interior, fields.f90:203-241: ix in 0..nx, ir in 1..ny-1
// (r_max local).  COPY = 1 also copies every element of B, ghosts included,
// to B_old before updating it (the half step's bxm_old = bxm,
// fields.f90:326-328, in the same pass): the grid then spans the whole
// array.
// ------------------------------------------------------------------------
template <class T, int COPY>
__global__ __launch_bounds__(256) void k_update_b(Grid g, cplx<T>* __restrict__ bx,
                                                  cplx<T>* __restrict__ br, cplx<T>* __restrict__ bt,
                                                  const cplx<T>* __restrict__ ex,
                                                  const cplx<T>* __restrict__ er,
                                                  const cplx<T>* __restrict__ et, cplx<T>* __restrict__ bx_old,
                                                  cplx<T>* __restrict__ br_old, cplx<T>* __restrict__ bt_old) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x + (COPY ? 1 - g.ng : 0);
  const int ir = blockIdx.y + (COPY ? 1 - g.ng : 1);
  const int m = blockIdx.z;
  if (ix > (COPY ? g.nx + g.ng : g.nx)) return;
  const long k = g.idx(ix, ir, m);
  const cplx<T> bx0 = bx[k], br0 = br[k], bt0 = bt[k];
  if (COPY) {
    bx_old[k] = bx0;
    br_old[k] = br0;
    bt_old[k] = bt0;
    if (ix < 0 || ix > g.nx || ir < 1 || ir > g.ny - 1) return;
  }
  const double r_dd = fabs((double)(ir - 1) * g.dy + g.y_grid_min);
  const double r_pd = r_dd + 0.5 * g.dy;
  const T r_d = (T)r_dd;
  const T a_x = (T)((double)m / r_dd);
  const T a_r = (T)((double)m / r_pd);
  const T dx = (T)g.dx, dy = (T)g.dy, dt = (T)g.dt, half = (T)0.5;
  const long krm = k - g.nx2;
  const long kxm = k - 1;
  const cplx<T> et0 = et[k], et_rm = et[krm], et_xm = et[kxm];
  const cplx<T> er0 = er[k], er_xm = er[kxm];
  const cplx<T> ex0 = ex[k], ex_rm = ex[krm];
  {  // (8) bxm
    cplx<T> t1 = imul(a_x, er0);
    cplx<T> t2 = cdivr(rmul(half, cadd(et0, et_rm)), r_d);
    cplx<T> t3 = cdivr(csub(et0, et_rm), dy);
    cplx<T> s = cadd(cadd(t1, t2), t3);
    bx[k] = csub(bx0, cmulr(cmulr(s, half), dt));
  }
  {  // (6) brm
    cplx<T> t1 = imul(a_r, ex0);
    cplx<T> t2 = cdivr(csub(et0, et_xm), dx);
    cplx<T> s = cadd(t1, t2);
    br[k] = cadd(br0, cmulr(cmulr(s, half), dt));
  }
  {  // (7) btm
    cplx<T> t1 = cneg(cdivr(csub(er0, er_xm), dx));
    cplx<T> t2 = cdivr(csub(ex0, ex_rm), dy);
    cplx<T> s = cadd(t1, t2);
    bt[k] = cadd(bt0, cmulr(cmulr(s, half), dt));
  }
}

// r=0 rows and reflected ghosts of update_b_field, fields.f90:249-310.
template <class T>
__global__ __launch_bounds__(256) void k_b_axis(Grid g, cplx<T>* __restrict__ bx,
                                                cplx<T>* __restrict__ br, cplx<T>* __restrict__ bt,
                                                const cplx<T>* __restrict__ ex,
                                                const cplx<T>* __restrict__ et) {
  const int ix = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1 - g.ng;
  if (ix > g.nx + g.ng) return;
  const int ng = g.ng;
  const double dt = g.dt, dx = g.dx;
  const double idy = 1.0 / g.dy;
  {
    const int m = 0;
    long k0 = g.idx(ix, 0, m), k1 = g.idx(ix, 1, m);
    br[k0] = czero<T>();
    bx[k0] = bx[k1];
    bt[k0] = cneg(bt[k1]);
    for (int ir = 1 - ng; ir <= -1; ++ir) {
      bt[g.idx(ix, ir, m)] = cneg(bt[g.idx(ix, -ir + 1, m)]);
      br[g.idx(ix, ir, m)] = cneg(br[g.idx(ix, -ir, m)]);
      bx[g.idx(ix, ir, m)] = bx[g.idx(ix, -ir + 1, m)];
    }
  }
  if (g.M > 1) {
    const int m = 1;
    long k0 = g.idx(ix, 0, m), k1 = g.idx(ix, 1, m);
    bx[k0] = cneg(bx[k1]);
    cplx<double> br0 = conv<double>(br[k0]);
    if (ix >= 2 - ng) {  // (22)
      cplx<double> t1 = imul(idy, conv<double>(ex[k0]));
      cplx<double> t2 = cdivr(csub(conv<double>(et[k0]), conv<double>(et[g.idx(ix - 1, 0, m)])), dx);
      cplx<double> s = cadd(t1, t2);
      br0 = cadd(br0, cmulr(cmulr(s, 0.5), dt));
      br[k0] = conv<T>(br0);
      br0 = conv<double>(br[k0]);
    }
    bt[k0] = conv<T>(csub(imul(-2.0, br0), conv<double>(bt[k1])));
    for (int ir = 1 - ng; ir <= -1; ++ir) {
      bt[g.idx(ix, ir, m)] = bt[g.idx(ix, -ir + 1, m)];
      br[g.idx(ix, ir, m)] = br[g.idx(ix, -ir, m)];
      bx[g.idx(ix, ir, m)] = cneg(bx[g.idx(ix, -ir + 1, m)]);
    }
  }
  double mode_sign = 1.0;
  for (int m = 2; m < g.M; ++m) {
    long k0 = g.idx(ix, 0, m), k1 = g.idx(ix, 1, m);
    bx[k0] = cneg(bx[k1]);
    br[k0] = czero<T>();
    bt[k0] = cneg(bt[k1]);
    for (int ir = 1 - ng; ir <= -1; ++ir) {
      bt[g.idx(ix, ir, m)] = conv<T>(cneg(rmul(mode_sign, conv<double>(bt[g.idx(ix, -ir + 1, m)]))));
      br[g.idx(ix, ir, m)] = conv<T>(cneg(rmul(mode_sign, conv<double>(br[g.idx(ix, -ir, m)]))));
      bx[g.idx(ix, ir, m)] = conv<T>(rmul(mode_sign, conv<double>(bx[g.idx(ix, -ir + 1, m)])));
    }
    mode_sign = -mode_sign;
  }
}

// ------------------------------------------------------------------------
// x halo exchange of one rank with itself when x is periodic
// (do_field_mpi_with_lengths_cmplx, boundary.F90:500-593), rows
// row_lo..row_hi.  One thread per (row, mode), ng columns each side.
// ------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void halo_x_body(const Grid& g, cplx<T>* __restrict__ f, int t, int row_lo, int row_hi,
                                            int do_max, int do_min) {
  const int nrow = row_hi - row_lo + 1;
  if (t >= nrow * g.M) return;
  const int j = row_lo + t % nrow, m = t / nrow;
  const int nx = g.nx, ng = g.ng;
  if (do_max)
    for (int i = 1; i <= ng; ++i) f[g.idx(nx + i, j, m)] = f[g.idx(i, j, m)];
  if (do_min)
    for (int i = 1; i <= ng; ++i) f[g.idx(i - ng, j, m)] = f[g.idx(nx - ng + i, j, m)];
}

template <class T>
__global__ __launch_bounds__(256) void k_halo_x(Grid g, cplx<T>* __restrict__ f, int row_lo,
                                                int row_hi, int do_max, int do_min) {
  halo_x_body(g, f, (int)(blockIdx.x * blockDim.x + threadIdx.x), row_lo, row_hi, do_max, do_min);
}

// field_mode_clamp_zero / field_mode_zero_gradient, boundary.F90:654-829,
// for x_min, x_max (one thread per (row, mode)) or y_max (one thread per
// (column, mode)).
template <class T>
__device__ __forceinline__ void edge_body(const Grid& g, cplx<T>* __restrict__ f, int t, int stag_x, int stag_y,
                                          int bnd, int clamp) {
  const int nx = g.nx, ny = g.ny, ng = g.ng;
  auto setv = [&](long a, long b) { f[a] = clamp ? cneg(f[b]) : f[b]; };
  if (bnd == XMIN || bnd == XMAX) {
    const int nrow = g.ny2;
    if (t >= nrow * g.M) return;
    const int j = 1 - ng + t % nrow, m = t / nrow;
    if (bnd == XMIN) {
      if (stag_x) {
        for (int i = 1; i <= ng - 1; ++i) setv(g.idx(i - ng, j, m), g.idx(ng - i, j, m));
        if (clamp) f[g.idx(0, j, m)] = czero<T>();
      } else {
        for (int i = 1; i <= ng; ++i) setv(g.idx(i - ng, j, m), g.idx(ng + 1 - i, j, m));
      }
    } else {
      if (stag_x) {
        if (clamp) f[g.idx(nx, j, m)] = czero<T>();
        for (int i = 1; i <= ng - 1; ++i) setv(g.idx(nx + i, j, m), g.idx(nx - i, j, m));
      } else {
        for (int i = 1; i <= ng; ++i) setv(g.idx(nx + i, j, m), g.idx(nx + 1 - i, j, m));
      }
    }
  } else {  // YMAX
    const int ncol = g.nx2;
    if (t >= ncol * g.M) return;
    const int ix = 1 - ng + t % ncol, m = t / ncol;
    if (stag_y) {
      if (clamp) f[g.idx(ix, ny, m)] = czero<T>();
      for (int i = 1; i <= ng - 1; ++i) setv(g.idx(ix, ny + i, m), g.idx(ix, ny - i, m));
    } else {
      for (int i = 1; i <= ng; ++i) setv(g.idx(ix, ny + i, m), g.idx(ix, ny + 1 - i, m));
    }
  }
}

// The halo copies and edge rules of one boundary call, as a list of
// operations on whole fields: one launch runs a group of operations that
// commute (blockIdx.y picks the operation), and a group ends where the next
// operation does not commute with one in it.  Operations on different
// fields commute; on one field only the x_min and x_max edges do (they touch
// disjoint columns, nx >= 2 ng); an x operation and a y_max edge of one field
// meet at the corners, so they keep their order.
enum { BC_OP_HALO = 0, BC_OP_EDGE = 1 };
template <class T>
struct BcOp {
  cplx<T>* f;
  int kind, n;              // n: threads
  int a, b, c, d;           // halo: row_lo, row_hi, do_max, do_min; edge: stag_x, stag_y, bnd, clamp
};
constexpr int MAX_BC_OPS = 12;
template <class T>
struct BcOpList {
  BcOp<T> op[MAX_BC_OPS];
};

template <class T>
__global__ __launch_bounds__(256) void k_bc_ops(Grid g, BcOpList<T> L) {
  const BcOp<T>& o = L.op[blockIdx.y];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= o.n) return;
  if (o.kind == BC_OP_HALO) halo_x_body(g, o.f, t, o.a, o.b, o.c, o.d);
  else edge_body(g, o.f, t, o.a, o.b, o.c, o.d);
}

// ------------------------------------------------------------------------
// setup_field_boundaries, setup.F90:393-420.  One thread per (row, mode).
// out[0..5] = ex, er, et, bx, br, bt snapshots at x_min; out[6..11] x_max.
// ------------------------------------------------------------------------
template <class T>
struct BndPtrs {
  cplx<T>* p[12];
};
template <class T>
__global__ void k_setup_bnd(Grid g, const cplx<T>* ex, const cplx<T>* er, const cplx<T>* et,
                            const cplx<T>* bx, const cplx<T>* br, const cplx<T>* bt, BndPtrs<T> o) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.ny2 * g.M) return;
  const int j = 1 - g.ng + t % g.ny2, m = t / g.ny2, nx = g.nx;
  const long s = g.idx2(j, m);
  const T h = (T)0.5;
  o.p[0][s] = rmul(h, cadd(ex[g.idx(1, j, m)], ex[g.idx(0, j, m)]));
  o.p[1][s] = er[g.idx(0, j, m)];
  o.p[2][s] = et[g.idx(0, j, m)];
  o.p[6][s] = rmul(h, cadd(ex[g.idx(nx, j, m)], ex[g.idx(nx + 1, j, m)]));
  o.p[7][s] = er[g.idx(nx, j, m)];
  o.p[8][s] = et[g.idx(nx, j, m)];
  o.p[3][s] = bx[g.idx(0, j, m)];
  o.p[4][s] = rmul(h, cadd(br[g.idx(1, j, m)], br[g.idx(0, j, m)]));
  o.p[5][s] = rmul(h, cadd(bt[g.idx(1, j, m)], bt[g.idx(0, j, m)]));
  o.p[9][s] = bx[g.idx(nx, j, m)];
  o.p[10][s] = rmul(h, cadd(br[g.idx(nx, j, m)], br[g.idx(nx + 1, j, m)]));
  o.p[11][s] = rmul(h, cadd(bt[g.idx(nx, j, m)], bt[g.idx(nx + 1, j, m)]));
}

// ------------------------------------------------------------------------
// outflow_bcs_x_min / x_max, laser.f90:411-633 (arithmetic in FP64).
// side = 0: x_min, 1: x_max.  src = source1(0:ny), source2(0:ny) or null.
// One thread per (row, mode); the bx column copy runs first in its own
// kernel because the Brm update reads bxm(1, ir+1).
// ------------------------------------------------------------------------
template <class T>
__global__ void k_outflow_x_bx(Grid g, int side, cplx<T>* bx, const cplx<T>* bnd_bx) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (g.ny + 1) * g.M) return;
  const int ir = t % (g.ny + 1), m = t / (g.ny + 1);
  bx[g.idx(side ? g.nx : 0, ir, m)] = bnd_bx[g.idx2(ir, m)];
}

template <class T>
__global__ void k_outflow_x(Grid g, int side, cplx<T>* bt, cplx<T>* br, const cplx<T>* bx,
                            const cplx<T>* er, const cplx<T>* et, const cplx<T>* jr,
                            const cplx<T>* jt, const cplx<T>* bnd_er, const cplx<T>* bnd_et,
                            const cplx<T>* bnd_br, const cplx<T>* bnd_bt,
                            const double* __restrict__ src) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int ny = g.ny;
  if (t >= ny * g.M) return;
  const int ir = 1 + t % ny, m = t / ny;
  const double c = C_LIGHT, dt = g.dt;
  const double dtc2 = dt * (c * c), lx = dtc2 / g.dx, lr = dtc2 / g.dy;
  const double sum = 1.0 / (lx + c), diff = lx - c, dt_eps = dt / EPSILON0;
  const int xb = side ? g.nx : 1;       // boundary column written
  const int xi = side ? g.nx - 1 : 2;   // interior neighbour (btm, brm)
  const int xs = side ? g.nx - 1 : 1;   // column of erm/etm/bxm/jrm/jtm
  auto D = [](cplx<T> a) { return conv<double>(a); };
  // Btheta, rows 1..ny.  x_min pairs source_t(ir); x_max pairs source_t(ir-1)
  // and both pair r_d_vals(ir-1) (non-conformable arrays in the reference).
  {
    int is = side ? ir - 1 : ir;
    cplx<double> st = (m == 1 && src) ? mk(src[is], src[ny + 1 + is]) : mk(0.0, 0.0);
    double r_d = fabs((double)(ir - 1 - 1) * g.dy + g.y_grid_min);
    cplx<double> b = rmul(2.0, cadd(D(bnd_er[g.idx2(ir, m)]), rmul(c, D(bnd_bt[g.idx2(ir, m)]))));
    cplx<double> d = rmul(2.0, D(er[g.idx(xs, ir, m)]));
    cplx<double> e = cdivr(imul(((double)m * (c * c)) * dt, D(bx[g.idx(xs, ir, m)])), r_d);
    cplx<double> h = rmul(dt_eps, D(jr[g.idx(xs, ir, m)]));
    cplx<double> k = rmul(diff, D(bt[g.idx(xi, ir, m)]));
    cplx<double> v;
    if (!side) {
      cplx<double> a = rmul(4.0, st);
      v = rmul(sum, cadd(cadd(cadd(csub(cadd(a, b), d), e), h), k));
    } else {
      cplx<double> a = rmul(-4.0, st);
      v = rmul(sum, cadd(csub(csub(cadd(csub(a, b), d), e), h), k));
    }
    bt[g.idx(xb, ir, m)] = conv<T>(v);
  }
  // Br, rows 1..ny-1 (r=0 and r_max are local)
  if (ir <= ny - 1) {
    cplx<double> sr = (m == 1 && src) ? mk(src[ny + 1 + ir], -src[ir]) : mk(0.0, 0.0);
    cplx<double> b = rmul(2.0, cadd(D(bnd_et[g.idx2(ir, m)]), rmul(c, D(bnd_br[g.idx2(ir, m)]))));
    cplx<double> d = rmul(2.0, D(et[g.idx(xs, ir, m)]));
    cplx<double> e = rmul(lr, csub(D(bx[g.idx(xs, ir + 1, m)]), D(bx[g.idx(xs, ir, m)])));
    cplx<double> h = rmul(dt_eps, D(jt[g.idx(xs, ir, m)]));
    cplx<double> k = rmul(diff, D(br[g.idx(xi, ir, m)]));
    cplx<double> v;
    if (!side) {
      cplx<double> a = rmul(-4.0, sr);
      v = rmul(sum, cadd(csub(csub(cadd(csub(a, b), d), e), h), k));
    } else {
      cplx<double> a = rmul(4.0, sr);
      v = rmul(sum, cadd(cadd(cadd(csub(cadd(a, b), d), e), h), k));
    }
    br[g.idx(xb, ir, m)] = conv<T>(v);
  }
}

// outflow_bcs_r_max, laser.f90:637-690.  One thread per (column, mode).
template <class T>
__global__ void k_outflow_r(Grid g, cplx<T>* bx, cplx<T>* bt, const cplx<T>* bx_old,
                            const cplx<T>* bt_old, const cplx<T>* br, const cplx<T>* br_old,
                            const cplx<T>* er, const cplx<T>* et, const cplx<T>* jx,
                            const cplx<T>* jx_old, const cplx<T>* jt, const cplx<T>* jt_old) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = g.nx, ny = g.ny;
  if (t >= nx * g.M) return;
  const int ix = 1 + t % nx, m = t / nx;
  const double c = C_LIGHT, dt = g.dt;
  const double dtc2 = dt * (c * c);
  const double inv_r = 1.0 / (((double)ny - 1.5) * g.dy + g.y_grid_min);
  const double dtc2_4r = 0.25 * dtc2 * inv_r;
  const double icdt = ((0.5 * c) * dt) * inv_r;
  const double lx = dtc2 / g.dx, ly = dtc2 / g.dy;
  const double sum_x = 1.0 / (ly + c), sum_t = 1.0 / ((ly + c) + dtc2_4r);
  const double dt_2eps = (0.5 * dt) / EPSILON0;
  auto D = [](cplx<T> a) { return conv<double>(a); };
  if (ix <= nx - 1) {
    cplx<double> a = cneg(cmulr(D(bx[g.idx(ix, ny - 1, m)]), c - ly));
    cplx<double> b = cmulr(D(bx_old[g.idx(ix, ny, m)]), -c + ly);
    cplx<double> d = cmulr(D(bx_old[g.idx(ix, ny - 1, m)]), -c - ly);
    cplx<double> e = rmul((c * dt) * inv_r, D(et[g.idx(ix, ny - 1, m)]));
    cplx<double> h = rmul(0.5 * lx, csub(cadd(csub(D(br[g.idx(ix + 1, ny - 1, m)]), D(br[g.idx(ix, ny - 1, m)])),
                                              D(br_old[g.idx(ix + 1, ny - 1, m)])),
                                         D(br_old[g.idx(ix, ny - 1, m)])));
    cplx<double> k = imul(icdt * (double)m, cadd(D(er[g.idx(ix, ny, m)]), D(er[g.idx(ix, ny - 1, m)])));
    cplx<double> l = rmul(dt_2eps, cadd(D(jt[g.idx(ix, ny - 1, m)]), D(jt_old[g.idx(ix, ny - 1, m)])));
    bx[g.idx(ix, ny, m)] = conv<T>(rmul(sum_x, csub(csub(cadd(csub(csub(csub(a, b), d), e), h), k), l)));
  }
  {
    cplx<double> a = cneg(cmulr(D(bt[g.idx(ix, ny - 1, m)]), (c - ly) + dtc2_4r));
    cplx<double> b = cmulr(D(bt_old[g.idx(ix, ny, m)]), (-c + ly) + dtc2_4r);
    cplx<double> d = cmulr(D(bt_old[g.idx(ix, ny - 1, m)]), (-c - ly) + dtc2_4r);
    cplx<double> e = rmul((0.5 * lx) / c, csub(csub(cadd(D(er[g.idx(ix, ny, m)]), D(er[g.idx(ix, ny - 1, m)])),
                                                    D(er[g.idx(ix - 1, ny, m)])),
                                               D(er[g.idx(ix - 1, ny - 1, m)])));
    cplx<double> k = imul((icdt * (double)m) * c, cadd(D(br[g.idx(ix, ny - 1, m)]), D(br_old[g.idx(ix, ny - 1, m)])));
    cplx<double> l = rmul(dt_2eps, cadd(D(jx[g.idx(ix, ny - 1, m)]), D(jx_old[g.idx(ix, ny - 1, m)])));
    bt[g.idx(ix, ny, m)] = conv<T>(rmul(sum_t, cadd(csub(csub(csub(csub(a, b), d), e), k), l)));
  }
}

// zero_b at r_max, boundary.F90:1528-1532
template <class T>
__global__ void k_zero_b_rmax(Grid g, cplx<T>* bx, cplx<T>* br, cplx<T>* bt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.nx2 * g.M) return;
  const int ix = 1 - g.ng + t % g.nx2, m = t / g.nx2;
  const long k = g.idx(ix, g.ny, m);
  bx[k] = czero<T>();
  br[k] = czero<T>();
  bt[k] = czero<T>();
}

// ------------------------------------------------------------------------
// currents
// ------------------------------------------------------------------------

// current_bcs_r_min_final, boundary.F90:1909-1959.  One thread per column,
// all modes (the fold of mode m only touches mode m).
template <class T>
__global__ __launch_bounds__(256) void k_current_axis(Grid g, cplx<T>* __restrict__ jx,
                                                      cplx<T>* __restrict__ jr,
                                                      cplx<T>* __restrict__ jt) {
  const int ix = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1 - g.ng;
  if (ix > g.nx + g.ng) return;
  const int ng = g.ng;
  T mode_sign = (T)1;
  for (int m = 0; m < g.M; ++m) {
    for (int j = 2; j <= ng; ++j) {
      long a = g.idx(ix, 1 - j, m);
      long kx = g.idx(ix, j - 1, m), kr = g.idx(ix, j, m);
      jx[kx] = cadd(jx[kx], rmul(mode_sign, jx[a]));
      jr[kr] = csub(jr[kr], rmul(mode_sign, jr[a]));
      jt[kx] = csub(jt[kx], rmul(mode_sign, jt[a]));
      jx[a] = czero<T>();
      jr[a] = czero<T>();
      jt[a] = czero<T>();
    }
    long k1 = g.idx(ix, 1, m);
    jr[k1] = csub(jr[k1], rmul(mode_sign, jr[g.idx(ix, 0, m)]));
    mode_sign = -mode_sign;
  }
  for (int m = 0; m < g.M; ++m) {
    long k0 = g.idx(ix, 0, m), k1 = g.idx(ix, 1, m), k2 = g.idx(ix, 2, m);
    if (m > 0) {
      jx[k0] = czero<T>();
    } else {
      jx[k0] = conv<T>(cdivr(csub(rmul(4.0, conv<double>(jx[k1])), conv<double>(jx[k2])), 3.0));
    }
    if (m == 1) {
      cplx<double> w = csub(rmul(9.0, conv<double>(jr[k1])), conv<double>(jr[k2]));
      cplx<double> jt0 = mk(w.im / 8.0, -(w.re / 8.0));
      jt[k0] = conv<T>(jt0);
      jt0 = conv<double>(jt[k0]);
      jr[k0] = conv<T>(csub(imul(2.0, jt0), conv<double>(jr[k1])));
    } else {
      jt[k0] = czero<T>();
      jr[k0] = cneg(jr[k1]);
    }
  }
}

// The three current components of one kernel's launch: blockIdx.y = 0, 1, 2
// picks jx, jr, jt (flip = blockIdx.y + 1); the components are independent.
template <class T>
struct Jxrt {
  cplx<T>* j[3];
};

// particle_reflection_bcs_complex (boundary.F90:918-1015) + the one-rank
// particle_periodic_bcs_complex (:1133-1245), each component (flip = 1, 2,
// 3 for jx, jr, jt), all modes.  x part: one thread per (row, mode).
template <class T>
__global__ void k_current_x(Grid g, Jxrt<T> jj) {
  cplx<T>* __restrict__ J = jj.j[blockIdx.y];
  const int flip = blockIdx.y + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nrow = g.ny2;
  if (t >= nrow * g.M) return;
  const int j = 1 - g.ng + t % nrow, m = t / nrow;
  const int nx = g.nx, ng = g.ng;
  if (g.bc_particle[XMIN] == BC_REFLECT) {
    if (flip == 1) {
      for (int i = 1; i <= ng - 1; ++i) {
        J[g.idx(i, j, m)] = csub(J[g.idx(i, j, m)], J[g.idx(1 - i, j, m)]);
        J[g.idx(1 - i, j, m)] = czero<T>();
      }
    } else {
      for (int i = 1; i <= ng - 1; ++i) {
        J[g.idx(i, j, m)] = cadd(J[g.idx(i, j, m)], J[g.idx(-i, j, m)]);
        J[g.idx(-i, j, m)] = czero<T>();
      }
    }
  }
  if (g.bc_particle[XMAX] == BC_REFLECT) {
    if (flip == 1) {
      for (int i = 1; i <= ng; ++i) {
        J[g.idx(nx + 1 - i, j, m)] = csub(J[g.idx(nx + 1 - i, j, m)], J[g.idx(nx + i, j, m)]);
        J[g.idx(nx + i, j, m)] = czero<T>();
      }
    } else {
      for (int i = 1; i <= ng; ++i) {
        J[g.idx(nx - i, j, m)] = cadd(J[g.idx(nx - i, j, m)], J[g.idx(nx + i, j, m)]);
        J[g.idx(nx + i, j, m)] = czero<T>();
      }
    }
  }
}

// r_max reflection part of particle_reflection_bcs_complex: one thread per
// (column, mode); runs after the x part (the reference applies x_min, x_max,
// then y_max within one call).
template <class T>
__global__ void k_current_rmax(Grid g, Jxrt<T> jj) {
  cplx<T>* __restrict__ J = jj.j[blockIdx.y];
  const int flip = blockIdx.y + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.nx2 * g.M) return;
  const int ix = 1 - g.ng + t % g.nx2, m = t / g.nx2;
  const int ny = g.ny, ng = g.ng;
  const double r_max = g.y_grid_min + ((double)ny - 0.5) * g.dy;
  for (int i = 1; i <= ng; ++i) {
    if (flip == 2) {
      double num = r_max + ((double)i - 0.5) * g.dy, den = r_max - ((double)i - 0.5) * g.dy;
      long a = g.idx(ix, ny + 1 - i, m);
      J[a] = conv<T>(csub(conv<double>(J[a]), cdivr(cmulr(conv<double>(J[g.idx(ix, ny + i, m)]), num), den)));
    } else if (flip == 1) {
      double num = r_max + (double)i * g.dy, den = r_max - (double)i * g.dy;
      long a = g.idx(ix, ny - i, m);
      J[a] = conv<T>(cadd(conv<double>(J[a]), cdivr(cmulr(conv<double>(J[g.idx(ix, ny + i, m)]), num), den)));
    } else {
      long a = g.idx(ix, ny - i, m);
      J[a] = cadd(J[a], J[g.idx(ix, ny + i, m)]);
    }
    J[g.idx(ix, ny + i, m)] = czero<T>();
  }
}

// periodic ghost sum on one rank, one thread per (row, mode)
template <class T>
__global__ void k_current_periodic(Grid g, Jxrt<T> jj) {
  cplx<T>* __restrict__ J = jj.j[blockIdx.y];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nrow = g.ny2;
  if (t >= nrow * g.M) return;
  const int j = 1 - g.ng + t % nrow, m = t / nrow;
  const int nx = g.nx, ng = g.ng;
  for (int i = 1; i <= ng; ++i) J[g.idx(i, j, m)] = cadd(J[g.idx(i, j, m)], J[g.idx(nx + i, j, m)]);
  for (int i = 1; i <= ng; ++i)
    J[g.idx(nx - ng + i, j, m)] = cadd(J[g.idx(nx - ng + i, j, m)], J[g.idx(i - ng, j, m)]);
}

// ------------------------------------------------------------------------
// fill / convert
// ------------------------------------------------------------------------
template <class T>
__global__ void k_fill_zero(cplx<T>* __restrict__ p, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = czero<T>();
}

__global__ void k_d2f(const cplx<double>* __restrict__ s, cplx<float>* __restrict__ d, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) d[i] = mk((float)s[i].re, (float)s[i].im);
}
__global__ void k_f2d(const cplx<float>* __restrict__ s, cplx<double>* __restrict__ d, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) d[i] = mk((double)s[i].re, (double)s[i].im);
}

// ------------------------------------------------------------------------
// Per-mode field energy (FP64 accumulation), one block per (row, mode),
// reduction over the row then one atomic per block.
// out[2*m] = sum over E terms, out[2*m+1] = sum over B terms (unscaled).
// ------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_energy(Grid g, const cplx<T>* ex, const cplx<T>* er,
                                                const cplx<T>* et, const cplx<T>* bx,
                                                const cplx<T>* br, const cplx<T>* bt,
                                                double* out) {
  const int ir = blockIdx.x + 1, m = blockIdx.y;
  const double r_c = ((double)ir - 0.5) * g.dy, r_e = (double)ir * g.dy;
  double se = 0.0, sb = 0.0;
  for (int ix = 1 + threadIdx.x; ix <= g.nx; ix += blockDim.x) {
    long k = g.idx(ix, ir, m);
    auto n2 = [](cplx<T> a) { return (double)a.re * (double)a.re + (double)a.im * (double)a.im; };
    se += n2(ex[k]) * r_e + n2(er[k]) * r_c + n2(et[k]) * r_e;
    sb += n2(bx[k]) * r_c + n2(br[k]) * r_e + n2(bt[k]) * r_c;
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = se;
  red[1][threadIdx.x] = sb;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&out[2 * m], red[0][0]);
    atomicAdd(&out[2 * m + 1], red[1][0]);
  }
}

// ------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------
static inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

template <class T>
void launch_update_e(const Grid& g, FieldSet<T>& F, hipStream_t s) {
  dim3 grid(nblk(g.nx + 1), g.ny, g.M);
  hipLaunchKernelGGL(k_update_e<T>, grid, dim3(256), 0, s, g, F.ex(), F.er(), F.et(), F.bx(), F.br(),
                     F.bt(), F.jx(), F.jr(), F.jt());
  hipLaunchKernelGGL(k_e_axis<T>, dim3(nblk(g.nx2)), dim3(256), 0, s, g, F.ex(), F.er(), F.et(), F.bt(),
                     F.jx());
}

template <class T>
void launch_update_b(const Grid& g, FieldSet<T>& F, int copy_old, hipStream_t s) {
  if (copy_old)
    hipLaunchKernelGGL((k_update_b<T, 1>), dim3(nblk(g.nx2), g.ny2, g.M), dim3(256), 0, s, g, F.bx(), F.br(),
                       F.bt(), F.ex(), F.er(), F.et(), F.bx_old(), F.br_old(), F.bt_old());
  else if (g.ny > 1)
    hipLaunchKernelGGL((k_update_b<T, 0>), dim3(nblk(g.nx + 1), g.ny - 1, g.M), dim3(256), 0, s, g, F.bx(),
                       F.br(), F.bt(), F.ex(), F.er(), F.et(), nullptr, nullptr, nullptr);
  hipLaunchKernelGGL(k_b_axis<T>, dim3(nblk(g.nx2)), dim3(256), 0, s, g, F.bx(), F.br(), F.bt(), F.ex(),
                     F.et());
}

template <class T>
class BcQueue {
 public:
  BcQueue(const Grid& g, hipStream_t s) : g_(g), s_(s) {}
  ~BcQueue() { flush(); }
  // halo rows row_lo..row_hi of f when x is periodic
  void halo(cplx<T>* f, int row_lo, int row_hi) {
    const int do_max = g_.bc_field[XMAX] == BC_PERIODIC, do_min = g_.bc_field[XMIN] == BC_PERIODIC;
    if (!do_max && !do_min) return;
    add(BcOp<T>{f, BC_OP_HALO, (row_hi - row_lo + 1) * g_.M, row_lo, row_hi, do_max, do_min});
  }
  void edge(cplx<T>* f, int sx, int sy, int bnd, int clamp) {
    if (g_.bc_field[bnd] == BC_PERIODIC) return;
    const int n = (bnd == YMAX ? g_.nx2 : g_.ny2) * g_.M;
    add(BcOp<T>{f, BC_OP_EDGE, n, sx, sy, bnd, clamp});
  }
  void flush() {
    if (!n_) return;
    BcOpList<T> L{};
    int most = 0;
    for (int i = 0; i < n_; ++i) {
      L.op[i] = ops_[i];
      most = std::max(most, ops_[i].n);
    }
    hipLaunchKernelGGL(k_bc_ops<T>, dim3(nblk(most), n_), dim3(256), 0, s_, g_, L);
    n_ = 0;
  }

 private:
  static bool commute(const BcOp<T>& x, const BcOp<T>& y) {
    if (x.f != y.f) return true;
    return x.kind == BC_OP_EDGE && y.kind == BC_OP_EDGE && x.c != YMAX && y.c != YMAX && x.c != y.c;
  }
  void add(const BcOp<T>& o) {
    for (int i = 0; i < n_; ++i)
      if (!commute(ops_[i], o)) {
        flush();
        break;
      }
    if (n_ == MAX_BC_OPS) flush();
    ops_[n_++] = o;
  }
  const Grid& g_;
  hipStream_t s_;
  BcOp<T> ops_[MAX_BC_OPS];
  int n_ = 0;
};

// stagger flags (x, y) of each component, constants.F90:291-299
enum { S_EX, S_ER, S_ET, S_BX, S_BR, S_BT };
static const int STX[6] = {0, 1, 1, 1, 0, 0};
static const int STY[6] = {1, 0, 1, 0, 1, 0};

template <class T>
static void generic_edges(BcQueue<T>& q, const Grid& g, cplx<T>* a, cplx<T>* b, cplx<T>* c, int sa, int sb,
                          int sc) {
  for (int i = XMIN; i <= YMAX; ++i) {
    if (i == YMIN) continue;
    int bc = g.bc_field[i];
    int clamp = (bc == BC_CLAMP || bc == BC_SIMPLE_LASER || bc == BC_SIMPLE_OUTFLOW);
    int zg = (bc == BC_ZERO_GRADIENT || bc == BC_CPML_LASER || bc == BC_CPML_OUTFLOW);
    if (clamp || zg) {
      q.edge(a, STX[sa], STY[sa], i, clamp);
      q.edge(b, STX[sb], STY[sb], i, clamp);
      q.edge(c, STX[sc], STY[sc], i, clamp);
    }
  }
}

// efield_bcs, boundary.F90:1355-1413
template <class T>
void launch_efield_bcs(const Grid& g, FieldSet<T>& F, hipStream_t s) {
  const int ny = g.ny, ng = g.ng;
  BcQueue<T> q(g, s);
  q.halo(F.ex(), 1 - ng, ny + ng - 1);
  q.halo(F.er(), 1 - ng, ny + ng);
  q.halo(F.et(), 1 - ng, ny + ng - 1);
  for (int i = XMIN; i <= XMAX; ++i)
    if (g.bc_field[i] == BC_CONDUCT) {
      q.edge(F.ex(), STX[S_EX], STY[S_EX], i, 1);
      q.edge(F.er(), STX[S_ER], STY[S_ER], i, 0);
      q.edge(F.et(), STX[S_ET], STY[S_ET], i, 0);
    }
  if (g.bc_field[YMAX] == BC_CONDUCT) {
    q.edge(F.ex(), STX[S_EX], STY[S_EX], YMAX, 0);
    q.edge(F.er(), STX[S_ER], STY[S_ER], YMAX, 1);
    q.edge(F.et(), STX[S_ET], STY[S_ET], YMAX, 0);
  }
  generic_edges(q, g, F.ex(), F.er(), F.et(), S_EX, S_ER, S_ET);
}

// bfield_bcs, boundary.F90:1417-1476
template <class T>
void launch_bfield_bcs(const Grid& g, FieldSet<T>& F, int mpi_only, hipStream_t s) {
  const int ny = g.ny, ng = g.ng;
  BcQueue<T> q(g, s);
  q.halo(F.bx(), 1 - ng, ny + ng);
  q.halo(F.br(), 1 - ng, ny + ng - 1);
  q.halo(F.bt(), 1 - ng, ny + ng);
  if (mpi_only) return;
  for (int i = XMIN; i <= XMAX; ++i)
    if (g.bc_field[i] == BC_CONDUCT) {
      q.edge(F.bx(), STX[S_BX], STY[S_BX], i, 0);
      q.edge(F.br(), STX[S_BR], STY[S_BR], i, 1);
      q.edge(F.bt(), STX[S_BT], STY[S_BT], i, 1);
    }
  if (g.bc_field[YMAX] == BC_CONDUCT) {
    q.edge(F.bx(), STX[S_BX], STY[S_BX], YMAX, 1);
    q.edge(F.br(), STX[S_BR], STY[S_BR], YMAX, 0);
    q.edge(F.bt(), STX[S_BT], STY[S_BT], YMAX, 1);
  }
  generic_edges(q, g, F.bx(), F.br(), F.bt(), S_BX, S_BR, S_BT);
}

template <class T>
void launch_setup_bnd(const Grid& g, FieldSet<T>& F, hipStream_t s) {
  BndPtrs<T> o;
  for (int k = 0; k < 6; ++k) {
    o.p[k] = F.bnd_xmin[k];
    o.p[6 + k] = F.bnd_xmax[k];
  }
  hipLaunchKernelGGL(k_setup_bnd<T>, dim3(nblk((long)g.ny2 * g.M)), dim3(256), 0, s, g, F.ex(), F.er(),
                     F.et(), F.bx(), F.br(), F.bt(), o);
}

// bfield_final_bcs, boundary.F90:1505-1537; src_* are device pointers or null
template <class T>
void launch_bfield_final_bcs(const Grid& g, FieldSet<T>& F, const double* src_xmin,
                             const double* src_xmax, hipStream_t s) {
  launch_bfield_bcs(g, F, 0, s);
  for (int side = 0; side < 2; ++side) {
    int b = g.bc_field[side];
    if (!(b == BC_SIMPLE_LASER || b == BC_SIMPLE_OUTFLOW)) continue;
    const double* src = (b == BC_SIMPLE_LASER) ? (side ? src_xmax : src_xmin) : nullptr;
    cplx<T>** bnd = side ? F.bnd_xmax : F.bnd_xmin;
    hipLaunchKernelGGL(k_outflow_x_bx<T>, dim3(nblk((long)(g.ny + 1) * g.M)), dim3(256), 0, s, g, side,
                       F.bx(), (const cplx<T>*)bnd[3]);
    hipLaunchKernelGGL(k_outflow_x<T>, dim3(nblk((long)g.ny * g.M)), dim3(256), 0, s, g, side, F.bt(),
                       F.br(), F.bx(), F.er(), F.et(), F.jr(), F.jt(), bnd[1], bnd[2], bnd[4], bnd[5], src);
  }
  int b = g.bc_field[YMAX];
  if (b == BC_SIMPLE_OUTFLOW) {
    hipLaunchKernelGGL(k_outflow_r<T>, dim3(nblk((long)g.nx * g.M)), dim3(256), 0, s, g, F.bx(), F.bt(),
                       F.bx_old(), F.bt_old(), F.br(), F.br_old(), F.er(), F.et(), F.jx(), F.jx_old(),
                       F.jt(), F.jt_old());
  } else if (b == BC_ZERO_B) {
    hipLaunchKernelGGL(k_zero_b_rmax<T>, dim3(nblk((long)g.nx2 * g.M)), dim3(256), 0, s, g, F.bx(), F.br(),
                       F.bt());
  }
  launch_bfield_bcs(g, F, 1, s);
}

template <class T>
void launch_current_axis(const Grid& g, FieldSet<T>& F, hipStream_t s) {
  hipLaunchKernelGGL(k_current_axis<T>, dim3(nblk(g.nx2)), dim3(256), 0, s, g, F.jx(), F.jr(), F.jt());
}

// current_finish, current_smooth.F90:29-46: current_bcs then J halos
template <class T>
void launch_current_finish(const Grid& g, FieldSet<T>& F, hipStream_t s) {
  const Jxrt<T> J{{F.jx(), F.jr(), F.jt()}};
  const bool xref = g.bc_particle[XMIN] == BC_REFLECT || g.bc_particle[XMAX] == BC_REFLECT;
  const bool rref = g.bc_particle[YMAX] == BC_REFLECT;
  const bool per = g.bc_particle[XMIN] == BC_PERIODIC && g.bc_particle[XMAX] == BC_PERIODIC;
  // per component: x reflection, r_max reflection, periodic sum, halos (the
  // reference's order); the components run side by side
  if (xref) hipLaunchKernelGGL(k_current_x<T>, dim3(nblk((long)g.ny2 * g.M), 3), dim3(256), 0, s, g, J);
  if (rref) hipLaunchKernelGGL(k_current_rmax<T>, dim3(nblk((long)g.nx2 * g.M), 3), dim3(256), 0, s, g, J);
  if (per) hipLaunchKernelGGL(k_current_periodic<T>, dim3(nblk((long)g.ny2 * g.M), 3), dim3(256), 0, s, g, J);
  BcQueue<T> q(g, s);
  for (int c = 0; c < 3; ++c) q.halo(J.j[c], 1 - g.ng, g.ny + g.ng);
}

// the three current arrays of a push, one launch
template <class T>
__global__ void k_zero_j(Jxrt<T> jj, long n) {
  cplx<T>* __restrict__ p = jj.j[blockIdx.y];
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = czero<T>();
}

template <class T>
void launch_zero_j(FieldSet<T>& F, long n, hipStream_t s) {
  unsigned b = nblk(n);
  if (b > 16384) b = 16384;
  hipLaunchKernelGGL(k_zero_j<T>, dim3(b, 3), dim3(256), 0, s, Jxrt<T>{{F.jx(), F.jr(), F.jt()}}, n);
}

template <class T>
void launch_fill_zero(cplx<T>* p, long n, hipStream_t s) {
  unsigned b = nblk(n);
  if (b > 65536) b = 65536;
  hipLaunchKernelGGL(k_fill_zero<T>, dim3(b), dim3(256), 0, s, p, n);
}

void launch_d2f(const cplx<double>* src, cplx<float>* dst, long n, hipStream_t s) {
  unsigned b = nblk(n);
  if (b > 65536) b = 65536;
  hipLaunchKernelGGL(k_d2f, dim3(b), dim3(256), 0, s, src, dst, n);
}
void launch_f2d(const cplx<float>* src, cplx<double>* dst, long n, hipStream_t s) {
  unsigned b = nblk(n);
  if (b > 65536) b = 65536;
  hipLaunchKernelGGL(k_f2d, dim3(b), dim3(256), 0, s, src, dst, n);
}

// ------------------------------------------------------------------------
// moving window: shift_field_modes (window.F90:132-150), out(i) = in(i + 1)
// for i in [1-ng, nx+ng-1], out(nx+ng) = in(nx+ng); and the x_max
// incoming-cell fix of shift_fields (window.F90:101-128)
template <class T>
__global__ void k_shift_x(Grid g, const cplx<T>* __restrict__ in, cplx<T>* __restrict__ out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.plane * g.M) return;
  const int i = (int)(e % g.nx2);  // 0-based column: ix = i + 1 - ng
  out[e] = i + 1 < g.nx2 ? in[e + 1] : in[e];
}

template <class T>
__global__ void k_window_fix(Grid g, cplx<T>* ex, cplx<T>* er, cplx<T>* et, cplx<T>* bx, cplx<T>* br, cplx<T>* bt,
                             const cplx<T>* bex, const cplx<T>* ber, const cplx<T>* bet, const cplx<T>* bbx,
                             const cplx<T>* bbr, const cplx<T>* bbt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.ny2 * g.M) return;
  const int m = t / g.ny2, j = t % g.ny2 + 1 - g.ng;
  const int nx = g.nx;
  const long s = g.idx2(j, m);
  const T h = (T)0.5;
  ex[g.idx(nx + 1, j, m)] = bex[s];
  er[g.idx(nx, j, m)] = ber[s];
  et[g.idx(nx, j, m)] = bet[s];
  ex[g.idx(nx, j, m)] = rmul(h, cadd(ex[g.idx(nx - 1, j, m)], ex[g.idx(nx + 1, j, m)]));
  er[g.idx(nx - 1, j, m)] = rmul(h, cadd(er[g.idx(nx - 2, j, m)], er[g.idx(nx, j, m)]));
  et[g.idx(nx - 1, j, m)] = rmul(h, cadd(et[g.idx(nx - 2, j, m)], et[g.idx(nx, j, m)]));
  bx[g.idx(nx, j, m)] = bbx[s];
  br[g.idx(nx + 1, j, m)] = bbr[s];
  bt[g.idx(nx + 1, j, m)] = bbt[s];
  bx[g.idx(nx - 1, j, m)] = rmul(h, cadd(bx[g.idx(nx - 2, j, m)], bx[g.idx(nx, j, m)]));
  br[g.idx(nx, j, m)] = rmul(h, cadd(br[g.idx(nx - 1, j, m)], br[g.idx(nx + 1, j, m)]));
  bt[g.idx(nx, j, m)] = rmul(h, cadd(bt[g.idx(nx - 1, j, m)], bt[g.idx(nx + 1, j, m)]));
}

template <class T>
void launch_shift_window_fields(const Grid& g, FieldSet<T>& F, cplx<T>*& spare, hipStream_t s) {
  const long n = g.plane * g.M;
  for (int k = 0; k < 9; ++k) {  // ex .. jt
    hipLaunchKernelGGL(k_shift_x<T>, dim3(nblk(n)), dim3(256), 0, s, g, F.f[k], spare);
    std::swap(F.f[k], spare);
  }
  hipLaunchKernelGGL(k_window_fix<T>, dim3((g.ny2 * g.M + 255) / 256), dim3(256), 0, s, g, F.ex(), F.er(), F.et(),
                     F.bx(), F.br(), F.bt(), F.bnd_xmax[0], F.bnd_xmax[1], F.bnd_xmax[2], F.bnd_xmax[3],
                     F.bnd_xmax[4], F.bnd_xmax[5]);
}

// ------------------------------------------------------------------------
// smooth_current (current_smooth.F90:50-58, smooth_mode_array :145-196): a
// pass writes the interior of `out` from the work array; the work array's
// interior is then refreshed from `out`
template <class T>
__global__ void k_smooth_pass(Grid g, const cplx<T>* __restrict__ wk, cplx<T>* __restrict__ out, T alpha, T beta,
                              int stride) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)g.nx * g.ny * g.M) return;
  const int ix = (int)(e % g.nx) + 1, iy = (int)((e / g.nx) % g.ny) + 1, m = (int)(e / ((long)g.nx * g.ny));
  const long k = g.idx(ix, iy, m);
  const long sx = stride, sy = (long)stride * g.nx2;
  const cplx<T> sum = cadd(cadd(cadd(wk[k - sx], wk[k + sx]), wk[k - sy]), wk[k + sy]);
  out[k] = cadd(rmul(alpha, wk[k]), rmul(beta, sum));
}

template <class T>
__global__ void k_copy_interior(Grid g, const cplx<T>* __restrict__ src, cplx<T>* __restrict__ dst) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)g.nx * g.ny * g.M) return;
  const int ix = (int)(e % g.nx) + 1, iy = (int)((e / g.nx) % g.ny) + 1, m = (int)(e / ((long)g.nx * g.ny));
  const long k = g.idx(ix, iy, m);
  dst[k] = src[k];
}

template <class T>
void launch_smooth_current(const Grid& g, FieldSet<T>& F, cplx<T>* wk, int its, int comp_its, const int* strides,
                           int nstrides, hipStream_t s) {
  const long n = g.plane * g.M, ni = (long)g.nx * g.ny * g.M;
  const bool per_max = g.bc_field[XMAX] == BC_PERIODIC, per_min = g.bc_field[XMIN] == BC_PERIODIC;
  cplx<T>* J[3] = {F.jx(), F.jr(), F.jt()};
  for (int c = 0; c < 3; ++c) {
    (void)hipMemcpyAsync(wk, J[c], n * sizeof(cplx<T>), hipMemcpyDeviceToDevice, s);
    double alpha = 0.5;
    const double beta = (1.0 - alpha) * 0.25;
    for (int it = 1; it <= its + comp_its; ++it) {
      for (int is = 0; is < nstrides; ++is) {
        if (per_max || per_min)  // field_mode_bc: the periodic exchange of the work array
          hipLaunchKernelGGL(k_halo_x<T>, dim3(nblk((long)g.ny2 * g.M)), dim3(256), 0, s, g, wk, 1 - g.ng,
                             g.ny + g.ng, (int)per_max, (int)per_min);
        hipLaunchKernelGGL(k_smooth_pass<T>, dim3(nblk(ni)), dim3(256), 0, s, g, wk, J[c], (T)alpha, (T)beta,
                           strides[is]);
        hipLaunchKernelGGL(k_copy_interior<T>, dim3(nblk(ni)), dim3(256), 0, s, g, J[c], wk);
      }
      if (it > its) alpha = (double)its * 0.5 + 1.0;
    }
  }
}

template <class T>
void launch_energy(const Grid& g, FieldSet<T>& F, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_energy<T>, dim3(g.ny, g.M), dim3(256), 0, s, g, F.ex(), F.er(), F.et(), F.bx(), F.br(),
                     F.bt(), out);
}

#define INST(T)                                                                               \
  template void launch_update_e<T>(const Grid&, FieldSet<T>&, hipStream_t);                   \
  template void launch_update_b<T>(const Grid&, FieldSet<T>&, int, hipStream_t);              \
  template void launch_efield_bcs<T>(const Grid&, FieldSet<T>&, hipStream_t);                 \
  template void launch_bfield_bcs<T>(const Grid&, FieldSet<T>&, int, hipStream_t);            \
  template void launch_setup_bnd<T>(const Grid&, FieldSet<T>&, hipStream_t);                  \
  template void launch_bfield_final_bcs<T>(const Grid&, FieldSet<T>&, const double*,          \
                                           const double*, hipStream_t);                       \
  template void launch_current_axis<T>(const Grid&, FieldSet<T>&, hipStream_t);               \
  template void launch_current_finish<T>(const Grid&, FieldSet<T>&, hipStream_t);             \
  template void launch_fill_zero<T>(cplx<T>*, long, hipStream_t);                             \
  template void launch_zero_j<T>(FieldSet<T>&, long, hipStream_t);                            \
  template void launch_energy<T>(const Grid&, FieldSet<T>&, double*, hipStream_t);            \
  template void launch_shift_window_fields<T>(const Grid&, FieldSet<T>&, cplx<T>*&, hipStream_t); \
  template void launch_smooth_current<T>(const Grid&, FieldSet<T>&, cplx<T>*, int, int, const int*, int,  \
                                         hipStream_t);
INST(double)
INST(float)

}  // namespace cylpic
