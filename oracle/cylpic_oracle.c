/*
 * cylpic_oracle.c -- FP64 CPU restatement of cylindrical EPOCH's quasi-3D PIC
 * timestep.  TEST INFRASTRUCTURE ONLY (see cylpic_oracle.h): the parity
 * checker for the HIP path, never part of the product.
 *
 * Every routine follows the Fortran text of the cited reference lines in the
 * same operation order (built with -ffp-contract=off), with one rank's
 * semantics for the MPI exchanges: the r=0 axis and r_max boundary are local,
 * x halos wrap onto this rank when x is periodic, r is never exchanged.
 *
 * Deviations from the reference, all at the last-ulp level:
 *   - the complex reciprocal 1/exp(-i theta) (particles.F90:323) is done with
 *     Smith's algorithm; the reference's compiler runtime may round it
 *     differently;
 *   - SQRT, ATAN2 come from this C library's libm.
 * These are "parity unpinned" at ulp level; everything else is pinned by the
 * tests under tests/ (closed forms, invariants, the worked deposition example
 * of the reference's documentation).
 */
#include "cylpic_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define C_LIGHT 2.99792458e8                                  /* constants.F90:175 */
#define EPSILON0 8.854187817620389850536563031710750e-12      /* constants.F90:179 */
#define PI 3.141592653589793238462643383279503                /* constants.F90:172 */
#define C_TINY 2.2250738585072014e-308                        /* TINY(1.0_num) */

/* Planted mistakes, for checking the checks: built with -DORACLE_PLANT=k
 * (oracle/Makefile target `plant`), mistake k replaces the restated term by
 * a plausible slip (a flipped sign, a dropped or swapped term, a wrong
 * factor).  tests/test_oracle_planted.py holds the table of mistakes and
 * requires each one to fail a committed test.  The default build (0) is the
 * restatement itself. */
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif
#define PLANT(k, good, bad) ((ORACLE_PLANT == (k)) ? (bad) : (good))

/* ------------------------------------------------------------------ */
/* indexing                                                           */
/* ------------------------------------------------------------------ */
#define NX2(g) ((g)->nx + 2 * (g)->ng)
#define NY2(g) ((g)->ny + 2 * (g)->ng)

/* Fortran f(ix, iy, m) with lower bounds (1-ng, 1-ng, 0) */
static inline size_t I3(const orc_grid *g, int ix, int iy, int m) {
  return ((size_t)m * (size_t)NY2(g) + (size_t)(iy + g->ng - 1)) * (size_t)NX2(g) +
         (size_t)(ix + g->ng - 1);
}
/* Fortran s(iy, m) with lower bounds (1-ng, 0) */
static inline size_t I2(const orc_grid *g, int iy, int m) {
  return (size_t)m * (size_t)NY2(g) + (size_t)(iy + g->ng - 1);
}

/* ------------------------------------------------------------------ */
/* complex arithmetic with Fortran's operand order                     */
/* ------------------------------------------------------------------ */
static inline ocplx cc(double re, double im) { ocplx z; z.re = re; z.im = im; return z; }
static inline ocplx cadd(ocplx a, ocplx b) { return cc(a.re + b.re, a.im + b.im); }
static inline ocplx csub(ocplx a, ocplx b) { return cc(a.re - b.re, a.im - b.im); }
static inline ocplx cneg(ocplx a) { return cc(-a.re, -a.im); }
static inline ocplx rmul(double s, ocplx a) { return cc(s * a.re, s * a.im); }
static inline ocplx cmulr(ocplx a, double s) { return cc(a.re * s, a.im * s); }
static inline ocplx cdivr(ocplx a, double s) { return cc(a.re / s, a.im / s); }
static inline ocplx cmul(ocplx a, ocplx b) {
  return cc(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
/* (0, s) * a : the product with a purely imaginary factor */
static inline ocplx imul(double s, ocplx a) { return cc(-(s * a.im), s * a.re); }
/* 1 / z by Smith's algorithm */
static inline ocplx crecip(ocplx z) {
  if (fabs(z.re) >= fabs(z.im)) {
    double r = z.im / z.re, den = z.re + z.im * r;
    return cc(1.0 / den, -r / den);
  } else {
    double r = z.re / z.im, den = z.re * r + z.im;
    return cc(r / den, -1.0 / den);
  }
}
static const ocplx CZERO = {0.0, 0.0};

/* nonzero once a push met a particle it cannot deposit (orc_take_error) */
static int orc_error = 0;
int orc_take_error(void) {
  const int e = orc_error;
  orc_error = 0;
  return e;
}

/* ------------------------------------------------------------------ */
/* field solver: src/fields.f90                                       */
/* ------------------------------------------------------------------ */

/* update_e_field, fields.f90:53-182 (E advanced by dt/2) */
void orc_update_e_field(const orc_grid *g, orc_fields *f) {
  const int nx = g->nx, ny = g->ny, M = g->n_mode, ng = g->ng;
  const double c2 = C_LIGHT * C_LIGHT, dt = g->dt, dx = g->dx, dy = g->dy;
  ocplx *ex = f->ex, *er = f->er, *et = f->et;
  const ocplx *bx = f->bx, *br = f->br, *bt = f->bt;
  const ocplx *jx = f->jx, *jr = f->jr, *jt = f->jt;

  /* interior, ir from 1 (r=0 axis is local) to ny, ix from 0 to nx  :62-113
   * (every element is independent: the rows run in parallel) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int m = 0; m < M; ++m) {
    for (int ir = 1; ir <= ny; ++ir) {
      double r_d = fabs((double)(ir - 1) * dy + g->y_grid_min);
      double r_p = r_d + 0.5 * dy;
      double fac_x = c2 / PLANT(5, r_p, r_d);
      double a_x = (double)m * fac_x;       /* im_fac_x = (0, a_x) */
      double a_r = ((double)m * c2) / PLANT(6, r_d, r_p);  /* im_fac_r = (0, a_r) */
      for (int ix = 0; ix <= nx; ++ix) {
        size_t k = I3(g, ix, ir, m);
        size_t kr1 = I3(g, ix, ir + 1, m);
        size_t kx1 = I3(g, ix + 1, ir, m);
        /* (11) exm */
        ocplx t1 = rmul(fac_x * 0.5, cadd(bt[kr1], bt[k]));
        ocplx t2 = imul(a_x, br[k]);
        ocplx t3 = cdivr(rmul(c2, csub(bt[kr1], bt[k])), dy);
        ocplx t4 = cdivr(jx[k], EPSILON0);
        ocplx s = csub(cadd(cadd(t1, t2), t3), t4);
        ex[k] = cadd(ex[k], cmulr(cmulr(s, 0.5), dt));
        /* (9) erm */
        t1 = cneg(imul(a_r, bx[k]));
        t2 = cdivr(rmul(c2, csub(bt[kx1], bt[k])), dx);
        t3 = cdivr(jr[k], EPSILON0);
        s = csub(csub(t1, t2), t3);
        er[k] = cadd(er[k], cmulr(cmulr(s, 0.5), dt));
        /* (10) etm */
        t1 = cdivr(rmul(c2, csub(br[kx1], br[k])), dx);
        t2 = cdivr(rmul(c2, csub(bx[kr1], bx[k])), dy);
        t3 = cdivr(jt[k], EPSILON0);
        s = csub(csub(t1, t2), t3);
        et[k] = cadd(et[k], cmulr(cmulr(s, 0.5), dt));
      }
    }
  }

  /* r_min rows and reflected ghosts, fields.f90:116-180 */
  for (int ix = 1 - ng; ix <= nx + ng; ++ix) {
    /* m = 0 */
    {
      const int m = 0;
      size_t k0 = I3(g, ix, 0, m), k1 = I3(g, ix, 1, m);
      ocplx s = csub(rmul((PLANT(1, 4.0, 2.0) * c2) / dy, bt[k1]), cdivr(jx[k0], EPSILON0));
      ex[k0] = cadd(ex[k0], cmulr(cmulr(s, 0.5), dt));
      et[k0] = CZERO;
      er[k0] = PLANT(4, cneg(er[k1]), er[k1]);
      for (int ir = 1 - ng; ir <= -1; ++ir) {
        et[I3(g, ix, ir, m)] = cneg(et[I3(g, ix, -ir, m)]);
        er[I3(g, ix, ir, m)] = cneg(er[I3(g, ix, -ir + 1, m)]);
        ex[I3(g, ix, ir, m)] = ex[I3(g, ix, -ir, m)];
      }
    }
    /* m = 1 */
    if (M > 1) {
      const int m = 1;
      size_t k0 = I3(g, ix, 0, m), k1 = I3(g, ix, 1, m), k2 = I3(g, ix, 2, m);
      ex[k0] = CZERO;
#if ORACLE_PLANT != 3
      er[k0] = csub(imul(2.0, et[k0]), er[k1]);
#endif
      ocplx w = csub(rmul(9.0, er[k1]), er[k2]);
      et[k0] = PLANT(2, cc(0.125 * w.im, -(0.125 * w.re)), cc(-(0.125 * w.im), 0.125 * w.re));
#if ORACLE_PLANT == 3
      er[k0] = csub(imul(2.0, et[k0]), er[k1]);
#endif
      for (int ir = 1 - ng; ir <= -1; ++ir) {
        et[I3(g, ix, ir, m)] = et[I3(g, ix, -ir, m)];
        er[I3(g, ix, ir, m)] = er[I3(g, ix, -ir + 1, m)];
        ex[I3(g, ix, ir, m)] = cneg(ex[I3(g, ix, -ir, m)]);
      }
    }
    /* m > 1 */
    double mode_sign = 1.0;
    for (int m = 2; m < M; ++m) {
      size_t k0 = I3(g, ix, 0, m), k1 = I3(g, ix, 1, m), k2 = I3(g, ix, 2, m);
      ex[k0] = CZERO;
      et[k0] = CZERO;
      er[k0] = cneg(er[k1]);
      er[k1] = cdivr(er[k2], PLANT(7, 9.0, 4.0));
      for (int ir = 1 - ng; ir <= -1; ++ir) {
        et[I3(g, ix, ir, m)] = cneg(rmul(mode_sign, et[I3(g, ix, -ir, m)]));
        er[I3(g, ix, ir, m)] = cneg(rmul(mode_sign, er[I3(g, ix, -ir + 1, m)]));
        ex[I3(g, ix, ir, m)] = rmul(mode_sign, ex[I3(g, ix, -ir, m)]);
      }
      mode_sign = -mode_sign;
    }
  }
}

/* update_b_field, fields.f90:186-312 (B advanced by dt/2) */
void orc_update_b_field(const orc_grid *g, orc_fields *f) {
  const int nx = g->nx, ny = g->ny, M = g->n_mode, ng = g->ng;
  const double dt = g->dt, dx = g->dx, dy = g->dy;
  ocplx *bx = f->bx, *br = f->br, *bt = f->bt;
  const ocplx *ex = f->ex, *er = f->er, *et = f->et;

  /* interior, ir from 1 to ny-1 (r_max is local), ix from 0 to nx  :203-241
   * (every element is independent: the rows run in parallel) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int m = 0; m < M; ++m) {
    for (int ir = 1; ir <= ny - 1; ++ir) {
      double r_d = fabs((double)(ir - 1) * dy + g->y_grid_min);
      double r_p = r_d + 0.5 * dy;
      double a_x = (double)m / r_d;  /* im_fac_x = (0, a_x) */
      double a_r = (double)m / r_p;  /* im_fac_r = (0, a_r) */
      for (int ix = 0; ix <= nx; ++ix) {
        size_t k = I3(g, ix, ir, m);
        size_t krm = I3(g, ix, ir - 1, m);
        size_t kxm = I3(g, ix - 1, ir, m);
        /* (8) bxm */
        ocplx t1 = imul(a_x, er[k]);
        ocplx t2 = cdivr(rmul(0.5, cadd(et[k], et[krm])), r_d);
        ocplx t3 = cdivr(csub(et[k], et[krm]), dy);
        ocplx s = cadd(cadd(t1, t2), t3);
        bx[k] = csub(bx[k], cmulr(cmulr(s, 0.5), dt));
        /* (6) brm */
        t1 = imul(a_r, ex[k]);
        t2 = cdivr(csub(et[k], et[kxm]), dx);
        s = cadd(t1, t2);
        br[k] = cadd(br[k], cmulr(cmulr(s, 0.5), dt));
        /* (7) btm */
        t1 = cneg(cdivr(csub(er[k], er[kxm]), dx));
        t2 = cdivr(csub(ex[k], ex[krm]), dy);
        s = cadd(t1, t2);
        bt[k] = cadd(bt[k], cmulr(cmulr(s, 0.5), dt));
      }
    }
  }

  /* r_min rows and reflected ghosts, fields.f90:249-310 */
  const double idy = 1.0 / dy;  /* imagi/dy = (0, 1/dy) */
  for (int ix = 1 - ng; ix <= nx + ng; ++ix) {
    {
      const int m = 0;
      size_t k0 = I3(g, ix, 0, m), k1 = I3(g, ix, 1, m);
      br[k0] = CZERO;
      bx[k0] = PLANT(8, bx[k1], cneg(bx[k1]));
      bt[k0] = cneg(bt[k1]);
      for (int ir = 1 - ng; ir <= -1; ++ir) {
        bt[I3(g, ix, ir, m)] = cneg(bt[I3(g, ix, -ir + 1, m)]);
        br[I3(g, ix, ir, m)] = cneg(br[I3(g, ix, -ir, m)]);
        bx[I3(g, ix, ir, m)] = bx[I3(g, ix, -ir + 1, m)];
      }
    }
    if (M > 1) {
      const int m = 1;
      size_t k0 = I3(g, ix, 0, m), k1 = I3(g, ix, 1, m);
      bx[k0] = cneg(bx[k1]);
      if (ix >= 2 - ng) {  /* (22) */
        ocplx t1 = imul(idy, ex[k0]);
        ocplx t2 = cdivr(csub(et[k0], et[I3(g, ix - 1, 0, m)]), dx);
        ocplx s = PLANT(10, cadd(t1, t2), csub(t1, t2));
        br[k0] = cadd(br[k0], cmulr(cmulr(s, 0.5), dt));
      }
      bt[k0] = PLANT(9, csub(imul(-2.0, br[k0]), bt[k1]), cadd(imul(-2.0, br[k0]), bt[k1]));
      for (int ir = 1 - ng; ir <= -1; ++ir) {
        bt[I3(g, ix, ir, m)] = bt[I3(g, ix, -ir + 1, m)];
        br[I3(g, ix, ir, m)] = br[I3(g, ix, -ir, m)];
        bx[I3(g, ix, ir, m)] = cneg(bx[I3(g, ix, -ir + 1, m)]);
      }
    }
    double mode_sign = 1.0;
    for (int m = 2; m < M; ++m) {
      size_t k0 = I3(g, ix, 0, m), k1 = I3(g, ix, 1, m);
      bx[k0] = cneg(bx[k1]);
      br[k0] = CZERO;
      bt[k0] = cneg(bt[k1]);
      for (int ir = 1 - ng; ir <= -1; ++ir) {
        bt[I3(g, ix, ir, m)] = cneg(rmul(mode_sign, bt[I3(g, ix, -ir + 1, m)]));
        br[I3(g, ix, ir, m)] = cneg(rmul(mode_sign, br[I3(g, ix, -ir, m)]));
        bx[I3(g, ix, ir, m)] = rmul(mode_sign, bx[I3(g, ix, -ir + 1, m)]);
      }
      mode_sign = -mode_sign;
    }
  }
}

/* ------------------------------------------------------------------ */
/* field boundaries: src/boundary.F90                                 */
/* ------------------------------------------------------------------ */

/* Stagger flags (stagger(c_dir_x/y, c_stagger_*m)), constants.F90:291-299
 * with the table of setup.F90:127-136. */
enum { S_EX, S_ER, S_ET, S_BX, S_BR, S_BT };
static const int STAG_X[6] = {0, 1, 1, 1, 0, 0};
static const int STAG_Y[6] = {1, 0, 1, 0, 1, 0};

/* do_field_mpi_with_lengths_cmplx on one rank, boundary.F90:500-593:
 * only the x exchange can be non-trivial (periodic onto itself).
 * rows [row_lo, row_hi] are the rows that reach the caller (the r-shift
 * copy of efield_bcs/bfield_bcs drops the top row, boundary.F90:1362-1371). */
static void halo_x(const orc_grid *g, ocplx *F, int row_lo, int row_hi) {
  const int nx = g->nx, ng = g->ng;
  for (int m = 0; m < g->n_mode; ++m)
    for (int j = row_lo; j <= row_hi; ++j) {
      if (g->bc_field[ORC_XMAX] == ORC_BC_PERIODIC)
        for (int i = 1; i <= ng; ++i) F[I3(g, nx + i, j, m)] = F[I3(g, i, j, m)];
      if (g->bc_field[ORC_XMIN] == ORC_BC_PERIODIC)
        for (int i = 1; i <= ng; ++i) F[I3(g, i - ng, j, m)] = F[I3(g, nx - ng + i, j, m)];
    }
}

/* field_mode_clamp_zero (sign = -1) and field_mode_zero_gradient (sign = +1),
 * boundary.F90:654-707 and :772-829; called only for x_min, x_max, y_max. */
static void mode_bc(const orc_grid *g, ocplx *F, int stag, int bnd, int clamp) {
  const int nx = g->nx, ny = g->ny, ng = g->ng, M = g->n_mode;
  if (g->bc_field[bnd] == ORC_BC_PERIODIC) return;
#define SETV(a, b) F[a] = clamp ? cneg(F[b]) : F[b]
  for (int m = 0; m < M; ++m) {
    if (bnd == ORC_XMIN) {
      for (int j = 1 - ng; j <= ny + ng; ++j) {
        if (STAG_X[stag]) {
          for (int i = 1; i <= ng - 1; ++i) SETV(I3(g, i - ng, j, m), I3(g, ng - i, j, m));
          if (clamp) F[I3(g, 0, j, m)] = CZERO;
        } else {
          for (int i = 1; i <= ng; ++i) SETV(I3(g, i - ng, j, m), I3(g, ng + 1 - i, j, m));
        }
      }
    } else if (bnd == ORC_XMAX) {
      for (int j = 1 - ng; j <= ny + ng; ++j) {
        if (STAG_X[stag]) {
          if (clamp) F[I3(g, nx, j, m)] = CZERO;
          for (int i = 1; i <= ng - 1; ++i) SETV(I3(g, nx + i, j, m), I3(g, nx - i, j, m));
        } else {
          for (int i = 1; i <= ng; ++i) SETV(I3(g, nx + i, j, m), I3(g, nx + 1 - i, j, m));
        }
      }
    } else if (bnd == ORC_YMAX) {
      if (STAG_Y[stag]) {
        if (clamp)
          for (int ix = 1 - ng; ix <= nx + ng; ++ix) F[I3(g, ix, ny, m)] = CZERO;
        for (int i = 1; i <= ng - 1; ++i)
          for (int ix = 1 - ng; ix <= nx + ng; ++ix) SETV(I3(g, ix, ny + i, m), I3(g, ix, ny - i, m));
      } else {
        for (int i = 1; i <= ng; ++i)
          for (int ix = 1 - ng; ix <= nx + ng; ++ix) SETV(I3(g, ix, ny + i, m), I3(g, ix, ny + 1 - i, m));
      }
    }
  }
#undef SETV
}

/* the generic clamp / zero-gradient loop shared by efield_bcs and bfield_bcs */
static void generic_edges(const orc_grid *g, ocplx *a, ocplx *b, ocplx *c,
                          int sa, int sb, int sc) {
  for (int i = ORC_XMIN; i <= ORC_YMAX; ++i) {
    if (i == ORC_YMIN) continue;
    int bc = g->bc_field[i];
    if (bc == ORC_BC_CLAMP || bc == ORC_BC_SIMPLE_LASER || bc == ORC_BC_SIMPLE_OUTFLOW) {
      mode_bc(g, a, sa, i, 1);
      mode_bc(g, b, sb, i, 1);
      mode_bc(g, c, sc, i, 1);
    }
    if (bc == ORC_BC_ZERO_GRADIENT) {
      mode_bc(g, a, sa, i, 0);
      mode_bc(g, b, sb, i, 0);
      mode_bc(g, c, sc, i, 0);
    }
  }
}

/* efield_bcs, boundary.F90:1355-1413 */
void orc_efield_bcs(const orc_grid *g, orc_fields *f) {
  const int ny = g->ny, ng = g->ng;
  halo_x(g, f->ex, 1 - ng, ny + ng - 1);
  halo_x(g, f->er, 1 - ng, ny + ng);
  halo_x(g, f->et, 1 - ng, ny + ng - 1);
  for (int i = ORC_XMIN; i <= ORC_XMAX; ++i) {
    if (g->bc_field[i] == ORC_BC_CONDUCT) {
      mode_bc(g, f->ex, S_EX, i, 1);
      mode_bc(g, f->er, S_ER, i, 0);
      mode_bc(g, f->et, S_ET, i, 0);
    }
  }
  if (g->bc_field[ORC_YMAX] == ORC_BC_CONDUCT) {
    mode_bc(g, f->ex, S_EX, ORC_YMAX, 0);
    mode_bc(g, f->er, S_ER, ORC_YMAX, 1);
    mode_bc(g, f->et, S_ET, ORC_YMAX, 0);
  }
  generic_edges(g, f->ex, f->er, f->et, S_EX, S_ER, S_ET);
}

/* bfield_bcs, boundary.F90:1417-1476 */
void orc_bfield_bcs(const orc_grid *g, orc_fields *f, int mpi_only) {
  const int ny = g->ny, ng = g->ng;
  halo_x(g, f->bx, 1 - ng, ny + ng);
  halo_x(g, f->br, 1 - ng, ny + ng - 1);
  halo_x(g, f->bt, 1 - ng, ny + ng);
  if (mpi_only) return;
  for (int i = ORC_XMIN; i <= ORC_XMAX; ++i) {
    if (g->bc_field[i] == ORC_BC_CONDUCT) {
      mode_bc(g, f->bx, S_BX, i, 0);
      mode_bc(g, f->br, S_BR, i, 1);
      mode_bc(g, f->bt, S_BT, i, 1);
    }
  }
  if (g->bc_field[ORC_YMAX] == ORC_BC_CONDUCT) {
    mode_bc(g, f->bx, S_BX, ORC_YMAX, 1);
    mode_bc(g, f->br, S_BR, ORC_YMAX, 0);
    mode_bc(g, f->bt, S_BT, ORC_YMAX, 1);
  }
  generic_edges(g, f->bx, f->br, f->bt, S_BX, S_BR, S_BT);
}

/* setup_field_boundaries, setup.F90:393-420 (no CPML: nx0 = 1, nx1 = nx) */
void orc_setup_field_boundaries(const orc_grid *g, orc_fields *f) {
  const int nx = g->nx, ny = g->ny, ng = g->ng;
  for (int m = 0; m < g->n_mode; ++m)
    for (int j = 1 - ng; j <= ny + ng; ++j) {
      size_t s = I2(g, j, m);
      f->bnd_xmin[0][s] = rmul(0.5, cadd(f->ex[I3(g, 1, j, m)], f->ex[I3(g, 0, j, m)]));
      f->bnd_xmin[1][s] = f->er[I3(g, 0, j, m)];
      f->bnd_xmin[2][s] = f->et[I3(g, 0, j, m)];
      f->bnd_xmax[0][s] = rmul(0.5, cadd(f->ex[I3(g, nx, j, m)], f->ex[I3(g, nx + 1, j, m)]));
      f->bnd_xmax[1][s] = f->er[I3(g, nx, j, m)];
      f->bnd_xmax[2][s] = f->et[I3(g, nx, j, m)];
      f->bnd_xmin[3][s] = f->bx[I3(g, 0, j, m)];
      f->bnd_xmin[4][s] = rmul(0.5, cadd(f->br[I3(g, 1, j, m)], f->br[I3(g, 0, j, m)]));
      f->bnd_xmin[5][s] = rmul(0.5, cadd(f->bt[I3(g, 1, j, m)], f->bt[I3(g, 0, j, m)]));
      f->bnd_xmax[3][s] = f->bx[I3(g, nx, j, m)];
      f->bnd_xmax[4][s] = rmul(0.5, cadd(f->br[I3(g, nx, j, m)], f->br[I3(g, nx + 1, j, m)]));
      f->bnd_xmax[5][s] = rmul(0.5, cadd(f->bt[I3(g, nx, j, m)], f->bt[I3(g, nx + 1, j, m)]));
    }
}

/* The m=1 plane-wave sources of laser.f90:476-489: source_t = s1 + i s2,
 * source_r = -i s1 + s2. */
static inline ocplx src_t(const double *src, int ny, int ir) {
  return cc(src[ir], PLANT(21, src[ny + 1 + ir], -src[ny + 1 + ir]));
}
static inline ocplx src_r(const double *src, int ny, int ir) {
  return cc(src[ny + 1 + ir], -src[ir]);
}

/* outflow_bcs_x_min, laser.f90:411-520.  src = source1(0:ny), source2(0:ny)
 * as evaluated by the host, or NULL (no laser attached: zero sources).
 * The r_d_vals(0:ny) term of the Btm update pairs with bxm(1,1:ny): the
 * arrays are non-conformable and the reference's element-wise evaluation
 * pairs row ir with r_d_vals(ir-1); that pairing is kept. */
void orc_outflow_bcs_x_min(const orc_grid *g, orc_fields *f, const double *src) {
  const int ny = g->ny, M = g->n_mode;
  const double c = C_LIGHT, dt = g->dt;
  const double dtc2 = dt * (c * c), lx = dtc2 / g->dx, lr = dtc2 / g->dy;
  const double sum = 1.0 / (lx + c), diff = lx - c, dt_eps = dt / EPSILON0;
  for (int m = 0; m < M; ++m)
    for (int ir = 0; ir <= ny; ++ir) f->bx[I3(g, 0, ir, m)] = f->bnd_xmin[3][I2(g, ir, m)];
  for (int m = 0; m < M; ++m) {
    for (int ir = 1; ir <= ny; ++ir) {
      ocplx st = (m == 1 && src) ? src_t(src, ny, ir) : CZERO;
      double r_d = fabs((double)(ir - 1 - PLANT(29, 1, 0)) * g->dy + g->y_grid_min);
      ocplx a = rmul(PLANT(20, 4.0, 2.0), st);
      ocplx b = rmul(2.0, cadd(f->bnd_xmin[1][I2(g, ir, m)], rmul(c, f->bnd_xmin[5][I2(g, ir, m)])));
      ocplx d = rmul(2.0, f->er[I3(g, 1, ir, m)]);
      ocplx e = cdivr(imul(((double)m * (c * c)) * dt, f->bx[I3(g, 1, ir, m)]), r_d);
      ocplx h = rmul(dt_eps, f->jr[I3(g, 1, ir, m)]);
      ocplx k = rmul(PLANT(22, diff, -diff), f->bt[I3(g, 2, ir, m)]);
      f->bt[I3(g, 1, ir, m)] = rmul(sum, cadd(cadd(cadd(csub(cadd(a, b), d), e), h), k));
    }
    const int ir_l = 1, ir_h = ny - 1;
    for (int ir = ir_l; ir <= ir_h; ++ir) {
      ocplx sr = (m == 1 && src) ? src_r(src, ny, ir) : CZERO;
      ocplx a = rmul(-4.0, sr);
      ocplx b = rmul(2.0, cadd(f->bnd_xmin[2][I2(g, ir, m)], rmul(c, f->bnd_xmin[4][I2(g, ir, m)])));
      ocplx d = rmul(2.0, f->et[I3(g, PLANT(23, 1, 2), ir, m)]);
      ocplx e = rmul(lr, csub(f->bx[I3(g, 1, ir + 1, m)], f->bx[I3(g, 1, ir, m)]));
      ocplx h = rmul(dt_eps, f->jt[I3(g, 1, ir, m)]);
      ocplx k = rmul(diff, f->br[I3(g, 2, ir, m)]);
      f->br[I3(g, 1, ir, m)] = rmul(sum, cadd(csub(csub(cadd(csub(a, b), d), e), h), k));
    }
  }
}

/* outflow_bcs_x_max, laser.f90:524-633.  Both source_t(0:ny) and
 * r_d_vals(0:ny) pair with rows 1:ny offset by one, as in the reference. */
void orc_outflow_bcs_x_max(const orc_grid *g, orc_fields *f, const double *src) {
  const int nx = g->nx, ny = g->ny, M = g->n_mode;
  const double c = C_LIGHT, dt = g->dt;
  const double dtc2 = dt * (c * c), lx = dtc2 / g->dx, lr = dtc2 / g->dy;
  const double sum = 1.0 / PLANT(24, lx + c, lx), diff = lx - c, dt_eps = dt / EPSILON0;
  for (int m = 0; m < M; ++m)
    for (int ir = 0; ir <= ny; ++ir) f->bx[I3(g, nx, ir, m)] = f->bnd_xmax[3][I2(g, ir, m)];
  for (int m = 0; m < M; ++m) {
    for (int ir = 1; ir <= ny; ++ir) {
      ocplx st = (m == 1 && src) ? src_t(src, ny, ir - 1) : CZERO;
      double r_d = fabs((double)(ir - 1 - 1) * g->dy + g->y_grid_min);
      ocplx a = rmul(-4.0, st);
      ocplx b = rmul(2.0, cadd(f->bnd_xmax[1][I2(g, ir, m)], rmul(c, f->bnd_xmax[5][I2(g, ir, m)])));
      ocplx d = rmul(2.0, f->er[I3(g, nx - 1, ir, m)]);
      ocplx e = cdivr(imul(((double)m * (c * c)) * dt, f->bx[I3(g, nx - 1, ir, m)]), r_d);
      ocplx h = rmul(dt_eps, f->jr[I3(g, nx - 1, ir, m)]);
      ocplx k = rmul(diff, f->bt[I3(g, nx - 1, ir, m)]);
      f->bt[I3(g, nx, ir, m)] = rmul(sum, cadd(csub(csub(cadd(csub(a, b), d), e), h), k));
    }
    const int ir_l = 1, ir_h = ny - 1;
    for (int ir = ir_l; ir <= ir_h; ++ir) {
      ocplx sr = (m == 1 && src) ? src_r(src, ny, ir) : CZERO;
      ocplx a = rmul(4.0, sr);
      ocplx b = rmul(2.0, cadd(f->bnd_xmax[2][I2(g, ir, m)], rmul(c, f->bnd_xmax[4][I2(g, ir, m)])));
      ocplx d = rmul(2.0, f->et[I3(g, nx - 1, ir, m)]);
      ocplx e = rmul(lr, csub(f->bx[I3(g, nx - 1, ir + 1, m)], f->bx[I3(g, nx - 1, ir, m)]));
      ocplx h = rmul(dt_eps, f->jt[I3(g, nx - 1, ir, m)]);
      ocplx k = rmul(diff, f->br[I3(g, nx - 1, ir, m)]);
      f->br[I3(g, nx, ir, m)] = rmul(sum, cadd(cadd(cadd(csub(cadd(a, b), d), e), h), k));
    }
  }
}

/* outflow_bcs_r_max, laser.f90:637-690 */
void orc_outflow_bcs_r_max(const orc_grid *g, orc_fields *f) {
  const int nx = g->nx, ny = g->ny, M = g->n_mode;
  const double c = C_LIGHT, dt = g->dt;
  const double dtc2 = dt * (c * c);
  const double inv_r = 1.0 / (((double)ny - 1.5) * g->dy + g->y_grid_min);
  const double dtc2_4r = 0.25 * dtc2 * inv_r;
  const double icdt = ((0.5 * c) * dt) * inv_r;  /* icdt_2r = (0, icdt) */
  const double lx = dtc2 / g->dx, ly = dtc2 / g->dy;
  const double sum_x = 1.0 / (ly + c), sum_t = 1.0 / ((ly + c) + dtc2_4r);
  const double dt_2eps = (0.5 * dt) / EPSILON0;
  const int ix_l = 1, ix_h = nx - 1;
  for (int m = 0; m < M; ++m) {
    for (int ix = ix_l; ix <= ix_h; ++ix) {
      ocplx a = cneg(cmulr(f->bx[I3(g, ix, ny - 1, m)], c - ly));
      ocplx b = cmulr(PLANT(27, f->bx_old, f->bx)[I3(g, ix, ny, m)], -c + ly);
      ocplx d = cmulr(f->bx_old[I3(g, ix, ny - 1, m)], PLANT(28, -c - ly, c + ly));
      ocplx e = rmul((c * dt) * inv_r, f->et[I3(g, ix, ny - 1, m)]);
      ocplx h = rmul(0.5 * lx,
                     csub(cadd(csub(f->br[I3(g, ix + 1, ny - 1, m)], f->br[I3(g, ix, ny - 1, m)]),
                               f->br_old[I3(g, ix + 1, ny - 1, m)]),
                          f->br_old[I3(g, ix, ny - 1, m)]));
      ocplx k = imul(icdt * (double)m, cadd(f->er[I3(g, ix, ny, m)], f->er[I3(g, ix, ny - 1, m)]));
      ocplx l = rmul(dt_2eps, cadd(f->jt[I3(g, ix, ny - 1, m)], f->jt_old[I3(g, ix, ny - 1, m)]));
      f->bx[I3(g, ix, ny, m)] =
          rmul(PLANT(25, sum_x, sum_t), csub(csub(cadd(csub(csub(csub(a, b), d), e), h), k), l));
    }
    for (int ix = 1; ix <= nx; ++ix) {
      ocplx a = cneg(cmulr(f->bt[I3(g, ix, ny - 1, m)], (c - ly) + dtc2_4r));
      ocplx b = cmulr(PLANT(30, f->bt_old, f->bt)[I3(g, ix, ny, m)], (-c + ly) + dtc2_4r);
      ocplx d = cmulr(f->bt_old[I3(g, ix, ny - 1, m)], (-c - ly) + dtc2_4r);
      ocplx e = rmul(PLANT(26, 1.0, -1.0) * ((0.5 * lx) / c),
                     csub(csub(cadd(f->er[I3(g, ix, ny, m)], f->er[I3(g, ix, ny - 1, m)]),
                               f->er[I3(g, ix - 1, ny, m)]),
                          f->er[I3(g, ix - 1, ny - 1, m)]));
      ocplx k = imul((icdt * (double)m) * c,
                     cadd(f->br[I3(g, ix, ny - 1, m)], f->br_old[I3(g, ix, ny - 1, m)]));
      ocplx l = rmul(dt_2eps, cadd(f->jx[I3(g, ix, ny - 1, m)], f->jx_old[I3(g, ix, ny - 1, m)]));
      f->bt[I3(g, ix, ny, m)] = rmul(sum_t, cadd(csub(csub(csub(csub(a, b), d), e), k), l));
    }
  }
}

/* bfield_final_bcs, boundary.F90:1505-1537 (laser phases are host-side) */
void orc_bfield_final_bcs(const orc_grid *g, orc_fields *f,
                          const double *src_xmin, const double *src_xmax) {
  orc_bfield_bcs(g, f, 0);
  int b = g->bc_field[ORC_XMIN];
  if (b == ORC_BC_SIMPLE_LASER || b == ORC_BC_SIMPLE_OUTFLOW)
    orc_outflow_bcs_x_min(g, f, b == ORC_BC_SIMPLE_LASER ? src_xmin : NULL);
  b = g->bc_field[ORC_XMAX];
  if (b == ORC_BC_SIMPLE_LASER || b == ORC_BC_SIMPLE_OUTFLOW)
    orc_outflow_bcs_x_max(g, f, b == ORC_BC_SIMPLE_LASER ? src_xmax : NULL);
  b = g->bc_field[ORC_YMAX];
  if (b == ORC_BC_SIMPLE_OUTFLOW) {
    orc_outflow_bcs_r_max(g, f);
  } else if (b == ORC_BC_ZERO_B) {
    for (int m = 0; m < g->n_mode; ++m)
      for (int ix = 1 - g->ng; ix <= g->nx + g->ng; ++ix) {
        size_t k = I3(g, ix, g->ny, m);
        f->bx[k] = CZERO; f->br[k] = CZERO; f->bt[k] = CZERO;
      }
  }
  orc_bfield_bcs(g, f, 1);
}

/* update_eb_fields_half, fields.f90:316-337 */
void orc_update_eb_fields_half(const orc_grid *g, orc_fields *f) {
  size_t n = (size_t)NX2(g) * NY2(g) * g->n_mode * sizeof(ocplx);
  orc_update_e_field(g, f);
  orc_efield_bcs(g, f);
  memcpy(f->bx_old, f->bx, n);
  memcpy(f->br_old, f->br, n);
  memcpy(f->bt_old, f->bt, n);
  orc_update_b_field(g, f);
  orc_bfield_bcs(g, f, 1);
}

/* update_eb_fields_final, fields.f90:341-353 */
void orc_update_eb_fields_final(const orc_grid *g, orc_fields *f,
                                const double *src_xmin, const double *src_xmax) {
  orc_update_b_field(g, f);
  orc_bfield_final_bcs(g, f, src_xmin, src_xmax);
  orc_update_e_field(g, f);
  orc_efield_bcs(g, f);
}

/* ------------------------------------------------------------------ */
/* currents: src/boundary.F90, src/housekeeping/current_smooth.F90     */
/* ------------------------------------------------------------------ */

/* particle_reflection_bcs_complex, boundary.F90:918-1015, one mode plane */
static void reflect_current(const orc_grid *g, ocplx *J, int m, int flip_dir) {
  const int nx = g->nx, ny = g->ny, ng = g->ng;
  if (g->bc_particle[ORC_XMIN] == ORC_BC_REFLECT) {
    for (int j = 1 - ng; j <= ny + ng; ++j) {
      if (flip_dir == 1) {
        for (int i = 1; i <= ng - 1; ++i) {
          J[I3(g, i, j, m)] = csub(J[I3(g, i, j, m)], J[I3(g, 1 - i, j, m)]);
          J[I3(g, 1 - i, j, m)] = CZERO;
        }
      } else {
        for (int i = 1; i <= ng - 1; ++i) {
          J[I3(g, i, j, m)] = cadd(J[I3(g, i, j, m)], J[I3(g, -i, j, m)]);
          J[I3(g, -i, j, m)] = CZERO;
        }
      }
    }
  }
  if (g->bc_particle[ORC_XMAX] == ORC_BC_REFLECT) {
    for (int j = 1 - ng; j <= ny + ng; ++j) {
      if (flip_dir == 1) {
        for (int i = 1; i <= ng; ++i) {
          J[I3(g, nx + 1 - i, j, m)] = csub(J[I3(g, nx + 1 - i, j, m)], J[I3(g, nx + i, j, m)]);
          J[I3(g, nx + i, j, m)] = CZERO;
        }
      } else {
        for (int i = 1; i <= ng; ++i) {
          J[I3(g, nx - i, j, m)] = cadd(J[I3(g, nx - i, j, m)], J[I3(g, nx + i, j, m)]);
          J[I3(g, nx + i, j, m)] = CZERO;
        }
      }
    }
  }
  if (g->bc_particle[ORC_YMAX] == ORC_BC_REFLECT) {
    const double r_max = g->y_grid_min + ((double)ny - 0.5) * g->dy;
    for (int i = 1; i <= ng; ++i) {
      for (int ix = 1 - ng; ix <= nx + ng; ++ix) {
        if (flip_dir == 2) {
          double num = r_max + ((double)i - 0.5) * g->dy;
          double den = r_max - ((double)i - 0.5) * g->dy;
          if (ORACLE_PLANT == 19) { double t = num; num = den; den = t; }
          J[I3(g, ix, ny + 1 - i, m)] =
              csub(J[I3(g, ix, ny + 1 - i, m)], cdivr(cmulr(J[I3(g, ix, ny + i, m)], num), den));
        } else if (flip_dir == 1) {
          double num = r_max + (double)i * g->dy;
          double den = r_max - (double)i * g->dy;
          J[I3(g, ix, ny - i, m)] =
              cadd(J[I3(g, ix, ny - i, m)], cdivr(cmulr(J[I3(g, ix, ny + i, m)], num), den));
        } else {
          J[I3(g, ix, ny - i, m)] = cadd(J[I3(g, ix, ny - i, m)], J[I3(g, ix, ny + i, m)]);
        }
        J[I3(g, ix, ny + i, m)] = CZERO;
      }
    }
  }
}

/* particle_periodic_bcs_complex on one rank, boundary.F90:1133-1245 */
static void periodic_current(const orc_grid *g, ocplx *J, int m) {
  const int nx = g->nx, ny = g->ny, ng = g->ng;
  if (!(g->bc_particle[ORC_XMIN] == ORC_BC_PERIODIC && g->bc_particle[ORC_XMAX] == ORC_BC_PERIODIC))
    return;
  for (int j = 1 - ng; j <= ny + ng; ++j)
    for (int i = 1; i <= ng; ++i)
      J[I3(g, i, j, m)] = cadd(J[I3(g, i, j, m)], J[I3(g, nx + i, j, m)]);
  if (ORACLE_PLANT == 18) return;
  for (int j = 1 - ng; j <= ny + ng; ++j)
    for (int i = 1; i <= ng; ++i)
      J[I3(g, nx - ng + i, j, m)] = cadd(J[I3(g, nx - ng + i, j, m)], J[I3(g, i - ng, j, m)]);
}

/* current_bcs without a species, boundary.F90:1893-1905 -> :1330-1351 */
void orc_current_bcs(const orc_grid *g, orc_fields *f) {
  for (int m = 0; m < g->n_mode; ++m) {
    reflect_current(g, f->jx, m, 1);
    periodic_current(g, f->jx, m);
    reflect_current(g, f->jr, m, 2);
    periodic_current(g, f->jr, m);
    reflect_current(g, f->jt, m, 3);
    periodic_current(g, f->jt, m);
  }
}

/* current_bcs_r_min_final, boundary.F90:1909-1959 */
void orc_current_bcs_r_min_final(const orc_grid *g, orc_fields *f) {
  const int nx = g->nx, ng = g->ng, M = g->n_mode;
  ocplx *jx = f->jx, *jr = f->jr, *jt = f->jt;
  double mode_sign = 1.0;
  for (int m = 0; m < M; ++m) {
    for (int j = 2; j <= ng; ++j)
      for (int ix = 1 - ng; ix <= nx + ng; ++ix) {
        size_t a = I3(g, ix, 1 - j, m);
        jx[I3(g, ix, j - 1, m)] = PLANT(12, cadd(jx[I3(g, ix, j - 1, m)], rmul(mode_sign, jx[a])),
                                        csub(jx[I3(g, ix, j - 1, m)], rmul(mode_sign, jx[a])));
        jr[I3(g, ix, j, m)] = csub(jr[I3(g, ix, j, m)], rmul(mode_sign, jr[a]));
        jt[I3(g, ix, j - 1, m)] = PLANT(11, csub(jt[I3(g, ix, j - 1, m)], rmul(mode_sign, jt[a])),
                                        cadd(jt[I3(g, ix, j - 1, m)], rmul(mode_sign, jt[a])));
        jx[a] = CZERO; jr[a] = CZERO; jt[a] = CZERO;
      }
    for (int ix = 1 - ng; ix <= nx + ng; ++ix)
      if (ORACLE_PLANT != 13)
        jr[I3(g, ix, 1, m)] = csub(jr[I3(g, ix, 1, m)], rmul(mode_sign, jr[I3(g, ix, 0, m)]));
    mode_sign = PLANT(17, -mode_sign, mode_sign);
  }
  for (int m = 0; m < M; ++m)
    for (int ix = 1 - ng; ix <= nx + ng; ++ix) {
      size_t k0 = I3(g, ix, 0, m), k1 = I3(g, ix, 1, m), k2 = I3(g, ix, 2, m);
      if (m > 0) jx[k0] = CZERO;
      else jx[k0] = cdivr(csub(rmul(4.0, jx[k1]), jx[k2]), PLANT(14, 3.0, 4.0));
      if (m == 1) {
        ocplx w = csub(rmul(9.0, jr[k1]), jr[k2]);
        jt[k0] = PLANT(15, cc(w.im / 8.0, -(w.re / 8.0)), cc(-(w.im / 8.0), w.re / 8.0));
        jr[k0] = PLANT(16, csub(imul(2.0, jt[k0]), jr[k1]), cadd(imul(2.0, jt[k0]), jr[k1]));
      } else {
        jt[k0] = CZERO;
        jr[k0] = cneg(jr[k1]);
      }
    }
}

/* current_finish, current_smooth.F90:29-46 (no smoothing, no correction) */
void orc_current_finish(const orc_grid *g, orc_fields *f) {
  orc_current_bcs(g, f);
  halo_x(g, f->jx, 1 - g->ng, g->ny + g->ng);
  halo_x(g, f->jr, 1 - g->ng, g->ny + g->ng);
  halo_x(g, f->jt, 1 - g->ng, g->ny + g->ng);
}

/* smooth_mode_array, current_smooth.F90:145-196: strided binomial passes
 * over the interior; the work array's ghost cells keep their values except
 * for the periodic exchange (field_mode_bc), strides <= ng.  alpha becomes
 * its/2 + 1 after each pass beyond `its` (beta stays 1/8), as written. */
static void smooth_mode_array(const orc_grid *g, ocplx *array, int its, int comp_its, const int *strides,
                              int nstrides) {
  const int nx = g->nx, ny = g->ny, M = g->n_mode;
  const size_t n = (size_t)g->n_mode * (nx + 2 * g->ng) * (ny + 2 * g->ng);
  ocplx *wk = malloc(n * sizeof(ocplx));
  memcpy(wk, array, n * sizeof(ocplx));
  double alpha = 0.5;
  const double beta = (1.0 - alpha) * 0.25;
  const int one = 1;
  if (nstrides <= 0) {
    strides = &one;
    nstrides = 1;
  }
  for (int it = 1; it <= its + comp_its; ++it) {
    for (int is = 0; is < nstrides; ++is) {
      halo_x(g, wk, 1 - g->ng, ny + g->ng);
      const int s = strides[is];
      for (int m = 0; m < M; ++m)
        for (int iy = 1; iy <= ny; ++iy)
          for (int ix = 1; ix <= nx; ++ix) {
            const ocplx sum = cadd(cadd(cadd(wk[I3(g, ix - s, iy, m)], wk[I3(g, ix + s, iy, m)]),
                                        wk[I3(g, ix, iy - s, m)]),
                                   wk[I3(g, ix, iy + s, m)]);
            array[I3(g, ix, iy, m)] = cadd(rmul(alpha, wk[I3(g, ix, iy, m)]), rmul(beta, sum));
          }
      for (int m = 0; m < M; ++m)
        for (int iy = 1; iy <= ny; ++iy)
          for (int ix = 1; ix <= nx; ++ix) wk[I3(g, ix, iy, m)] = array[I3(g, ix, iy, m)];
    }
    if (it > PLANT(43, its, its - 1)) alpha = (double)its * 0.5 + 1.0;
  }
  free(wk);
}

/* smooth_current, current_smooth.F90:50-58 */
void orc_smooth_current(const orc_grid *g, orc_fields *f, int its, int comp_its, const int *strides,
                        int nstrides) {
  smooth_mode_array(g, f->jx, its, comp_its, strides, nstrides);
  smooth_mode_array(g, f->jr, its, comp_its, strides, nstrides);
  smooth_mode_array(g, f->jt, its, comp_its, strides, nstrides);
}

/* ------------------------------------------------------------------ */
/* particle push and deposit: src/particles.F90                       */
/* ------------------------------------------------------------------ */

/* per-step geometry tables, particles.F90:190-217.  Arrays are offset so
 * that element [iy + ng] is Fortran index iy (inv_area_xt from 0-jng). */
void orc_geometry(const orc_grid *g, double *inv_area_rt, double *inv_area_xt,
                  double *inv_volume, double *ratio_area_xt) {
  const int ny = g->ny, jng = g->ng;
  const double dx = g->dx, dy = g->dy;
  double r_low = g->y_grid_min - (double)jng * dy;
  inv_area_xt[0 - jng + jng] = 1.0 / (((2.0 * PI) * fabs(r_low)) * dx);
  for (int iy = 1 - jng; iy <= ny + jng; ++iy) {
    if (lround((2.0 * r_low) / dy) == -1) {
      double h = 0.5 * dy;
      inv_area_rt[iy + jng] = 1.0 / (PI * (h * h));
    } else {
      double rh = r_low + dy;
      inv_area_rt[iy + jng] = 1.0 / (PI * fabs(rh * rh - r_low * r_low));
    }
    inv_area_xt[iy + jng] = 1.0 / (((2.0 * PI) * fabs(r_low + dy)) * dx);
    r_low = r_low + dy;
  }
  for (int iy = 1 - jng; iy <= ny + jng; ++iy) {
    inv_volume[iy + jng] = inv_area_rt[iy + jng] / dx;
    ratio_area_xt[iy + jng] = inv_area_xt[iy + jng] / inv_area_xt[iy - 1 + jng];
  }
}

/* triangle/gx.inc (centred weights, sum 2) */
static inline void weights3(double frac, double *w /* [-1..1] */) {
  double cf2 = frac * frac;
  w[-1] = (0.25 + cf2) + frac;
  w[0] = 1.5 - 2.0 * cf2;
  w[1] = (0.25 + cf2) - frac;
}

/* The DO im / DO iy / DO ix nest of particles.F90:586-665. */
void orc_deposit_from_weights(const orc_grid *g, orc_fields *f,
                              const double *gx_, const double *hx_,
                              const double *gy_, const double *hy_,
                              int cell_x2, int cell_y2, int dcellx, int dcelly,
                              double fcx, double fcy, double fcz,
                              ocplx exp_itheta_05, ocplx exp_idtheta,
                              double dtheta,
                              const double *inv_area_rt,
                              const double *inv_area_xt,
                              const double *inv_volume,
                              const double *ratio_area_xt) {
  const double *gx = gx_ + 2, *hx = hx_ + 2, *gy = gy_ + 2, *hy = hy_ + 2;
  const int jng = g->ng;
  const double third = 1.0 / 3.0, sixth = 0.5 * third;
  const int xmin = -1 + (dcellx - 1) / 2, xmax = 1 + (dcellx + 1) / 2;
  const int ymin = -1 + (dcelly - 1) / 2, ymax = 1 + (dcelly + 1) / 2;
  ocplx exp_imtheta0 = cc(1.0, 0.0), exp_imdtheta = cc(1.0, 0.0);
  ocplx m_fac_1 = CZERO, m_fac_2 = CZERO, m_fac_3 = CZERO, m_fac_4 = CZERO;
  double inv_mdth = 0.0, inv_m2dth2 = 0.0;

  for (int im = 0; im < g->n_mode; ++im) {
    double mdth = (double)im * dtheta;
    double m2dth2 = mdth * mdth;
    int small_angle_change = fabs(mdth) < 1.0e-4;
    if (!small_angle_change && im > 0) {
      inv_mdth = 1.0 / mdth;
      inv_m2dth2 = inv_mdth * inv_mdth;
    }
    if (im == 0) {
      exp_imtheta0 = cc(1.0, 0.0);
      exp_imdtheta = cc(1.0, 0.0);
    } else {
      exp_imtheta0 = cmul(exp_imtheta0, PLANT(35, exp_itheta_05, cc(exp_itheta_05.re, -exp_itheta_05.im)));
      exp_imdtheta = cmul(exp_imdtheta, exp_idtheta);
      if (small_angle_change) {
        m_fac_1 = rmul(PLANT(34, 2.0, 1.0), exp_imtheta0);
        m_fac_2 = cmul(m_fac_1, cc(1.0 - sixth * m2dth2, 0.5 * mdth));
        m_fac_3 = cmul(m_fac_1, cc(0.5 - 0.125 * m2dth2, PLANT(36, third, 0.25) * mdth));
        m_fac_4 = cmul(m_fac_1, cc(third - 0.1 * m2dth2, 0.25 * mdth));
      } else {
        m_fac_1 = rmul(PLANT(34, 2.0, 1.0) * inv_mdth, exp_imtheta0);
        m_fac_2 = cmul(m_fac_1, cc(exp_imdtheta.im, -(exp_imdtheta.re - 1.0)));
        ocplx t = cmul(exp_imdtheta, cc(1.0, -mdth));
        m_fac_3 = cmul(m_fac_1, rmul(inv_mdth, cc(t.re - 1.0, t.im)));
        ocplx u = cmul(exp_imdtheta, cc(-m2dth2 + 2.0, -(2.0 * mdth)));
        u.re = u.re - 2.0;
        m_fac_4 = cmul(m_fac_1, imul(inv_m2dth2, u));
      }
    }

    ocplx jyh_[5] = {CZERO, CZERO, CZERO, CZERO, CZERO};
    ocplx *jyh = jyh_ + 2;
    for (int iy = ymin; iy <= ymax; ++iy) {
      int cy = cell_y2 + iy;
      ocplx w_rt, ym_fac_1;
      if (im == 0) {
        w_rt = cc(gy[iy] + 0.5 * hy[iy], 0.0);
        ym_fac_1 = cc(0.5 * gy[iy] + third * hy[iy], 0.0);
      } else {
        w_rt = cadd(cmulr(m_fac_2, gy[iy]), cmulr(m_fac_3, hy[iy]));
        ym_fac_1 = cadd(cmulr(m_fac_3, gy[iy]), cmulr(m_fac_4, hy[iy]));
      }
      double fjx = fcx * inv_area_rt[cy + jng];
      double fjy = (fcy * hy[iy]) * inv_area_xt[cy + jng];
      double fjz = fcz * inv_volume[cy + jng];
      ocplx jxh = CZERO;
      for (int ix = xmin; ix <= xmax; ++ix) {
        int cx = cell_x2 + ix;
        ocplx w_xt;
        if (im == 0) w_xt = cc(gx[ix] + 0.5 * hx[ix], 0.0);
        else w_xt = cadd(cmulr(m_fac_2, gx[ix]), cmulr(m_fac_3, hx[ix]));
        ocplx w_xr = cadd(rmul(gx[ix], w_rt), rmul(hx[ix], ym_fac_1));
        jxh = csub(jxh, rmul(fjx * hx[ix], w_rt));
        jyh[ix] = csub(cmulr(jyh[ix], ratio_area_xt[cy + jng]), rmul(PLANT(38, fjy, -fjy), w_xt));
        ocplx jzh = rmul(fjz, w_xr);
        size_t kx = I3(g, cx + 1, cy, im), kr = I3(g, cx, cy + 1, im), kt = I3(g, cx, cy, im);
        f->jx[kx] = cadd(f->jx[kx], jxh);
        f->jr[kr] = cadd(f->jr[kr], jyh[ix]);
        f->jt[kt] = cadd(f->jt[kt], jzh);
      }
    }
  }
}

/* Gather of one field component: sum over a 3x3 patch of mode m,
 * triangle/e_part.inc and b_part.inc. */
static inline ocplx patch(const orc_grid *g, const ocplx *F, int cx, int cy, int m,
                          const double *wx, const double *wy) {
  ocplx acc = CZERO;
  for (int jy = -1; jy <= 1; ++jy) {
    ocplx row = cadd(cadd(rmul(wx[-1], F[I3(g, cx - 1, cy + jy, m)]),
                          rmul(wx[0], F[I3(g, cx, cy + jy, m)])),
                     rmul(wx[1], F[I3(g, cx + 1, cy + jy, m)]));
    acc = (jy == -1) ? rmul(wy[jy], row) : cadd(acc, rmul(wy[jy], row));
  }
  return acc;
}
static inline double re_mul(ocplx a, ocplx b) { return a.re * b.re - a.im * b.im; }

/* push_particles, particles.F90:159-731 (probes, tracers, photons,
 * HC_PUSH, delta-f and work-done options are not part of the default build
 * and are not restated). */
/* The constants of one species' push (particles.F90:159-261) and the
 * geometry tables, shared by the particles of a push. */
typedef struct {
  double part_q, part_mc, ipart_mc, cmratio, ccmratio;
  double c, dtco2, idx, idy, idt, fac;
  const double *inv_area_rt, *inv_area_xt, *inv_volume, *ratio_area_xt;
} push_consts;

/* One particle's push and deposit (particles.F90:268-665): the gathers read
 * f on grid g; the current goes into fj on grid gj, whose column ix is the
 * global column ix + xoff (gj = g, fj = f, xoff = 0 for the plain push).
 * Returns 0, or the error bit of a skipped particle (1: a move of more than
 * one cell, 2: a stencil outside the arrays). */
static int push_one(const orc_grid *g, const orc_fields *f, const orc_grid *gj, orc_fields *fj, int xoff,
                    orc_species *s, int64_t ip, const push_consts *k) {
  const int M = g->n_mode;
  const double part_q = k->part_q, part_mc = k->part_mc, ipart_mc = k->ipart_mc;
  const double cmratio = k->cmratio, ccmratio = k->ccmratio;
  const double c = k->c, dtco2 = k->dtco2, idx = k->idx, idy = k->idy, idt = k->idt, fac = k->fac;
  double part_weight = s->w[ip];
  double part_x = s->x[ip], part_y = s->y[ip], part_z = s->z[ip];
  double part_ux = s->px[ip] * ipart_mc;
  double part_uy = s->py[ip] * ipart_mc;
  double part_uz = s->pz[ip] * ipart_mc;

  double gamma_rel = sqrt(((part_ux * part_ux + part_uy * part_uy) + part_uz * part_uz) + 1.0);
  double root = dtco2 / gamma_rel;
  part_x = part_x + part_ux * root;
  part_y = part_y + part_uy * root;
  part_z = part_z + part_uz * root;

  double part_x_local = part_x - g->x_grid_min;
  double part_r = sqrt(part_y * part_y + part_z * part_z);
  double part_r_local = part_r - g->y_grid_min;
  ocplx exp_min_itheta = cc(part_y / part_r, (-part_z) / part_r);
  double theta_05 = atan2(part_z, part_y);
  ocplx exp_min_itheta_05 = exp_min_itheta;
  ocplx exp_itheta_05 = crecip(exp_min_itheta_05);

  double cell_x_r = part_x_local * idx;
  double cell_y_r = part_r_local * idy;
  int cell_x1 = (int)floor(cell_x_r + 0.5);
  double cell_frac_x = (double)cell_x1 - cell_x_r;
  cell_x1 += 1;
  int cell_y1 = (int)floor(cell_y_r + 0.5);
  double cell_frac_y = (double)cell_y1 - cell_y_r;
  cell_y1 += 1;
  double gx_[5] = {0}, gy_[5] = {0}, hx_[5] = {0}, hy_[5] = {0};
  double *gx = gx_ + 2, *gy = gy_ + 2, *hx = hx_ + 2, *hy = hy_ + 2;
  weights3(cell_frac_x, gx);
  weights3(cell_frac_y, gy);

  int cell_x2 = (int)floor(cell_x_r);
  cell_frac_x = ((double)cell_x2 - cell_x_r) + 0.5;
  cell_x2 += 1;
  int cell_y2 = (int)floor(cell_y_r);
  cell_frac_y = ((double)cell_y2 - cell_y_r) + 0.5;
  cell_y2 += 1;
  weights3(cell_frac_x, hx);
  weights3(cell_frac_y, hy);
  if (!(fabs(cell_x_r) < 1e9 && fabs(cell_y_r) < 1e9) || cell_x1 < 2 - g->ng || cell_x1 > g->nx + g->ng - 1 ||
      cell_y1 < 2 - g->ng || cell_y1 > g->ny + g->ng - 1 || cell_x2 < 2 - g->ng || cell_x2 > g->nx + g->ng - 1 ||
      cell_y2 < 2 - g->ng || cell_y2 > g->ny + g->ng - 1) {
    /* a stencil outside the arrays (a position the reference's grid
     * cannot hold, or not finite): flag and skip */
    return 2;
  }

  /* e_part.inc / b_part.inc */
  ocplx e_m = cc(1.0, 0.0);
  double ex_part = 0.0, er_part = 0.0, et_part = 0.0;
  for (int im = 0; im < M; ++im) {
    ex_part = ex_part + re_mul(e_m, patch(g, f->ex, cell_x1, cell_y2, im, gx, hy));
    er_part = er_part + re_mul(e_m, patch(g, f->er, cell_x2, cell_y1, im, hx, gy));
    et_part = et_part + re_mul(e_m, patch(g, f->et, cell_x2, cell_y2, im, hx, hy));
    e_m = cmul(e_m, exp_min_itheta);
  }
  double ey_part = er_part * exp_min_itheta.re + PLANT(37, et_part, -et_part) * exp_min_itheta.im;
  double ez_part = -(er_part * exp_min_itheta.im) + et_part * exp_min_itheta.re;
  e_m = cc(1.0, 0.0);
  double bx_part = 0.0, br_part = 0.0, bt_part = 0.0;
  for (int im = 0; im < M; ++im) {
    bx_part = bx_part + re_mul(e_m, patch(g, f->bx, cell_x2, cell_y1, im, hx, gy));
    br_part = br_part + re_mul(e_m, patch(g, f->br, cell_x1, cell_y2, im, gx, hy));
    bt_part = bt_part + re_mul(e_m, patch(g, f->bt, cell_x1, cell_y1, im, gx, gy));
    e_m = cmul(e_m, exp_min_itheta);
  }
  double by_part = br_part * exp_min_itheta.re + bt_part * exp_min_itheta.im;
  double bz_part = -(br_part * exp_min_itheta.im) + bt_part * exp_min_itheta.re;

  /* Boris, particles.F90:404-451 */
  double uxm = part_ux + cmratio * ex_part;
  double uym = part_uy + cmratio * ey_part;
  double uzm = part_uz + cmratio * ez_part;
  gamma_rel = PLANT(31, sqrt(((uxm * uxm + uym * uym) + uzm * uzm) + 1.0),
                    sqrt(((part_ux * part_ux + part_uy * part_uy) + part_uz * part_uz) + 1.0));
  root = ccmratio / gamma_rel;
  double taux = bx_part * root, tauy = by_part * root, tauz = bz_part * root;
  double taux2 = taux * taux, tauy2 = tauy * tauy, tauz2 = tauz * tauz;
  double tau = 1.0 / (((1.0 + taux2) + tauy2) + tauz2);
  double uxp = ((((1.0 + taux2) - tauy2) - tauz2) * uxm +
                2.0 * ((taux * tauy + PLANT(33, tauz, -tauz)) * uym + (taux * tauz - tauy) * uzm)) * tau;
  double uyp = ((((1.0 - taux2) + tauy2) - tauz2) * uym +
                2.0 * ((tauy * tauz + taux) * uzm + (tauy * taux - tauz) * uxm)) * tau;
  double uzp = ((((1.0 - taux2) - tauy2) + tauz2) * uzm +
                2.0 * ((tauz * taux + tauy) * uxm + (tauz * tauy - taux) * uym)) * tau;
  part_ux = uxp + PLANT(32, cmratio, -cmratio) * ex_part;
  part_uy = uyp + cmratio * ey_part;
  part_uz = uzp + cmratio * ez_part;

  double part_u2 = (part_ux * part_ux + part_uy * part_uy) + part_uz * part_uz;
  gamma_rel = sqrt(part_u2 + 1.0);
  double igamma = 1.0 / gamma_rel;
  root = dtco2 * igamma;
  double delta_x = part_ux * root, delta_y = part_uy * root, delta_z = part_uz * root;
  part_x = part_x + delta_x;
  part_y = part_y + delta_y;
  part_z = part_z + delta_z;
  s->x[ip] = part_x; s->y[ip] = part_y; s->z[ip] = part_z;
  s->px[ip] = part_mc * part_ux; s->py[ip] = part_mc * part_uy; s->pz[ip] = part_mc * part_uz;

  double part_vy = (part_uy * c) * igamma;
  double part_vz = (part_uz * c) * igamma;
  part_r = sqrt(part_y * part_y + part_z * part_z);
  ocplx exp_itheta_10 = cc(part_y / part_r, part_z / part_r);
  double part_vt = -(part_vy * exp_itheta_10.im) + part_vz * exp_itheta_10.re;

  /* current at t+1.5dt, particles.F90:504-584 */
  part_x_local = (part_x + delta_x) - g->x_grid_min;
  part_y = part_y + delta_y;
  part_z = part_z + delta_z;
  part_r = sqrt(part_y * part_y + part_z * part_z);
  part_r_local = part_r - g->y_grid_min;
  ocplx exp_itheta_15 = cc(part_y / part_r, part_z / part_r);
  double theta_15 = atan2(part_z, part_y);
  ocplx exp_idtheta = cmul(exp_itheta_15, exp_min_itheta_05);
  double dtheta = theta_15 - theta_05;

  for (int k = -2; k <= 2; ++k) { gx[k] = hx[k]; gy[k] = hy[k]; }
  cell_x_r = part_x_local * idx;
  cell_y_r = part_r_local * idy;
  int cell_x3 = (int)floor(cell_x_r);
  cell_frac_x = ((double)cell_x3 - cell_x_r) + 0.5;
  cell_x3 += 1;
  int cell_y3 = (int)floor(cell_y_r);
  cell_frac_y = ((double)cell_y3 - cell_y_r) + 0.5;
  cell_y3 += 1;
  for (int k = -2; k <= 2; ++k) { hx[k] = 0.0; hy[k] = 0.0; }
  int dcellx = cell_x3 - cell_x2, dcelly = cell_y3 - cell_y2;
  if (dcellx < -1 || dcellx > 1 || dcelly < -1 || dcelly > 1) {
    /* a move of more than one cell per step (or a non-finite position):
     * the reference's weight arrays would be overrun; flag and skip */
    return 1;
  }
  weights3(cell_frac_x, hx + dcellx);
  weights3(cell_frac_y, hy + dcelly);
  for (int k = -2; k <= 2; ++k) { hx[k] = hx[k] - gx[k]; hy[k] = hy[k] - gy[k]; }

  double q_weight_fac = (part_q * part_weight) * fac;
  double fcx = q_weight_fac * idt, fcy = q_weight_fac * idt, fcz = q_weight_fac * part_vt;
  orc_deposit_from_weights(gj, fj, gx_, hx_, gy_, hy_, cell_x2 - xoff, cell_y2, dcellx, dcelly,
                           fcx, fcy, fcz, exp_itheta_05, exp_idtheta, dtheta,
                           k->inv_area_rt, k->inv_area_xt, k->inv_volume, k->ratio_area_xt);
  return 0;
}

static push_consts species_consts(const orc_grid *g, const orc_species *s, const double *tab, int ntab) {
  push_consts k;
  const double c = C_LIGHT, dt = g->dt;
  const double dto2 = dt / 2.0;
  const double fac = 0.25;                 /* (0.5)**c_ndims */
  const double dtfac = (0.5 * dt) * fac;
  k.c = c;
  k.idx = 1.0 / g->dx;
  k.idy = 1.0 / g->dy;
  k.idt = 1.0 / dt;
  k.dtco2 = c * dto2;
  k.fac = fac;
  k.part_q = s->q;
  k.part_mc = c * s->m;
  k.ipart_mc = 1.0 / k.part_mc;
  k.cmratio = (k.part_q * dtfac) * k.ipart_mc;
  k.ccmratio = c * k.cmratio;
  k.inv_area_rt = tab;
  k.inv_area_xt = tab + ntab;
  k.inv_volume = tab + 2 * ntab;
  k.ratio_area_xt = tab + 3 * ntab;
  return k;
}

void orc_push_particles(const orc_grid *g, orc_fields *f, orc_species *sp, int n_species) {
  const int ny = g->ny, jng = g->ng;
  size_t nbytes = (size_t)NX2(g) * NY2(g) * g->n_mode * sizeof(ocplx);
  memcpy(f->jx_old, f->jx, nbytes);
  memcpy(f->jr_old, f->jr, nbytes);
  memcpy(f->jt_old, f->jt, nbytes);
  memset(f->jx, 0, nbytes);
  memset(f->jr, 0, nbytes);
  memset(f->jt, 0, nbytes);

  int ntab = ny + 2 * jng + 1;
  double *tab = calloc(4 * (size_t)ntab, sizeof(double));
  orc_geometry(g, tab, tab + ntab, tab + 2 * ntab, tab + 3 * ntab);

  for (int is = 0; is < n_species; ++is) {
    orc_species *s = &sp[is];
    if (s->immobile) continue;
    const push_consts k = species_consts(g, s, tab, ntab);
    for (int64_t ip = 0; ip < s->n; ++ip) orc_error |= push_one(g, f, g, f, 0, s, ip, &k);
    /* current_bcs(species=ispecies): a no-op unless particle bcs are mixed */
  }
  free(tab);

  orc_current_bcs_r_min_final(g, f);
  orc_particle_bcs(g, sp, n_species);
}

/* The CPU baseline's push on every host core (timing only; the tests use
 * orc_push_particles).  Each species' particles are cut by the x column of
 * their position into nthreads slabs of columns, like the reference's MPI
 * ranks; each thread pushes its slab's particles (stable order) and deposits
 * into its own zeroed current arrays over the slab's columns and ng ghost
 * columns each side; the copies are then added into J in slab order for
 * every element, so the result does not depend on scheduling.  It differs
 * from orc_push_particles only in the order of the current sums. */
void orc_push_particles_threads(const orc_grid *g, orc_fields *f, orc_species *sp, int n_species, int nthreads) {
  const int nx = g->nx, ny = g->ny, M = g->n_mode, jng = g->ng;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > nx) nthreads = nx;
  size_t nbytes = (size_t)NX2(g) * NY2(g) * M * sizeof(ocplx);
  memcpy(f->jx_old, f->jx, nbytes);
  memcpy(f->jr_old, f->jr, nbytes);
  memcpy(f->jt_old, f->jt, nbytes);
  memset(f->jx, 0, nbytes);
  memset(f->jr, 0, nbytes);
  memset(f->jt, 0, nbytes);

  int ntab = ny + 2 * jng + 1;
  double *tab = calloc(4 * (size_t)ntab, sizeof(double));
  orc_geometry(g, tab, tab + ntab, tab + 2 * ntab, tab + 3 * ntab);
  /* slab t: columns c0[t] .. c0[t+1]-1 */
  int *c0 = malloc(sizeof(int) * (size_t)(nthreads + 1));
  for (int t = 0; t <= nthreads; ++t) c0[t] = 1 + (int)(((int64_t)nx * t) / nthreads);
  orc_grid *gs = malloc(sizeof(orc_grid) * (size_t)nthreads);
  orc_fields *fs = calloc((size_t)nthreads, sizeof(orc_fields));
  for (int t = 0; t < nthreads; ++t) {
    gs[t] = *g;
    gs[t].nx = c0[t + 1] - c0[t];
    size_t n = (size_t)NX2(&gs[t]) * NY2(g) * M;
    fs[t].jx = calloc(n, sizeof(ocplx));
    fs[t].jr = calloc(n, sizeof(ocplx));
    fs[t].jt = calloc(n, sizeof(ocplx));
  }
  int err = 0;
  for (int is = 0; is < n_species; ++is) {
    orc_species *s = &sp[is];
    if (s->immobile || s->n <= 0) continue;
    const push_consts k = species_consts(g, s, tab, ntab);
    /* the particles of each slab, in their order */
    int *slab = malloc(sizeof(int) * (size_t)s->n);
    int64_t *start = calloc((size_t)nthreads + 1, sizeof(int64_t));
    for (int64_t ip = 0; ip < s->n; ++ip) {
      double cx = floor((s->x[ip] - g->x_grid_min) / g->dx + 0.5) + 1.0;
      int c = cx < 1.0 ? 1 : (cx > (double)nx ? nx : (int)cx);
      int t = (int)(((int64_t)(c - 1) * nthreads) / nx);
      while (t + 1 < nthreads && c >= c0[t + 1]) ++t;
      while (t > 0 && c < c0[t]) --t;
      slab[ip] = t;
      start[t + 1]++;
    }
    for (int t = 0; t < nthreads; ++t) start[t + 1] += start[t];
    int64_t *order = malloc(sizeof(int64_t) * (size_t)s->n);
    int64_t *fill = malloc(sizeof(int64_t) * (size_t)nthreads);
    for (int t = 0; t < nthreads; ++t) fill[t] = start[t];
    for (int64_t ip = 0; ip < s->n; ++ip) order[fill[slab[ip]]++] = ip;
#pragma omp parallel for num_threads(nthreads) schedule(static, 1) reduction(| : err)
    for (int t = 0; t < nthreads; ++t)
      for (int64_t q = start[t]; q < start[t + 1]; ++q)
        err |= push_one(g, f, &gs[t], &fs[t], c0[t] - 1, s, order[q], &k);
    free(slab); free(start); free(order); free(fill);
  }
  /* J += the slabs' currents, slab by slab for each element */
#pragma omp parallel for num_threads(nthreads) collapse(2) schedule(static)
  for (int m = 0; m < M; ++m)
    for (int iy = 1 - jng; iy <= ny + jng; ++iy)
      for (int t = 0; t < nthreads; ++t)
        for (int lx = 1 - jng; lx <= gs[t].nx + jng; ++lx) {
          const int ix = lx + c0[t] - 1;
          if (ix < 1 - jng || ix > nx + jng) continue;
          size_t a = I3(g, ix, iy, m), b = I3(&gs[t], lx, iy, m);
          f->jx[a] = cadd(f->jx[a], fs[t].jx[b]);
          f->jr[a] = cadd(f->jr[a], fs[t].jr[b]);
          f->jt[a] = cadd(f->jt[a], fs[t].jt[b]);
        }
  for (int t = 0; t < nthreads; ++t) {
    free(fs[t].jx); free(fs[t].jr); free(fs[t].jt);
  }
  free(fs); free(gs); free(c0); free(tab);
  orc_error |= err;

  orc_current_bcs_r_min_final(g, f);
  orc_particle_bcs(g, sp, n_species);
}

/* particle_bcs on one rank, boundary.F90:1541-1889 (periodic, reflect,
 * open; thermal boundaries are not part of this path).  Particles that wrap
 * through a periodic boundary are re-appended at the end of their list,
 * first those that left through x_min then through x_max, which is the
 * order the reference's self-exchange produces. */
void orc_particle_bcs(const orc_grid *g, orc_species *sp, int n_species) {
  const double dx = g->dx, dy = g->dy;
  const int png = 3;
  const double bshift_x = dx * (double)((1 + png + 0) / 2);
  const double x_min_outer = g->x_min - bshift_x, x_max_outer = g->x_max + bshift_x;
  const double x_shift = g->x_max - g->x_min;
  const double bshift_y = dy * (double)((1 + png + 0) / 2);
  const double y_max_outer = g->y_max + bshift_y;

  for (int is = 0; is < n_species; ++is) {
    orc_species *s = &sp[is];
    int64_t n = s->n;
    int8_t *state = malloc((size_t)(n > 0 ? n : 1));  /* 0 keep, 1 remove, 2 left -x, 3 left +x */
    for (int64_t ip = 0; ip < n; ++ip) {
      int xbd = 0, out_of_bounds = 0, bc = 0;
      double part_pos = s->x[ip];
      if (part_pos < g->x_min) {
        xbd = -1;
        xbd = 0;
        bc = g->bc_particle[ORC_XMIN];
        if (bc == ORC_BC_REFLECT) {
          s->x[ip] = 2.0 * g->x_min - part_pos;
          s->px[ip] = -s->px[ip];
        } else if (bc == ORC_BC_PERIODIC) {
          xbd = -1;
          s->x[ip] = part_pos - (-1.0) * x_shift;
        }
        if (part_pos < x_min_outer && bc != ORC_BC_PERIODIC) out_of_bounds = 1;
      }
      if (part_pos >= g->x_max) {
        xbd = 1;
        xbd = 0;
        bc = g->bc_particle[ORC_XMAX];
        if (bc == ORC_BC_REFLECT) {
          s->x[ip] = 2.0 * g->x_max - part_pos;
          s->px[ip] = -s->px[ip];
        } else if (bc == ORC_BC_PERIODIC) {
          xbd = 1;
          s->x[ip] = part_pos - 1.0 * x_shift;
        }
        if (part_pos >= x_max_outer && bc != ORC_BC_PERIODIC) out_of_bounds = 1;
      }
      part_pos = sqrt(s->y[ip] * s->y[ip] + s->z[ip] * s->z[ip]);
      if (part_pos >= g->y_max) {
        bc = g->bc_particle[ORC_YMAX];
        if (bc == ORC_BC_REFLECT) {
          double radial_reduction = 2.0 * g->y_max / part_pos - 1.0;
          s->y[ip] = s->y[ip] * radial_reduction;
          s->z[ip] = s->z[ip] * radial_reduction;
          double inv_final_r = 1.0 / sqrt(s->y[ip] * s->y[ip] + s->z[ip] * s->z[ip]);
          double cos_theta = s->y[ip] * inv_final_r;
          double sin_theta = s->z[ip] * inv_final_r;
          double part_pr = s->py[ip] * cos_theta + s->pz[ip] * sin_theta;
          double part_pt = -(s->py[ip] * sin_theta) + s->pz[ip] * cos_theta;
          if (ORACLE_PLANT == 39) part_pr = -part_pr;
          s->py[ip] = -(part_pr * cos_theta) - part_pt * sin_theta;
          s->pz[ip] = -(part_pr * sin_theta) + part_pt * cos_theta;
        }
        if (part_pos >= y_max_outer && bc != ORC_BC_PERIODIC) out_of_bounds = 1;
      }
      state[ip] = out_of_bounds ? 1 : (xbd < 0 ? 2 : (xbd > 0 ? 3 : 0));
    }
    /* compact, then append the two self-exchanged groups */
    double *arrs[7] = {s->x, s->y, s->z, s->px, s->py, s->pz, s->w};
    int64_t *tmpi = malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t cnt = 0;
    for (int pass = 0; pass < 3; ++pass) {
      int want = pass == 0 ? 0 : (pass == 1 ? 2 : 3);
      for (int64_t ip = 0; ip < n; ++ip)
        if (state[ip] == want) tmpi[cnt++] = ip;
    }
    double *tmp = malloc(sizeof(double) * (size_t)(cnt > 0 ? cnt : 1));
    for (int a = 0; a < 7; ++a) {
      for (int64_t k = 0; k < cnt; ++k) tmp[k] = arrs[a][tmpi[k]];
      memcpy(arrs[a], tmp, sizeof(double) * (size_t)cnt);
    }
    if (s->id) {
      int64_t *tid = malloc(sizeof(int64_t) * (size_t)(cnt > 0 ? cnt : 1));
      for (int64_t k = 0; k < cnt; ++k) tid[k] = s->id[tmpi[k]];
      memcpy(s->id, tid, sizeof(int64_t) * (size_t)cnt);
      free(tid);
    }
    s->n = cnt;
    free(tmp); free(tmpi); free(state);
  }
}

/* ------------------------------------------------------------------ */
/* cell key (split_particle.F90:61-62) and stable counting sort         */
/* ------------------------------------------------------------------ */
void orc_cell_keys(const orc_grid *g, int64_t n, const double *x, const double *y,
                   const double *z, int tile_x, int tile_r, int shift, int32_t *keys) {
  const int nx = g->nx, ny = g->ny;
  const int ntx = (nx + shift + tile_x - 1) / tile_x;
  for (int64_t i = 0; i < n; ++i) {
    double r = sqrt(y[i] * y[i] + z[i] * z[i]);
    int cx = (int)floor((x[i] - g->x_grid_min) / g->dx + PLANT(40, 1.5, 1.0));
    int cy = (int)floor((r - g->y_grid_min) / g->dy + 1.5);
    /* a particle on the domain edge within rounding maps to the edge cell */
    cx = cx < 1 ? 1 : (cx > nx ? nx : cx);
    cy = cy < 1 ? 1 : (cy > ny ? ny : cy);
    int lx = cx - 1 + shift, ly = cy - 1;
    int tx = lx / tile_x, ty = ly / tile_r;
    keys[i] = (int32_t)((((int64_t)ty * ntx + tx) * tile_r + (ly - ty * tile_r)) * tile_x +
                        (lx - tx * tile_x));
  }
}

void orc_stable_sort_keys(int64_t n, const int32_t *keys, int32_t nkeys, int64_t *perm) {
  int64_t *start = calloc((size_t)nkeys + 1, sizeof(int64_t));
  for (int64_t i = 0; i < n; ++i) start[keys[i] + 1]++;
  for (int32_t k = 0; k < nkeys; ++k) start[k + 1] += start[k];
  for (int64_t i = 0; i < n; ++i) perm[start[keys[i]]++] = i;
  free(start);
}

/* ------------------------------------------------------------------ */
/* set_dt, setup.F90:629-673                                          */
/* ------------------------------------------------------------------ */
double orc_set_dt(double dx, double dy, double dt_multiplier,
                  double dt_plasma_frequency, double dt_laser) {
  double dt = ((0.9 * (dx < dy ? dx : dy)) / sqrt(2.0)) / C_LIGHT;
  if (dt_plasma_frequency > C_TINY && dt_plasma_frequency < dt) dt = dt_plasma_frequency;
  if (dt_laser > C_TINY && dt_laser < dt) dt = dt_laser;
  return dt_multiplier * dt;
}

/* ------------------------------------------------------------------ */
/* mode energies: W_m = kappa_m * sum (eps0/2 |E_m|^2 + |B_m|^2/(2 mu0))  */
/* * r dr dx over interior cells, each component at its own staggered   */
/* radius; kappa_0 = 2 pi, kappa_m>0 = pi (the theta integral of         */
/* Re[F_m exp(-i m theta)]^2).                                           */
/* ------------------------------------------------------------------ */
void orc_mode_energy(const orc_grid *g, const orc_fields *f, double *w_e, double *w_b) {
  const double mu0 = 4.0e-7 * PI;
  for (int m = 0; m < g->n_mode; ++m) {
    double kappa = (m == 0) ? PLANT(46, 2.0 * PI, PI) : PI;
    double se = 0.0, sb = 0.0;
    for (int ir = 1; ir <= g->ny; ++ir) {
      double r_c = ((double)ir - 0.5) * g->dy;   /* er, bx, bt */
      double r_e = (double)ir * g->dy;           /* ex, et, br */
      for (int ix = 1; ix <= g->nx; ++ix) {
        size_t k = I3(g, ix, ir, m);
        ocplx a = f->ex[k], b = f->er[k], c = f->et[k];
        se += (a.re * a.re + a.im * a.im) * r_e + (b.re * b.re + b.im * b.im) * r_c +
              (c.re * c.re + c.im * c.im) * r_e;
        a = f->bx[k]; b = f->br[k]; c = f->bt[k];
        sb += (a.re * a.re + a.im * a.im) * r_c + (b.re * b.re + b.im * b.im) * r_e +
              (c.re * c.re + c.im * c.im) * r_c;
      }
    }
    w_e[m] = kappa * 0.5 * EPSILON0 * se * g->dx * g->dy;
    w_b[m] = kappa * 0.5 / mu0 * sb * g->dx * g->dy;
  }
}

/* ------------------------------------------------------------------ */
/* moving window, window.F90:62-153, 330-350                           */
/* ------------------------------------------------------------------ */
/* shift_field_modes, window.F90:132-150: f(i) = f(i+1) for i in
 * [1-ng, nx+ng-1]; f(nx+ng) keeps its value.  field_mode_bc is an MPI halo
 * exchange, a no-op on one rank with non-periodic x. */
static void shift_field_modes(const orc_grid *g, ocplx *f) {
  const int nx = g->nx, ny = g->ny, ng = g->ng;
  for (int m = 0; m < g->n_mode; ++m)
    for (int j = 1 - ng; j <= ny + ng; ++j)
      for (int i = 1 - ng; i <= nx + ng - 1; ++i) f[I3(g, i, j, m)] = f[I3(g, i + 1, j, m)];
}

/* shift_fields, window.F90:101-128 (x_max is a domain boundary on one rank) */
static void shift_fields(const orc_grid *g, orc_fields *f) {
  ocplx *arr[9] = {f->ex, f->er, f->et, f->bx, f->br, f->bt, f->jx, f->jr, f->jt};
  for (int k = 0; k < 9; ++k) shift_field_modes(g, arr[k]);
  const int nx = g->nx, ny = g->ny, ng = g->ng;
  for (int m = 0; m < g->n_mode; ++m)
    for (int j = 1 - ng; j <= ny + ng; ++j) {
      const size_t s = I2(g, j, m);
      f->ex[I3(g, nx + 1, j, m)] = f->bnd_xmax[0][s];
      f->er[I3(g, nx, j, m)] = f->bnd_xmax[1][s];
      f->et[I3(g, nx, j, m)] = f->bnd_xmax[2][s];
      f->ex[I3(g, nx, j, m)] = PLANT(41, rmul(0.5, cadd(f->ex[I3(g, nx - 1, j, m)], f->ex[I3(g, nx + 1, j, m)])),
                                     f->ex[I3(g, nx - 1, j, m)]);
      f->er[I3(g, nx - 1, j, m)] = rmul(0.5, cadd(f->er[I3(g, nx - 2, j, m)], f->er[I3(g, nx, j, m)]));
      f->et[I3(g, nx - 1, j, m)] = rmul(0.5, cadd(f->et[I3(g, nx - 2, j, m)], f->et[I3(g, nx, j, m)]));
      f->bx[I3(g, nx, j, m)] = f->bnd_xmax[3][s];
      f->br[I3(g, nx + 1, j, m)] = f->bnd_xmax[4][s];
      f->bt[I3(g, nx + 1, j, m)] = f->bnd_xmax[5][s];
      f->bx[I3(g, nx - 1, j, m)] = rmul(0.5, cadd(f->bx[I3(g, nx - 2, j, m)], f->bx[I3(g, nx, j, m)]));
      f->br[I3(g, nx, j, m)] = rmul(0.5, cadd(f->br[I3(g, nx - 1, j, m)], f->br[I3(g, nx + 1, j, m)]));
      f->bt[I3(g, nx, j, m)] = rmul(0.5, cadd(f->bt[I3(g, nx - 1, j, m)], f->bt[I3(g, nx + 1, j, m)]));
    }
}

/* remove_particles, window.F90:330-350: drop x < x_min, keeping the order */
static void remove_left(const orc_grid *g, orc_species *sp, int n_species) {
  for (int s = 0; s < n_species; ++s) {
    orc_species *p = &sp[s];
    int64_t k = 0;
    for (int64_t i = 0; i < p->n; ++i) {
      if (p->x[i] < g->x_min) continue;
      p->x[k] = p->x[i];
      p->y[k] = p->y[i];
      p->z[k] = p->z[i];
      p->px[k] = p->px[i];
      p->py[k] = p->py[i];
      p->pz[k] = p->pz[i];
      p->w[k] = p->w[i];
      if (p->id) p->id[k] = p->id[i];
      ++k;
    }
    p->n = k;
  }
}

/* shift_window, window.F90:62-97 (no CPML: x_min = xb_min) */
void orc_shift_window(orc_grid *g, orc_fields *f, orc_species *sp, int n_species, int ncells) {
  for (int c = 0; c < ncells; ++c) {
    /* x_global(1) + dx, xb_global(1) + dx, xb_global(nx + 1) */
    g->x_grid_min = g->x_grid_min + g->dx;
    g->x_min = g->x_min + g->dx;
    g->x_max = g->x_min + (double)g->nx * g->dx;
    remove_left(g, sp, n_species);
    shift_fields(g, f);
  }
}

/* ------------------------------------------------------------------ */
/* output diagnostics: src/io/calc_df.F90, src/io/dist_fn.F90         */
/* ------------------------------------------------------------------ */

/* calc_boundary_modes without mixed particle bcs (calc_df.F90:35-55): the
 * real processor_summation_bcs (boundary.F90:1305-1326) on the real and
 * imaginary parts, i.e. particle_reflection_bcs without a flip direction
 * (:833-914, cell-centred folds) then the one-rank particle_periodic_bcs
 * (:1019-1129). */
static void density_summation_bcs(const orc_grid *g, ocplx *a) {
  const int nx = g->nx, ny = g->ny, ng = g->ng;
  for (int m = 0; m < g->n_mode; ++m) {
    if (g->bc_particle[ORC_XMIN] == ORC_BC_REFLECT)
      for (int j = 1 - ng; j <= ny + ng; ++j)
        for (int i = 1; i <= ng - 1; ++i) {
          a[I3(g, i, j, m)] = cadd(a[I3(g, i, j, m)], a[I3(g, 1 - i, j, m)]);
          a[I3(g, 1 - i, j, m)] = CZERO;
        }
    if (g->bc_particle[ORC_XMAX] == ORC_BC_REFLECT)
      for (int j = 1 - ng; j <= ny + ng; ++j)
        for (int i = 1; i <= ng; ++i) {
          a[I3(g, nx + 1 - i, j, m)] = cadd(a[I3(g, nx + 1 - i, j, m)], a[I3(g, nx + i, j, m)]);
          a[I3(g, nx + i, j, m)] = CZERO;
        }
    if (g->bc_particle[ORC_YMAX] == ORC_BC_REFLECT)
      for (int i = 1; i <= ng; ++i)
        for (int ix = 1 - ng; ix <= nx + ng; ++ix) {
          a[I3(g, ix, ny + 1 - i, m)] = cadd(a[I3(g, ix, ny + 1 - i, m)], a[I3(g, ix, ny + i, m)]);
          a[I3(g, ix, ny + i, m)] = CZERO;
        }
    if (g->bc_particle[ORC_XMIN] == ORC_BC_PERIODIC && g->bc_particle[ORC_XMAX] == ORC_BC_PERIODIC) {
      for (int j = 1 - ng; j <= ny + ng; ++j)
        for (int i = 1; i <= ng; ++i) a[I3(g, i, j, m)] = cadd(a[I3(g, i, j, m)], a[I3(g, nx + i, j, m)]);
      for (int j = 1 - ng; j <= ny + ng; ++j)
        for (int i = 1; i <= ng; ++i)
          a[I3(g, nx - ng + i, j, m)] = cadd(a[I3(g, nx - ng + i, j, m)], a[I3(g, i - ng, j, m)]);
    }
  }
}

/* field_mode_zero_gradient for a cell-centred array (boundary.F90:654-707)
 * on x_min, x_max, y_min, y_max in that order, skipping periodic ones. */
static void centre_zero_gradient(const orc_grid *g, ocplx *a) {
  const int nx = g->nx, ny = g->ny, ng = g->ng;
  for (int m = 0; m < g->n_mode; ++m) {
    if (g->bc_field[ORC_XMIN] != ORC_BC_PERIODIC)
      for (int j = 1 - ng; j <= ny + ng; ++j)
        for (int i = 1; i <= ng; ++i) a[I3(g, i - ng, j, m)] = a[I3(g, ng + 1 - i, j, m)];
    if (g->bc_field[ORC_XMAX] != ORC_BC_PERIODIC)
      for (int j = 1 - ng; j <= ny + ng; ++j)
        for (int i = 1; i <= ng; ++i) a[I3(g, nx + i, j, m)] = a[I3(g, nx + 1 - i, j, m)];
    if (g->bc_field[ORC_YMIN] != ORC_BC_PERIODIC)
      for (int i = 1; i <= ng; ++i)
        for (int ix = 1 - ng; ix <= nx + ng; ++ix) a[I3(g, ix, i - ng, m)] = a[I3(g, ix, ng + 1 - i, m)];
    if (g->bc_field[ORC_YMAX] != ORC_BC_PERIODIC)
      for (int i = 1; i <= ng; ++i)
        for (int ix = 1 - ng; ix <= nx + ng; ++ix) a[I3(g, ix, ny + i, m)] = a[I3(g, ix, ny + 1 - i, m)];
  }
}

/* calc_number_density_modes, calc_df.F90:588-661, with particle_to_grid.inc
 * and triangle/gxfac.inc (weights summing to one, the r < dy fold) and
 * get_macro_part_volume (partlist.F90:999-1013): each particle adds
 * gx gy w / (2 pi dx dy r) times 1 (m = 0) or 2 e^{i m theta} (m > 0). */
void orc_number_density_modes(const orc_grid *g, const orc_species *sp, int n_species,
                              int species, ocplx *out) {
  const size_t n = (size_t)g->n_mode * NX2(g) * NY2(g);
  memset(out, 0, n * sizeof(ocplx));
  const int s0 = species < 0 ? 0 : species, s1 = species < 0 ? n_species - 1 : species;
  for (int is = s0; is <= s1; ++is) {
    const orc_species *s = &sp[is];
    for (int64_t ip = 0; ip < s->n; ++ip) {
      const double y = s->y[ip], z = s->z[ip];
      const double part_r = sqrt(y * y + z * z);
      const double cxr = (s->x[ip] - g->x_grid_min) / g->dx;
      const double cyr = (part_r - g->y_grid_min) / g->dy;
      int cx = (int)floor(cxr + 0.5), cy = (int)floor(cyr + 0.5);
      const double fx = (double)cx - cxr, fy = (double)cy - cyr;
      cx += 1;
      cy += 1;
      double gx[3], gy[3];
      const double cx2 = fx * fx, cy2 = fy * fy;
      gx[0] = 0.5 * ((0.25 + cx2) + fx);
      gx[1] = 0.75 - cx2;
      gx[2] = 0.5 * ((0.25 + cx2) - fx);
      gy[0] = 0.5 * ((0.25 + cy2) + fy);
      gy[1] = 0.75 - cy2;
      gy[2] = 0.5 * ((0.25 + cy2) - fy);
      if (part_r < PLANT(44, g->dy, 0.5 * g->dy)) {
        gy[1] = gy[1] + gy[0];
        gy[0] = 0.0;
      }
      if (cx - 1 < 1 - g->ng || cx + 1 > g->nx + g->ng || cy - 1 < 1 - g->ng || cy + 1 > g->ny + g->ng)
        continue; /* outside the array: no valid particle gets here */
      const double vol = (((2.0 * PI) * g->dx) * g->dy) * part_r;
      const double pnd = s->w[ip] / vol;
      const ocplx eith = cc(y / part_r, z / part_r);
      ocplx eim = cc(1.0, 0.0), mf = cc(1.0, 0.0);
      for (int m = 0; m < g->n_mode; ++m) {
        if (m == 0) {
          mf = cc(1.0, 0.0);
        } else {
          eim = cmul(eim, eith);
          mf = rmul(2.0, eim);
        }
        for (int iy = -1; iy <= 1; ++iy)
          for (int ix = -1; ix <= 1; ++ix) {
            const size_t k = I3(g, cx + ix, cy + iy, m);
            out[k] = cadd(out[k], rmul((gx[ix + 1] * gy[iy + 1]) * pnd, mf));
          }
      }
    }
  }
  density_summation_bcs(g, out);
  centre_zero_gradient(g, out);
}

/* the particle_data(:) of general_dist_fn for a non-photon particle
 * (dist_fn.F90:486-553); angles only where `need` asks; returns 0 when an
 * angle is needed and its momentum pair has |p| < tiny (CYCLE). */
static int df_particle_data(const orc_species *s, int64_t ip, const int *need, const double *amax,
                            double *pd) {
  const double mc = s->m * C_LIGHT, mc2 = mc * C_LIGHT;
  const double px = s->px[ip], py = s->py[ip], pz = s->pz[ip];
  const double ux = px / mc, uy = py / mc, uz = pz / mc;
  const double u2 = (ux * ux + uy * uy) + uz * uz;
  const double gam = sqrt(u2 + 1.0);
  const double gm1 = u2 / (gam + 1.0);
  pd[ORC_DIR_X] = s->x[ip];
  pd[ORC_DIR_Y] = s->y[ip];
  pd[ORC_DIR_PX] = px;
  pd[ORC_DIR_PY] = py;
  pd[ORC_DIR_PZ] = pz;
  pd[ORC_DIR_EN] = gm1 * mc2;
  pd[ORC_DIR_GAMMA_M1] = gm1;
  pd[ORC_DIR_MOD_P] = sqrt((px * px + py * py) + pz * pz);
  const double a1[3] = {px, py, pz}, a2[3] = {py, pz, px};
  for (int k = 0; k < 3; ++k) {
    pd[ORC_DIR_XY_ANGLE + k] = 0.0;
    if (!need[k]) continue;
    const double p = sqrt(a1[k] * a1[k] + a2[k] * a2[k]);
    if (p < C_TINY) return 0;
    double th = asin(a2[k] / p);
    if (a1[k] < 0.0) th = copysign(1.0, a2[k]) * PI - th;
    if (amax && th + 2.0 * PI < amax[k]) th = th + 2.0 * PI;
    pd[ORC_DIR_XY_ANGLE + k] = th;
  }
  return 1;
}

static int df_restricted(const double *pd, const double *restr, const int *use) {
  for (int d = 1; d <= ORC_DF_MAXDIRS; ++d)
    if (use[d - 1] && (pd[d] < restr[2 * (d - 1)] || pd[d] > restr[2 * (d - 1) + 1])) return 1;
  return 0;
}

int orc_dist_fn(const orc_grid *g, const orc_species *s, int ndims, const int *direction,
                double *ranges, int *resolution, const double *restrictions,
                const int *use_restrictions, double *out) {
  if (ndims < 1 || ndims > ORC_DF_MAXDIMS) return 1;
  if (s->n < 1) return 2;
  int calc[3] = {0, 0, 0}, any_calc = 0, parallel[3] = {0, 0, 0};
  int need[3] = {0, 0, 0};
  double dgrid[3] = {1.0, 1.0, 1.0};
  for (int d = 0; d < ndims; ++d) {
    const int dir = direction[d];
    double *r = ranges + 2 * d;
    if (dir == ORC_DIR_X || dir == ORC_DIR_Y) {
      const double h = dir == ORC_DIR_X ? g->dx : g->dy;
      const double lo = dir == ORC_DIR_X ? g->x_min : g->y_min;
      const double hi = dir == ORC_DIR_X ? g->x_max : g->y_max;
      const double gmin = dir == ORC_DIR_X ? g->x_grid_min : g->y_grid_min;
      const double gmax = gmin + (double)((dir == ORC_DIR_X ? g->nx : g->ny) - 1) * h;
      if (fabs(r[0] - r[1]) <= C_TINY) { /* dist_fn.F90:200-205, 248-253 */
        r[0] = gmin - 0.5 * h;
        r[1] = gmax + 0.5 * h;
      } else { /* :206-229: snap to cell boundaries inside the domain */
        r[0] = fmax(r[0], lo);
        r[1] = fmin(r[1], hi);
        r[0] = lo + floor((r[0] - lo) / h) * h;
        r[1] = lo + ceil((r[1] - lo) / h) * h;
        r[0] = fmax(r[0], gmin - 0.5 * h);
        r[1] = fmin(r[1], gmax + 0.5 * h);
      }
      resolution[d] = (int)lround((r[1] - r[0]) / h);
      if (resolution[d] <= 0) resolution[d] = 0;
      dgrid[d] = h;
      parallel[d] = 1;
      continue;
    }
    if (dir < ORC_DIR_PX || dir > ORC_DIR_MOD_P) return 1;
    if (fabs(r[0] - r[1]) <= C_TINY) { calc[d] = 1; any_calc = 1; }
    if (dir >= ORC_DIR_XY_ANGLE && dir <= ORC_DIR_ZX_ANGLE) need[dir - ORC_DIR_XY_ANGLE] = 1;
  }
  for (int k = 0; k < 3; ++k)
    if (use_restrictions[ORC_DIR_XY_ANGLE - 1 + k]) need[k] = 1;
  double pd[ORC_DF_MAXDIRS + 1];
  if (any_calc) { /* :341-453 */
    for (int d = 0; d < ndims; ++d)
      if (calc[d]) { ranges[2 * d] = 1.7976931348623157e308 * 0.4; ranges[2 * d + 1] = -1.7976931348623157e308 * 0.4; }
    for (int64_t ip = 0; ip < s->n; ++ip) {
      if (!df_particle_data(s, ip, need, NULL, pd)) continue;
      if (df_restricted(pd, restrictions, use_restrictions)) continue;
      for (int d = 0; d < ndims; ++d)
        if (calc[d]) {
          const double v = pd[direction[d]];
          if (v < ranges[2 * d]) ranges[2 * d] = v;
          if (v > ranges[2 * d + 1]) ranges[2 * d + 1] = v;
        }
    }
  }
  double amax[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < ndims; ++d) { /* :455-472 */
    double *r = ranges + 2 * d;
    if (fabs(r[0] - r[1]) <= C_TINY) { r[0] = -1.0; r[1] = 1.0; }
    const int dir = direction[d];
    if (dir >= ORC_DIR_XY_ANGLE && dir <= ORC_DIR_ZX_ANGLE) amax[dir - ORC_DIR_XY_ANGLE] = r[1];
    if (!parallel[d]) dgrid[d] = (r[1] - r[0]) / (double)resolution[d];
  }
  for (int d = ndims; d < 3; ++d) resolution[d] = 1;
  if (!out) return 0;
  const int64_t ncell = (int64_t)resolution[0] * resolution[1] * resolution[2];
  for (int64_t k = 0; k < ncell; ++k) out[k] = 0.0;
  if (ncell == 0) return 0;
  for (int64_t ip = 0; ip < s->n; ++ip) { /* :482-578 */
    if (!df_particle_data(s, ip, need, amax, pd)) continue;
    if (df_restricted(pd, restrictions, use_restrictions)) continue;
    int64_t cell[3] = {0, 0, 0};
    int ok = 1;
    for (int d = 0; d < ndims; ++d) {
      const int64_t c = (int64_t)floor((pd[direction[d]] - ranges[2 * d]) / dgrid[d]);
      if (c < 0 || c >= resolution[d]) { ok = 0; break; }
      cell[d] = c;
    }
    if (!ok) continue;
    out[(cell[2] * resolution[1] + cell[1]) * resolution[0] + cell[0]] += s->w[ip];
  }
  return 0;
}
