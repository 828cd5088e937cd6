/*
 * cylpic_oracle.h -- FP64 CPU restatement of cylindrical EPOCH's quasi-3D
 * particle-in-cell timestep (Boris push, charge-conserving mode deposit,
 * per-mode FDTD with the r=0 axis rows, field/current/particle boundaries).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * load it.  The product never links, calls or falls back to it.
 *
 * Conventions (identical to the reference Fortran):
 *   - x is the cylinder axis, y is the radius ("r"), particles carry
 *     Cartesian (x, y, z) with r = sqrt(y^2 + z^2)   (src/particles.F90:315-323)
 *   - every mode array has the Fortran shape
 *       COMPLEX(8) f(1-ng:nx+ng, 1-ng:ny+ng, 0:n_mode-1)
 *     stored column-major (x fastest), complex interleaved (re, im)
 *                                                    (src/shared_data.F90:474-478)
 *   - ng = jng = 5 (triangle shape: png = 3, ng = png + 2)
 *                                                    (src/constants.F90:524-545)
 *   - one rank: x_min_local = x_min, y_grid_min_local = y_grid_min, the r=0
 *     axis and r_max boundaries are both local.
 */
#ifndef CYLPIC_ORACLE_H
#define CYLPIC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary codes: the values of c_bc_* in src/constants.F90:56-72. */
enum {
  ORC_BC_PERIODIC = 1, ORC_BC_OTHER = 2, ORC_BC_SIMPLE_LASER = 3,
  ORC_BC_SIMPLE_OUTFLOW = 4, ORC_BC_OPEN = 5, ORC_BC_ZERO_GRADIENT = 7,
  ORC_BC_CLAMP = 8, ORC_BC_REFLECT = 9, ORC_BC_CONDUCT = 10,
  ORC_BC_THERMAL = 11, ORC_BC_ZERO_B = 16
};
/* Boundary slots: c_bd_x_min .. c_bd_y_max minus one. */
enum { ORC_XMIN = 0, ORC_XMAX = 1, ORC_YMIN = 2, ORC_YMAX = 3 };

typedef struct { double re, im; } ocplx;

typedef struct {
  int nx, ny, n_mode, ng;          /* ng == jng == 5                        */
  double dx, dy, dt;
  double x_min, x_max, y_min, y_max;
  double x_grid_min, y_grid_min;   /* centre of cell 1 (setup.F90:177-180)  */
  int bc_field[4];                 /* after setup_boundaries' mapping       */
  int bc_particle[4];              /* shared by all species (no "mixed")    */
  double time;
} orc_grid;

typedef struct {
  /* nine live arrays, six *_old copies, each (nx+2ng)*(ny+2ng)*n_mode */
  ocplx *ex, *er, *et, *bx, *br, *bt, *jx, *jr, *jt;
  ocplx *bx_old, *br_old, *bt_old, *jx_old, *jr_old, *jt_old;
  /* boundary snapshots of setup_field_boundaries, (ny+2ng)*n_mode each,
   * order: ex, er, et, bx, br, bt; x_min then x_max      (setup.F90:393-420) */
  ocplx *bnd_xmin[6], *bnd_xmax[6];
} orc_fields;

typedef struct {
  int64_t n;
  double *x, *y, *z, *px, *py, *pz, *w;
  int64_t *id;
  double q, m;
  int immobile;
} orc_species;

/* Returns and clears the error flag of the last pushes: nonzero when a
 * particle moved more than one cell in a step (or had a non-finite
 * position), which the reference's deposit cannot represent; that
 * particle's update was skipped. */
int orc_take_error(void);

/* ---- field solver (src/fields.f90) ---------------------------------- */
void orc_update_e_field(const orc_grid *g, orc_fields *f);
void orc_update_b_field(const orc_grid *g, orc_fields *f);
void orc_update_eb_fields_half(const orc_grid *g, orc_fields *f);
/* src_xmin / src_xmax: host-evaluated laser sources, 2*(ny+1) doubles each
 * (source1(0:ny) then source2(0:ny), laser.f90:444-461), or NULL for 0. */
void orc_update_eb_fields_final(const orc_grid *g, orc_fields *f,
                                const double *src_xmin, const double *src_xmax);

/* ---- field boundaries (src/boundary.F90, src/laser.f90) -------------- */
void orc_efield_bcs(const orc_grid *g, orc_fields *f);
void orc_bfield_bcs(const orc_grid *g, orc_fields *f, int mpi_only);
void orc_bfield_final_bcs(const orc_grid *g, orc_fields *f,
                          const double *src_xmin, const double *src_xmax);
void orc_outflow_bcs_x_min(const orc_grid *g, orc_fields *f, const double *src);
void orc_outflow_bcs_x_max(const orc_grid *g, orc_fields *f, const double *src);
void orc_outflow_bcs_r_max(const orc_grid *g, orc_fields *f);
void orc_setup_field_boundaries(const orc_grid *g, orc_fields *f);

/* ---- particles (src/particles.F90, src/boundary.F90) ----------------- */
void orc_push_particles(const orc_grid *g, orc_fields *f,
                        orc_species *sp, int n_species);
/* The same push with the particles cut into nthreads x slabs pushed in
 * parallel (OpenMP), each depositing into its own current arrays, added in
 * slab order: for timing the CPU baseline on every core; it differs from
 * orc_push_particles only in the order of the current sums. */
void orc_push_particles_threads(const orc_grid *g, orc_fields *f, orc_species *sp, int n_species, int nthreads);
void orc_current_bcs(const orc_grid *g, orc_fields *f);
void orc_current_bcs_r_min_final(const orc_grid *g, orc_fields *f);
void orc_current_finish(const orc_grid *g, orc_fields *f);
/* smooth_current (current_smooth.F90:50-58, smooth_mode_array :145-196) */
void orc_smooth_current(const orc_grid *g, orc_fields *f, int its, int comp_its,
                        const int *strides, int nstrides);
void orc_particle_bcs(const orc_grid *g, orc_species *sp, int n_species);
/* ---- moving window (src/housekeeping/window.F90) -------------------- */
/* shift_window for ncells cells: per cell the grid moves by dx, particles
 * with x < x_min are removed and E, B, J shift one cell towards x_min with
 * the x_max incoming-cell fix.  The injection of new particles is the
 * caller's (append them to the species before the call). */
void orc_shift_window(orc_grid *g, orc_fields *f, orc_species *sp, int n_species, int ncells);
void orc_geometry(const orc_grid *g, double *inv_area_rt, double *inv_area_xt,
                  double *inv_volume, double *ratio_area_xt);

/* Deposit one particle's current from its weights (the DO im / DO iy /
 * DO ix nest of particles.F90:586-665).  gx,hx,gy,hy are 5-element arrays
 * indexed -2..2 (pass pointer to element -2).  cell_x2, cell_y2 are the
 * t+1/2 staggered cells; dcellx/dcelly the t+1.5 shifts. */
void orc_deposit_from_weights(const orc_grid *g, orc_fields *f,
                              const double *gx, const double *hx,
                              const double *gy, const double *hy,
                              int cell_x2, int cell_y2, int dcellx, int dcelly,
                              double fcx, double fcy, double fcz,
                              ocplx exp_itheta_05, ocplx exp_idtheta,
                              double dtheta,
                              const double *inv_area_rt,
                              const double *inv_area_xt,
                              const double *inv_volume,
                              const double *ratio_area_xt);

/* ---- cell key sort (key: src/housekeeping/split_particle.F90:61-62) --- */
/* Tile-major cell key: tiles of tile_x * tile_r cells, tiles row-major in x,
 * cells inside a tile row-major in x; the tile columns start `shift` cells
 * before cell 1.  Cells are clamped into 1..nx, 1..ny (a particle on the
 * outer edge within rounding). */
void orc_cell_keys(const orc_grid *g, int64_t n, const double *x,
                   const double *y, const double *z, int tile_x, int tile_r,
                   int shift, int32_t *keys);
/* Stable counting sort of keys in [0, nkeys): perm[i] = source index. */
void orc_stable_sort_keys(int64_t n, const int32_t *keys, int32_t nkeys,
                          int64_t *perm);

/* ---- setup scalars (src/housekeeping/setup.F90:629-673) -------------- */
double orc_set_dt(double dx, double dy, double dt_multiplier,
                  double dt_plasma_frequency, double dt_laser);

/* ---- diagnostics with no reference routine (the reference's Cartesian
 * energy sums are never updated in the cylindrical code) -------------- */
void orc_mode_energy(const orc_grid *g, const orc_fields *f, double *w_e,
                     double *w_b);

/* ---- output diagnostics (src/io/calc_df.F90, src/io/dist_fn.F90) ----- */
/* calc_number_density_modes (calc_df.F90:588-661): species >= 0 one
 * species, < 0 the sum over all; out: one mode array. */
void orc_number_density_modes(const orc_grid *g, const orc_species *sp, int n_species,
                              int species, ocplx *out);
/* Directions of a distribution function: c_dir_* (constants.F90:231-247). */
enum {
  ORC_DIR_X = 1, ORC_DIR_Y = 2, ORC_DIR_PX = 3, ORC_DIR_PY = 4, ORC_DIR_PZ = 5,
  ORC_DIR_EN = 6, ORC_DIR_GAMMA_M1 = 7, ORC_DIR_XY_ANGLE = 8, ORC_DIR_YZ_ANGLE = 9,
  ORC_DIR_ZX_ANGLE = 10, ORC_DIR_MOD_P = 11, ORC_DF_MAXDIRS = 11, ORC_DF_MAXDIMS = 3
};
/* general_dist_fn on one rank (dist_fn.F90:118-715).  ranges[2*3] and
 * resolution[3] are the requested values on entry and the ones used on
 * return; restrictions[2*11] / use_restrictions[11] per direction (index
 * direction - 1).  out (resolution[0]*resolution[1]*resolution[2] doubles,
 * first index fastest) is written unless NULL.  Returns 0, 1 for an unknown
 * direction, 2 for an empty species (nothing written). */
int orc_dist_fn(const orc_grid *g, const orc_species *s, int ndims, const int *direction,
                double *ranges, int *resolution, const double *restrictions,
                const int *use_restrictions, double *out);

#ifdef __cplusplus
}
#endif
#endif
