/*
 * cylpic.h -- C ABI of the MI355X (gfx950) quasi-3D PIC timestep library.
 *
 * The library replaces the argument-less Fortran hot-path subroutines of
 * cylindrical EPOCH (src/epoch2d.F90:213-265 main loop) with calls on an
 * opaque handle that owns GPU-resident fields and particles.  Plain C types
 * only: doubles, ints, pointers and sizes.  Every call returns 0 on success
 * or a nonzero CYLPIC_E* code; cylpic_last_error() gives the message.  No C++
 * exception crosses this boundary.  The reference aborts through
 * abort_code(errcode) (src/housekeeping/utilities.f90:261-281); a Fortran
 * caller maps a nonzero return to abort_code(c_err_other).
 *
 * Conventions (src/shared_data.F90:474-478, src/constants.F90:524-545):
 *   x is the cylinder axis, y the radius; particles carry Cartesian
 *   (x, y, z) and momenta (px, py, pz) in SI units, weight in real
 *   particles per macro-particle.  Mode fields are COMPLEX(8)
 *   f(1-ng:nx+ng, 1-ng:ny+ng, 0:n_mode-1), column-major, interleaved
 *   (re, im) doubles: exactly what c_loc(exm) hands over from Fortran.
 *   ng = jng = 5 (triangle shape).
 */
#ifndef CYLPIC_H
#define CYLPIC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cylpic_ctx cylpic_ctx;

/* return codes */
#define CYLPIC_OK 0
#define CYLPIC_EINVAL 1     /* bad argument                         */
#define CYLPIC_EHIP 2       /* a HIP runtime call failed            */
#define CYLPIC_ESTATE 3     /* call out of order / handle not ready */
#define CYLPIC_ENOMEM 4     /* device allocation failed             */
#define CYLPIC_EPHYS 5      /* a particle moved more than the CFL allows */

/* precision of the device path */
#define CYLPIC_FP64 0       /* FP64 everywhere, the reference's method          */
#define CYLPIC_MIXED 1      /* int32 cell + FP32 offsets, FP32 fields/momenta,
                               compensated FP32 current tiles, FP64 axis rows   */

/* boundary codes = the reference's c_bc_* (src/constants.F90:56-72), given
 * as the deck sets them; the library applies setup_boundaries' mapping
 * (src/boundary.F90:30-131) itself. */

/* field identifiers for upload/download (src/shared_data.F90:474-478) */
enum {
  CYLPIC_EXM = 0, CYLPIC_ERM, CYLPIC_ETM, CYLPIC_BXM, CYLPIC_BRM, CYLPIC_BTM,
  CYLPIC_JXM, CYLPIC_JRM, CYLPIC_JTM,
  CYLPIC_BXM_OLD, CYLPIC_BRM_OLD, CYLPIC_BTM_OLD,
  CYLPIC_JXM_OLD, CYLPIC_JRM_OLD, CYLPIC_JTM_OLD,
  CYLPIC_NFIELDS
};

/* Create a handle on HIP device `device`.  Replaces the field allocation of
 * mpi_initialise (src/housekeeping/mpi_routines.F90:373-403) and the grid
 * scalars of setup_grid (src/housekeeping/setup.F90:164-208) for one rank
 * that owns the whole domain (the r=0 axis and r_max are local).
 * bc_field / bc_particle: x_min, x_max, y_min, y_max codes as in the deck.
 * All fields start at zero. */
int cylpic_init(cylpic_ctx **out, int nx, int ny, int n_mode, int ng,
                double dx, double dy, double dt,
                double x_min, double x_max, double y_min, double y_max,
                const int bc_field[4], const int bc_particle[4],
                int precision, int device);
int cylpic_finalize(cylpic_ctx *h);
const char *cylpic_last_error(void);

/* dt after set_dt (src/housekeeping/setup.F90:629-673) on the host */
int cylpic_set_dt(cylpic_ctx *h, double dt);
int cylpic_set_time(cylpic_ctx *h, double time);

/* Restart.  The reference's restart dump holds the fields (E, B, B_old, J,
 * J_old: restart_data, src/housekeeping/setup.F90:1199-1362), the
 * particles, time and step (the dump's header, :815), x_grid_min and the
 * window's shift fraction (:1094-1098).  Here the fields and particles go through
 * cylpic_download_field / cylpic_download_particles and back through the
 * uploads; the host keeps the shift fraction; and these two calls carry the
 * rest: time, the window's position (x_min, x_grid_min; x_max follows), the
 * push count (its remainder modulo the sort interval is the sort phase) and
 * the tile columns' offset.  On the new handle: cylpic_init with the same
 * grid, tile and tuning calls, cylpic_setup_field_boundaries while the
 * fields are still zero (the reference snapshots them in after_deck_last
 * before restart_data reads the dump, src/epoch2d.F90:113-117),
 * cylpic_set_restart_state before any particle upload, then the uploads. */
int cylpic_get_restart_state(cylpic_ctx *h, double *time, double *x_min, double *x_grid_min,
                             int64_t *push_count, int *tile_shift);
int cylpic_set_restart_state(cylpic_ctx *h, double time, double x_min, double x_grid_min,
                             int64_t push_count, int tile_shift);

/* Register species `id` (0-based, ids dense): charge, mass, immobile flag
 * (src/shared_data.F90:190-280); capacity = maximum particle count. */
int cylpic_set_species(cylpic_ctx *h, int id, double charge, double mass,
                       int immobile, int64_t capacity);

/* Copy a species' particles to / from the device, SoA staging arrays filled
 * from the linked list (src/housekeeping/partlist.F90:414-488 wire order:
 * pos3, p3, weight).  ids may be NULL (no identity tracking).  On the mixed
 * path every momentum component must be finite with |p| / (m c) < 1e18
 * (CYLPIC_EINVAL otherwise, nothing copied); the same holds for
 * cylpic_inject_particles. */
int cylpic_upload_particles(cylpic_ctx *h, int species, int64_t n,
                            const double *x, const double *y, const double *z,
                            const double *px, const double *py, const double *pz,
                            const double *w, const int64_t *ids);
int cylpic_download_particles(cylpic_ctx *h, int species, int64_t capacity,
                              int64_t *n_out, double *x, double *y, double *z,
                              double *px, double *py, double *pz, double *w,
                              int64_t *ids);
int cylpic_particle_count(cylpic_ctx *h, int species, int64_t *n_out);

/* One mode array in the Fortran layout above (c_loc(exm) etc.). */
int cylpic_upload_field(cylpic_ctx *h, int which, const double *f);
int cylpic_download_field(cylpic_ctx *h, int which, double *f);

/* setup_field_boundaries (src/housekeeping/setup.F90:393-420): snapshot the
 * x_min/x_max columns the outflow boundaries use.  Call after the initial
 * fields are uploaded. */
int cylpic_setup_field_boundaries(cylpic_ctx *h);

/* update_eb_fields_half (src/fields.f90:316-337): E to t+dt/2, efield_bcs,
 * B_old = B, B to t+dt/2, B halos. */
int cylpic_update_eb_fields_half(cylpic_ctx *h);

/* push_particles (src/particles.F90:28-734) for all mobile species:
 * J_old = J, J = 0, Boris push, charge-conserving mode deposit,
 * current_bcs_r_min_final (src/boundary.F90:1909-1959) and particle_bcs
 * (src/boundary.F90:1541-1889: periodic, reflect, open). */
int cylpic_push_particles(cylpic_ctx *h);

/* current_finish (src/housekeeping/current_smooth.F90:29-46): current_bcs
 * (reflection + periodic ghost sums) and J halos; no smoothing. */
int cylpic_current_finish(cylpic_ctx *h);

/* update_eb_fields_final (src/fields.f90:341-353) with bfield_final_bcs
 * (src/boundary.F90:1505-1537).  src_xmin / src_xmax: the laser sources the
 * host evaluated for this step through the deck parser, source1(0:ny) then
 * source2(0:ny) (src/laser.f90:444-461), 2*(ny+1) doubles; NULL = no laser. */
int cylpic_update_eb_fields_final(cylpic_ctx *h, const double *src_xmin,
                                  const double *src_xmax);

/* Current smoothing inside current_finish (smooth_current,
 * src/housekeeping/current_smooth.F90:50-58 with smooth_mode_array
 * :145-196): the deck's smooth_its, smooth_comp_its and smooth_strides (NULL
 * or nstrides = 0: stride 1; each stride in 1..ng).  its + comp_its = 0 turns
 * smoothing off (the default). */
int cylpic_set_smoothing(cylpic_ctx *h, int its, int comp_its, const int *strides, int nstrides);

/* The four calls above in main-loop order plus time += dt
 * (src/epoch2d.F90:213-265 without optional physics, output or window). */
int cylpic_step(cylpic_ctx *h, const double *src_xmin, const double *src_xmax);

/* Moving window (src/housekeeping/window.F90).  The host keeps the
 * reference's moving_window bookkeeping (window.F90:353-376: the shift
 * fraction, start and stop times, the after-move boundaries) and its particle
 * injection (insert_particles, window.F90:157-325: parser densities and the
 * RNG), and calls:
 *   cylpic_inject_particles  - before the shift, with the new particles of the
 *                              ncells columns beyond x_max (SI, as in upload);
 *                              they join the species without a full re-sort:
 *                              the batches since the last sort form one
 *                              region after the sorted particles, stable-
 *                              sorted by current tile key at each injection;
 *   cylpic_move_window       - shift_window for ncells cells: per cell the
 *                              grid moves by dx, E, B and J move one cell
 *                              towards x_min with the x_max incoming-cell fix
 *                              (shift_fields, window.F90:101-150), and the
 *                              particles left of x_min are removed
 *                              (remove_particles, window.F90:330-350).
 * x must not be periodic. */
int cylpic_inject_particles(cylpic_ctx *h, int species, int64_t n,
                            const double *x, const double *y, const double *z,
                            const double *px, const double *py, const double *pz,
                            const double *w, const int64_t *ids);
int cylpic_move_window(cylpic_ctx *h, int ncells);

/* Re-sort every species' particles by tile key (stable), dropping removed
 * particles.  The key of a particle is its cell (src/housekeeping/
 * split_particle.F90:61-62) mapped to its tile.  Sorting happens on its own
 * every `interval` steps (default 1); this forces one now. */
int cylpic_sort_particles(cylpic_ctx *h);
int cylpic_set_sort_interval(cylpic_ctx *h, int interval);
/* Sort method (no counterpart in the reference, which keeps no order):
 * CYLPIC_SORT_RADIX always runs the full LSD radix sort; CYLPIC_SORT_MIGRATE
 * (default) moves particles by one counting and one scatter pass when every
 * particle's tile moved at most one tile in x and r since the last sort
 * (always true for one push under the CFL limit), and falls back to the radix
 * sort otherwise.  Both give the stable order by tile key. */
#define CYLPIC_SORT_RADIX 0
#define CYLPIC_SORT_MIGRATE 1
/* CYLPIC_SORT_CELLS: every sort is the upload's sort (cell key and half-cell
 * quadrant, then each tile laid out in rounds of one particle per cell), so
 * that the push's wavefronts keep spanning distinct cells as particles drift
 * between cells; counted as a radix sort. */
#define CYLPIC_SORT_CELLS 2
int cylpic_set_sort_method(cylpic_ctx *h, int method);
/* Number of sorts done so far by each method. */
int cylpic_sort_counts(cylpic_ctx *h, int64_t *radix, int64_t *migrate);
/* Tile shape (cells) of the push/deposit work decomposition. */
int cylpic_set_tile(cylpic_ctx *h, int tile_x, int tile_r);
/* acc_mode (mixed path): per-tile current sums in FP32 (0), FP32 with
 * a 2Sum-compensated low word (1), FP64 (2), int64 fixed point (3, default:
 * exact, order-independent tile sums) or two 32-bit fixed-point halves packed
 * in one int64 per complex cell (4: half the atomics, ~1e-5 relative
 * resolution); gather_lds (both paths): stage the tile's fields in LDS (1,
 * the mixed path's default) or gather from L1/L2 (0, the FP64 default); 2
 * (mixed path, acc_mode 3) stages E only and gathers B from L1/L2.  Has no
 * counterpart in the reference (its deposit is serial FP64). */
int cylpic_tune(cylpic_ctx *h, int acc_mode, int gather_lds);

/* Autotuner (no counterpart in the reference): times `pushes` pushes (plus
 * one warm-up) of the current particles for each candidate tile shape and
 * gather source of this precision that fits in LDS, restores the particles,
 * the currents and their *_old copies after each, and keeps the fastest
 * candidate (the particles re-sorted under it).  Sorting is suspended
 * while it times, so every candidate is timed on the same pushes.  Returns
 * the choice and its time per push.  Call it with profiling off; the step
 * counters are restored. */
int cylpic_autotune(cylpic_ctx *h, int pushes, int *tile_x, int *tile_r, int *gather_lds,
                    double *ms_per_push);

/* Per-mode field energies (FP64 reduction; the reference's Cartesian
 * energy sums are never updated in cylindrical mode, src/housekeeping/
 * setup.F90:526-571), w_e[n_mode], w_b[n_mode] in joules. */
int cylpic_mode_energy(cylpic_ctx *h, double *w_e, double *w_b);

/* Output diagnostics from the device-resident particles.  A species that a
 * push has left with removed particles is sorted first (as for a download).
 *
 * number_density_mode (calc_number_density_modes, src/io/calc_df.F90:588-661,
 * with the ghost sums of calc_boundary_modes :35-55): species >= 0 one
 * species, < 0 the sum over all species; out: one complex mode array of the
 * field shape (as cylpic_download_field). */
int cylpic_number_density_modes(cylpic_ctx *h, int species, double *out);

/* Distribution functions (general_dist_fn, src/io/dist_fn.F90:118-715, one
 * rank).  direction[ndims] (ndims 1..3): CYLPIC_DIR_*; ranges[2*ndims]:
 * (lo, hi) per direction, lo == hi asks for the automatic range (momentum
 * directions) or the whole domain (x, y); resolution[ndims]: bins of each
 * momentum direction (x and y bin per cell).  On return ranges and
 * resolution hold the values used and *nbins their product;
 * restrictions[2*11] / use_restrictions[11] (or NULL) restrict each
 * direction (index direction - 1) to [lo, hi].  out (capacity doubles, first
 * direction fastest) receives the summed weights unless NULL.  An empty
 * species writes nothing and returns *nbins = 0. */
enum {
  CYLPIC_DIR_X = 1, CYLPIC_DIR_Y = 2, CYLPIC_DIR_PX = 3, CYLPIC_DIR_PY = 4, CYLPIC_DIR_PZ = 5,
  CYLPIC_DIR_EN = 6, CYLPIC_DIR_GAMMA_M1 = 7, CYLPIC_DIR_XY_ANGLE = 8, CYLPIC_DIR_YZ_ANGLE = 9,
  CYLPIC_DIR_ZX_ANGLE = 10, CYLPIC_DIR_MOD_P = 11
};
/* Continuity check of one push (no counterpart in the reference; it checks
 * the identity its charge-conserving deposit keeps for m = 0,
 * src/particles.F90:508-665).  cylpic_continuity_begin, called right before
 * cylpic_push_particles, stores the m = 0 charge of every mobile species at
 * x + v dt/2 on the staggered grid (quadratic-spline fractions; the mixed
 * path uses its push's own FP32 half step and weights); then
 * cylpic_continuity_end, right after the push and before
 * cylpic_current_finish, returns max |Q(t+3dt/2) - Q(t+dt/2) + dt (net Jx, Jr
 * current out of the cell)|, max |Q(t+3dt/2) - Q(t+dt/2)| and max
 * |Q(t+3dt/2)| in coulombs over the cells at least ng from every edge.
 *
 * The same pair also checks every mode m; cylpic_continuity_modes returns
 * the last check's per-mode maxima (arrays of n_mode, each may be null).
 * A particle's charge in mode m is q w Sx Sy c_m e^{i m theta} (c_0 = 1,
 * c_m = 2), the deposit moves it from t+dt/2 to t+3dt/2 along straight
 * weights and an angle theta + s dtheta (dtheta as the deposit takes it),
 * and its x-r currents carry all of the change except
 *   Z_m = i m dtheta q w c_m e^{i m theta_a} (f2 A + f3 B + f4 C)
 * (f_k(x) = int_0^1 s^(k-2) e^{i x s} ds at x = m dtheta, the reference's
 * m-factors; A, B, C the shape products of the deposit).  So
 *   residual[m] = max |Q_m(t+3dt/2) - Q_m(t+dt/2) + dt (net Jx, Jr) - Z_m|
 * is round-off for every mode (Z_0 = 0: the m = 0 identity above), with
 *   dq[m] = max |Q_m(t+3dt/2) - Q_m(t+dt/2)| its scale.  The deposited
 * J_theta carries v_theta(t+dt), not r dtheta / dt: Zp_m, Z_m with dtheta
 * replaced by v_theta dt / r at t+dt, is that current's azimuthal transport
 * particle by particle, and
 *   jtheta[m] = max |Z_m - Zp_m|,  z[m] = max |Z_m|
 * measure the difference; per particle it is r dtheta / (v_theta dt) - 1 =
 * (dt / r)^2 (3 v_r^2 - v_theta^2) / 12 + ..., at most (v_perp dt / r)^2 / 4
 * to leading order. */
int cylpic_continuity_begin(cylpic_ctx *h);
int cylpic_continuity_end(cylpic_ctx *h, double *residual, double *dq, double *q);
int cylpic_continuity_modes(cylpic_ctx *h, double *residual, double *dq, double *jtheta, double *z);

/* Gauss's law for m = 0 (no counterpart in the reference; the quantity its
 * charge-conserving deposit keeps, src/particles.F90:508-665, with the E
 * updates of src/fields.f90:53-113, 316-353, which apply the previous and
 * the new current for half a step each).  Called between steps: Q is the
 * mean of the m = 0 charges of every mobile species at x - v dt/2 and at
 * x + v dt/2 (staggered quadratic-spline fractions, as in the continuity
 * check) plus every immobile species at x; returns max |eps0 (net Ex, Er
 * flux out of the cell) - Q|, max |Q| and max |eps0 net flux| in coulombs
 * over the cells at least ng from every edge.  The first keeps its initial
 * value (zero for a neutral start at rest with E = 0 and J = 0) while the
 * charge stays away from the edges. */
int cylpic_gauss_residual_m0(cylpic_ctx *h, double *residual, double *q, double *flux);

/* cylpic_gauss_residual_m0 without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same kernels and result; the two charge planes and the
 * result words live in a 16 MiB array that the library holds on each
 * device until the process ends, so the grid's (nx + 2 ng) (ny + 2 ng) must
 * not exceed 1048576 cells (CYLPIC_EINVAL otherwise).  Every call on one
 * device shares that array: calls that overlap on a device (other threads,
 * other handles) race. */
int cylpic_gauss_residual_m0_noalloc(cylpic_ctx *h, double *residual, double *q, double *flux);

int cylpic_dist_fn(cylpic_ctx *h, int species, int ndims, const int *direction, double *ranges,
                   int *resolution, const double *restrictions, const int *use_restrictions,
                   double *out, int64_t capacity, int64_t *nbins);

/* Wait for all queued device work of this handle. */
int cylpic_synchronize(cylpic_ctx *h);

/* Kernel timing (HIP events on the library's stream).  After enabling,
 * cylpic_kernel_time(h, name, &ms, &count) returns the accumulated time and
 * launch count of a kernel family ("push_deposit", "update_e", ...). */
int cylpic_profile(cylpic_ctx *h, int enable);
int cylpic_kernel_time(cylpic_ctx *h, const char *name, double *ms, int64_t *count);

/* Batches: many independent runs (a parameter scan) stepped together.  No
 * counterpart in the reference, which runs one simulation per job; each
 * batch step is one cylpic_step (src/epoch2d.F90:213-265) of every run.
 * Steps are captured into HIP graphs and launched whole: each run's kernels
 * on its own stream, and the mixed-precision pushes of all runs with the
 * same mode count (default 16 x 4 tiles, accumulator and gather) in one
 * persistent launch over their (run, tile) items.  A run in a batch must
 * not be stepped on its own or have calls queued while batch steps are in
 * flight (cylpic_batch_synchronize first).
 *   cylpic_batch_create   the runs share one device; allocates what a step
 *                         could allocate lazily
 *   cylpic_batch_step     one step; src_xmin[r], src_xmax[r]: run r's laser
 *                         sources as in cylpic_update_eb_fields_final
 *                         (either array or any entry may be null); captured
 *                         and instantiated afresh every call
 *   cylpic_batch_steps    nsteps steps in chunks of P = lcm of the runs'
 *                         2 sort_interval (at most 64).  A chunk that starts
 *                         from a state seen before, and ends in the state it
 *                         started from (no window move, injection, upload or
 *                         download in between), relaunches the graph kept
 *                         for it: no host work per step.  src: null (no
 *                         laser) or nsteps blocks, one per step, each holding
 *                         per run in batch order 2 (ny + 1) doubles of x_min
 *                         sources then 2 (ny + 1) of x_max sources
 *   cylpic_batch_stats    steps taken, graphs instantiated, chunks that
 *                         relaunched a kept graph, batched push groups */
typedef struct cylpic_batch cylpic_batch;
int cylpic_batch_create(cylpic_batch **out, cylpic_ctx *const *runs, int nruns);
int cylpic_batch_step(cylpic_batch *b, const double *const *src_xmin, const double *const *src_xmax);
int cylpic_batch_steps(cylpic_batch *b, int nsteps, const double *src);
int cylpic_batch_synchronize(cylpic_batch *b);
int cylpic_batch_stats(cylpic_batch *b, int64_t *steps, int64_t *instantiations, int64_t *replays, int *push_groups);
int cylpic_batch_destroy(cylpic_batch *b);

#ifdef __cplusplus
}
#endif
#endif
