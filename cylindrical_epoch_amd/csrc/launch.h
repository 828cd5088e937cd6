// Device-state containers and kernel launcher declarations.
#pragma once

#include "common.h"

namespace cylpic {

// The fifteen mode arrays of the reference (shared_data.F90:474-478) in
// the device precision T, plus the outflow-boundary snapshots.
template <class T>
struct FieldSet {
  cplx<T>* f[15] = {};
  cplx<T>* bnd_xmin[6] = {};
  cplx<T>* bnd_xmax[6] = {};
  cplx<T>* ex() { return f[0]; }
  cplx<T>* er() { return f[1]; }
  cplx<T>* et() { return f[2]; }
  cplx<T>* bx() { return f[3]; }
  cplx<T>* br() { return f[4]; }
  cplx<T>* bt() { return f[5]; }
  cplx<T>* jx() { return f[6]; }
  cplx<T>* jr() { return f[7]; }
  cplx<T>* jt() { return f[8]; }
  cplx<T>* bx_old() { return f[9]; }
  cplx<T>* br_old() { return f[10]; }
  cplx<T>* bt_old() { return f[11]; }
  cplx<T>* jx_old() { return f[12]; }
  cplx<T>* jr_old() { return f[13]; }
  cplx<T>* jt_old() { return f[14]; }
};

// Geometry tables of push_particles (particles.F90:190-217), FP64,
// element [iy + jng] holds Fortran index iy.
struct Geo {
  const double* inv_area_rt;
  const double* inv_area_xt;
  const double* inv_volume;
  const double* ratio_area_xt;
};

// Per-species constants of push_particles (particles.F90:159-261)
struct PushConst {
  double part_q, part_mc, ipart_mc, cmratio, ccmratio;
  double idx, idy, idt, dtco2;
};

// FP64 particle arrays (SoA).  key = tile of the particle's cell, or
// ntiles for a removed particle.
struct PartF64 {
  double *x, *y, *z, *px, *py, *pz, *w;
  int* key;
};

// Mixed-precision particle arrays: x as integer cell (1..nx) plus an FP32
// offset from the cell centre in units of dx; y, z in units of dy; momenta
// as u = p / (m c); FP32 weight.
struct PartMixed {
  int* cell;
  float *ox, *y, *z, *ux, *uy, *uz, *w;
  int* key;
};

// fields.hip
template <class T> void launch_update_e(const Grid&, FieldSet<T>&, hipStream_t);
// copy_old: also B_old = B (every element) before the update, in the same pass
template <class T> void launch_update_b(const Grid&, FieldSet<T>&, int copy_old, hipStream_t);
template <class T> void launch_efield_bcs(const Grid&, FieldSet<T>&, hipStream_t);
template <class T> void launch_bfield_bcs(const Grid&, FieldSet<T>&, int mpi_only, hipStream_t);
template <class T> void launch_setup_bnd(const Grid&, FieldSet<T>&, hipStream_t);
template <class T>
void launch_bfield_final_bcs(const Grid&, FieldSet<T>&, const double* src_xmin, const double* src_xmax,
                             hipStream_t);
template <class T> void launch_current_axis(const Grid&, FieldSet<T>&, hipStream_t);
template <class T> void launch_current_finish(const Grid&, FieldSet<T>&, hipStream_t);
template <class T> void launch_fill_zero(cplx<T>*, long n, hipStream_t);
template <class T> void launch_zero_j(FieldSet<T>&, long n, hipStream_t);  // jx, jr, jt = 0
template <class T> void launch_energy(const Grid&, FieldSet<T>&, double* out, hipStream_t);
// moving window: shift E, B, J one cell towards x_min (through `spare`, a
// field-sized buffer that is swapped with each array) and fix the x_max
// incoming cells from the stored boundary columns
template <class T> void launch_shift_window_fields(const Grid&, FieldSet<T>&, cplx<T>*& spare, hipStream_t);
// smooth_current: `its` binomial iterations plus `comp_its` compensating
// ones with the given strides (host array, each in 1..ng); wk: a field-sized
// work buffer
template <class T>
void launch_smooth_current(const Grid&, FieldSet<T>&, cplx<T>* wk, int its, int comp_its, const int* strides,
                           int nstrides, hipStream_t);
void launch_d2f(const cplx<double>* src, cplx<float>* dst, long n, hipStream_t);
void launch_f2d(const cplx<float>* src, cplx<double>* dst, long n, hipStream_t);

// push.hip
// Neighbour-migration direction counts a push leaves for the next sort
// (counts: [ntiles][MIG_DIRS] of the tiles' main ranges, counts_x of their
// extra ranges; flag: 2 ints zeroed before the push); a null counts pointer
// skips the counting.
constexpr int MIG_DIRS = 10;
constexpr int MIG_REMOVED = 9;
struct MigCount {
  uint32_t* counts;
  uint32_t* counts_x;
  int* flag;
};

// Particles appended since the last sort (the moving window's injection):
// tile t's extra range is [base + start[t], base + start[t + 1]), in key
// order after the main ranges; start = null: none.
struct ExtraRange {
  const int* start;
  int base;
};
void launch_push_deposit_f64(const Grid&, const Tiles&, const PushConst&, const PartF64&,
                             const int* tile_start, ExtraRange, FieldSet<double>&, const Geo&, int ldsf, int* err,
                             MigCount, hipStream_t);
// key of every particle: its tile (cell_key 0), or 4 x its tile-major cell
// (Tiles::key) + half-cell quadrant (1; 2: the same, except that a particle
// whose key marks it removed, >= ntiles, gets ntiles * cells * 4)
void launch_keys_f64(const Grid&, const Tiles&, const PartF64&, long n, int cell_key, hipStream_t);
// push_mixed.hip
int launch_push_deposit_mixed(const Grid&, const Tiles&, const PushConst&, double charge, const PartMixed&,
                              const int* tile_start, ExtraRange, FieldSet<float>&, const float* geo, int ntab,
                              int acc, int ldsf, int* err, MigCount, hipStream_t);
void launch_keys_mixed(const Grid&, const Tiles&, const PartMixed&, long n, int cell_key, hipStream_t);
// The batched mixed push: the arguments of one (run, species) push as an
// opaque descriptor of push_desc_mixed_bytes(), many of which one persistent
// launch works through (tiles of 16 x 4, the int64 box, the LDS gather: the
// defaults).  push_batch_blocks: the launch's workgroups (resident ones)
// for M modes and `lds` bytes; counters: NUM_XCD ints of scratch.
size_t push_desc_mixed_bytes();
void fill_push_desc_mixed(void* out, const Grid&, const Tiles&, const PushConst&, double charge, const PartMixed&,
                          const int* tile_start, ExtraRange, FieldSet<float>&, const float* geo, int ntab, int* err,
                          MigCount);
bool batch_push_supported(const Tiles&, int M, int acc, int ldsf);
int push_batch_blocks(int M, size_t lds);
void launch_push_deposit_mixed_batch(const void* descs, const int* item_start, int ndesc, int M, size_t lds,
                                     int* counters, int nblocks, hipStream_t);
// moving window: remove the particles left of the moved x_min (key = ntiles);
// the mixed path's cell numbers drop by ncells
void launch_window_mixed(const PartMixed&, long n, int ncells, int ntiles, hipStream_t);
void launch_window_f64(const Grid&, const PartF64&, long n, int ntiles, hipStream_t);
void launch_to_mixed(const Grid&, double mass, const double* const* src, const PartMixed&, long n, hipStream_t);
void launch_from_mixed(const Grid&, double mass, const PartMixed&, double* const* dst, long n, hipStream_t);
size_t push_lds_bytes_mixed(const Grid&, const Tiles&, int acc, int ldsf);
size_t push_lds_bytes(const Grid&, const Tiles&, int ldsf);

// A device-side condition on a launch: the kernel runs when p is null or
// when (p[0] != 0) == (when != 0).  The neighbour migration sort and its
// radix fallback are queued together with opposite guards on the
// migration flag, so the device picks one and the host never waits to read
// the flag.
struct Guard {
  const int* p = nullptr;
  int when = 0;
};
__device__ __forceinline__ bool guard_skip(const Guard& gd) { return gd.p && ((gd.p[0] != 0) != (gd.when != 0)); }

// sort.hip
struct SortScratch {
  uint32_t *k[2], *v[2];
  uint32_t* counts;
  uint32_t* scan_tmp;
  long cap;
};
size_t sort_scratch_words(long n);
// Stable sort of the composite keys (tile << slot_bits) | (i mod 2^slot_bits)
// of n particles with tile keys in [0, ntiles] (ntiles = removed); slot_bits
// = 0 sorts by tile alone.  On return perm[i] is the source index of sorted
// position i and tiles_out[i] its tile key.
void radix_sort(const uint32_t* tiles_in, long n, int ntiles, int slot_bits, SortScratch& s,
                uint32_t* tiles_out, uint32_t** perm_out, hipStream_t, Guard = Guard{});
void launch_tile_bounds(const uint32_t* sorted_tiles, long n, int ntiles, int* tile_start, hipStream_t,
                        Guard = Guard{});
// keys of the slots [tile_start[ntiles], n_main) set to ntiles (removed): the
// slots a migration sort left behind its kept particles, before a radix sort
// over the host's upper bound n_main
void launch_mark_stale(uint32_t* key, const int* tile_start, int ntiles, long n_main, hipStream_t, Guard = Guard{});
// counter[0] += 1 when the guard passes (device-side sort statistics)
void launch_guard_count(int* counter, hipStream_t, Guard);
// n 32-bit words set to zero by a kernel
void launch_zero_words(void* p, long n, hipStream_t);
// tile columns relabelled one down (keys of n particles, null: ranges only);
// ts_out = the ranges under the new labels
void launch_relabel_tiles(uint32_t* key, long n, int ntiles, const int* ts_in, int* ts_out, hipStream_t);
// key[i] /= div (tile-major cell keys -> tile keys)
void launch_div_keys(uint32_t* key, long n, uint32_t div, hipStream_t);
// map[dest(i)] = perm[i], dest(i) the place of sorted index i under the
// per-tile rounds interleave, from the sorted tile-major cell keys (x 4 +
// quadrant) of n particles, removed ones keyed ntiles * cells * 4: the
// gather by map (in place of perm) writes that layout
void launch_round_map(const uint32_t* sorted_cell_keys, const uint32_t* perm, long n, int ntiles, int cells,
                      uint32_t* map, hipStream_t);
// Payload permutation dst[i] = src[perm[i]]
void launch_gather_f64(const uint32_t* perm, long n, const double* const* src, double* const* dst,
                       int narr, hipStream_t, Guard = Guard{});
void launch_gather_32(const uint32_t* perm, long n, const uint32_t* const* src, uint32_t* const* dst,
                      int narr, hipStream_t, Guard = Guard{});
void launch_gather_64(const uint32_t* perm, long n, const uint64_t* src, uint64_t* dst, hipStream_t,
                      Guard = Guard{});

// Neighbour migration sort (keys tile-sorted before the last push).
// launch_mig_count fills the direction counts (as the push kernels do);
// mc.flag must be zeroed first.  mc.flag[0] != 0: a particle's tile moved
// beyond the 3 x 3 neighbourhood (sort with radix_sort instead).
// launch_mig_scatter then writes the sorted payload dst/key_out/id_out (ids
// optional) and tile_start_new (ntiles + 2 ints); scratch holds
// mig_scratch_words(ntiles) words.
size_t mig_scratch_words(int ntiles);
void launch_mig_count(const uint32_t* key, const int* tile_start, ExtraRange xr, const Tiles& tl, MigCount mc,
                      hipStream_t);
void launch_mig_scatter(const uint32_t* key, const int* tile_start, ExtraRange xr, const Tiles& tl, MigCount mc,
                        uint32_t* scratch, int* tile_start_new, int word_bytes, const void* const* src,
                        void* const* dst, int narr, uint32_t* key_out, const uint64_t* id_in, uint64_t* id_out,
                        hipStream_t, Guard = Guard{});

// diag.hip: output diagnostics
// number density modes (complex FP64, field-shaped `out`, zeroed by the
// caller); then launch_density_bcs applies the ghost sums and zero-gradient
// ghosts of calc_boundary_modes / calc_number_density_modes
// (one workgroup per tile over the tile ranges when tile_start is given)
void launch_number_density_f64(const Grid&, const Tiles&, const PartF64&, const int* tile_start, ExtraRange,
                               long n, cplx<double>* out, hipStream_t);
void launch_number_density_mixed(const Grid&, const Tiles&, const PartMixed&, const int* tile_start, ExtraRange,
                                 long n, cplx<double>* out, hipStream_t);
void launch_density_bcs(const Grid&, cplx<double>* a, hipStream_t);
// continuity check: the m = 0 charge at x + vs v dt/2 on the staggered grid
// (one plane, zeroed by the caller; particles with key >= ntiles skipped;
// vs = 1 for the continuity check, -1 for the half step behind, 0 for a
// species that does not move), then the residual maxima (out[3], bit
// patterns, zeroed by the caller)
void launch_charge_half_f64(const Grid&, const Tiles&, const PartF64&, const int* tile_start, ExtraRange, long n,
                            double mass, double q, double vs, double* out, hipStream_t);
// the mixed path's charge with the push's own FP32 half step (push_mixed.hip)
void launch_charge_half_mixed(const Grid&, const Tiles&, const PartMixed&, const int* tile_start, ExtraRange,
                              long n, double charge, double vs, double* out, hipStream_t);
template <class T>
void launch_continuity_m0(const Grid&, const double* q0, const double* q1, const cplx<T>* jx, const cplx<T>* jr,
                          const Geo& geo, unsigned long long* out, hipStream_t);
// Continuity of every mode (cont_modes.h): the charge modes at the push's
// half step before it (Q0: M complex planes), then after it the charge modes
// at t + 3dt/2 (Q1) and the azimuthal terms Z, Zp (M complex planes each),
// all zeroed by the caller; then per mode m the maxima over the same cells
// of |Q1 - Q0 + dt (x-r flux of J_m) - Z|, |Q1 - Q0|, |Z - Zp| and |Z| into
// out[4 m .. 4 m + 3] (bit patterns, zeroed by the caller)
// (ts, xr: the species' tile ranges and appended ranges, which hold its
// particles; ts null: every slot below n is a particle)
// (planes: the charge set at the end, then Z and Zp, `stride` words apart;
// with tile ranges a workgroup per tile sums into LDS first)
void launch_cont_modes_begin_f64(const Grid&, const Tiles&, const PartF64&, long n, const int* ts, ExtraRange xr,
                                 double mass, double q, double* q0, hipStream_t);
void launch_cont_modes_end_f64(const Grid&, const Tiles&, const PartF64&, long n, const int* ts, ExtraRange xr,
                               double mass, double q, double* planes, long stride, hipStream_t);
void launch_cont_modes_begin_mixed(const Grid&, const Tiles&, const PartMixed&, long n, const int* ts, ExtraRange xr,
                                   double q, double* q0, hipStream_t);
void launch_cont_modes_end_mixed(const Grid&, const Tiles&, const PartMixed&, long n, const int* ts, ExtraRange xr,
                                 double q, double* planes, long stride, hipStream_t);
template <class T>
void launch_continuity_modes(const Grid&, const double* q0, const double* q1, const double* z, const double* zp,
                             const cplx<T>* jx, const cplx<T>* jr, const Geo& geo, unsigned long long* out,
                             hipStream_t);
// Gauss's law for m = 0 on the same cells: max |eps0 (net flux of Ex, Er out
// of the cell) - (qa + qb) / 2|, max |(qa + qb) / 2|, max |eps0 net flux|
// (out[3] as above)
template <class T>
void launch_gauss_m0(const Grid&, const double* qa, const double* qb, const cplx<T>* ex, const cplx<T>* er,
                     const Geo& geo, unsigned long long* out, hipStream_t);
// n doubles set to zero by a kernel (vector stores), for buffers that must
// be cleared without hipMemsetAsync
void launch_zero_f64(double* p, long n, hipStream_t);

// Distribution functions: c_dir_* numbering (constants.F90:231-247)
enum {
  DF_X = 1, DF_Y = 2, DF_PX = 3, DF_PY = 4, DF_PZ = 5, DF_EN = 6, DF_GAMMA_M1 = 7, DF_XY_ANGLE = 8,
  DF_YZ_ANGLE = 9, DF_ZX_ANGLE = 10, DF_MOD_P = 11, DF_MAXDIRS = 11, DF_MAXDIMS = 3
};
constexpr int DF_LDS_BINS = 4096;  // histograms up to this size are summed per workgroup in LDS
struct DistFn {
  int ndims;
  int direction[3];
  double range[3][2];
  double dgrid[3];
  int resolution[3];
  int calc_range[3];
  int need_angle[3];  // xy, yz, zx
  double angle_max[3];
  double restriction[DF_MAXDIRS][2];
  int use_restriction[DF_MAXDIRS];
};
// per-block (min, max) of each auto-ranged direction: part[nblocks][6]
void launch_df_range_f64(const Grid&, const PartF64&, long n, double mc, const DistFn&, double* part, int nblocks,
                         hipStream_t);
void launch_df_range_mixed(const Grid&, const PartMixed&, long n, double mc, const DistFn&, double* part,
                           int nblocks, hipStream_t);
// adds each particle's weight into its bin of `out` (zeroed by the caller)
void launch_df_bin_f64(const Grid&, const PartF64&, long n, double mc, const DistFn&, double* out, int nblocks,
                       hipStream_t);
void launch_df_bin_mixed(const Grid&, const PartMixed&, long n, double mc, const DistFn&, double* out,
                         int nblocks, hipStream_t);

}  // namespace cylpic
