// C ABI of the gfx950 quasi-3D PIC library (declared in include/cylpic.h).
// Host-side state management: device allocations, the call sequence of the
// reference's main loop (src/epoch2d.F90:213-265) and the sort schedule.
#include "../../include/cylpic.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <map>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "launch.h"

using namespace cylpic;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(CYLPIC_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

struct Species {
  double q = 0, m = 0;
  int immobile = 0;
  long cap = 0;
  long n = 0;             // particles held; an upper bound while !n_exact
  bool n_exact = true;    // false after a device-decided sort until refresh_count
  bool ids = false;
  // FP64: x, y, z, px, py, pz, w (cur / alternate)
  double* a[7] = {};
  double* b[7] = {};
  // mixed: cell (int32 stored as uint32) + ox, y, z, ux, uy, uz, w (float)
  uint32_t* ma[8] = {};
  uint32_t* mb[8] = {};
  int64_t* id[2] = {};
  uint32_t* key[2] = {};      // tile key per particle (ntiles = removed), sort ping-pong
  int* tile_start = nullptr;      // ntiles + 2
  int* tile_start_alt = nullptr;  // ntiles + 2, migration sort output
  int* d_count = nullptr;         // == tile_start + ntiles
  bool dirty = false;             // pushed since the last sort (removed particles may remain)
  bool ranged = false;            // tile_start describes the tile ranges of the keys' last sort
  uint32_t* mig_counts = nullptr; // [ntiles][MIG_DIRS] direction counts for the migration sort
  uint32_t* mig_counts_x = nullptr;  // the same for the extra ranges
  int* extra_start = nullptr;     // ntiles + 2: tile ranges of the appended (injected) particles
  long n_extra = 0;               // particles appended after the main ranges since the last sort
  int* mig_flag = nullptr;        // 2 ints: farther move seen, particles kept
  bool counted = false;           // the last push left fresh direction counts
};

struct ProfEntry {
  double ms = 0;
  int64_t count = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

}  // namespace

struct cylpic_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  Grid g{};
  Tiles tl{};
  int prec = CYLPIC_FP64;
  FieldSet<double> f64;
  FieldSet<float> f32;
  double* geo = nullptr;
  float* geof = nullptr;  // FP32 copy for the mixed path
  int ntab = 0;
  double* src_dev[2] = {};
  double* src_host[2] = {};       // two pinned slots per boundary, used in turn
  hipEvent_t src_ev[2][2] = {};   // [boundary][slot]: that slot's copy to the device is done
  int src_slot = 0;
  int* d_err = nullptr;
  double* d_energy = nullptr;
  std::vector<Species> sp;
  SortScratch sc{};
  uint32_t* sc_mem = nullptr;
  long sc_cap = -1;
  int sort_interval = 1;
  int slot_bits = 0;
  uint32_t* mig = nullptr;  // migration sort scratch (mig_scratch_words(ntiles))
  int* d_sortstat = nullptr;  // device-decided sorts: [radix, migrate]
  bool migrate = true;      // use the neighbour migration sort when it applies
  bool sort_cells = false;  // every sort re-keys by cell and re-interleaves the tiles
  double* stage = nullptr;  // FP64 staging of host particles for the mixed path
  long stage_cap = 0;
  cplx<double>* spare64 = nullptr;  // field-sized buffer of the moving window's shifts and smoothing
  cplx<float>* spare32 = nullptr;
  cplx<double>* diag64 = nullptr;  // field-sized FP64 buffer of the number-density diagnostic
  double* diag_part = nullptr;      // per-block partial ranges of the distribution functions
  double* diag_hist = nullptr;      // distribution function bins
  long diag_hist_cap = 0;
  double* cont_q = nullptr;               // two m = 0 charge planes of the continuity check
  double* gauss_q = nullptr;              // two m = 0 charge planes of the Gauss check
  unsigned long long* cont_out = nullptr;  // the three maxima of either check
  bool cont_armed = false;
  double* cm_planes = nullptr;             // continuity of every mode: Q0, Q1, Z, Zp (M complex planes each)
  unsigned long long* cm_out = nullptr;    // its per-mode maxima (4 per mode)
  std::vector<double> cm_last;             // the last check's per-mode maxima, empty before one
  int smooth_its = 0, smooth_comp_its = 0;  // current smoothing (current_smooth.F90)
  std::vector<int> smooth_strides{1};
  long sort_radix = 0, sort_migrate = 0;
  int acc_mode = 3;    // mixed path: current accumulator (0 FP32, 1 FP32 + 2Sum, 2 FP64, 3 int64, 4 2x32 packed)
  int gather_lds = 1;  // gather from the LDS-staged field region (1) or L1/L2 (0); init sets it per precision
  int lds_max = 65536;  // LDS bytes per workgroup on the device
  long push_count = 0;
  bool sorted_once = false;
  bool profiling = false;
  std::map<std::string, ProfEntry> prof;
  std::vector<hipEvent_t> event_pool;
  double time = 0;
};

namespace {

// -------------------------------------------------------------- profiling
hipEvent_t get_event(cylpic_ctx* h) {
  if (!h->event_pool.empty()) {
    hipEvent_t e = h->event_pool.back();
    h->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct Scope {
  cylpic_ctx* h;
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
  Scope(cylpic_ctx* h_, const char* n) : h(h_), name(n) {
    if (h->profiling) {
      a = get_event(h);
      b = get_event(h);
      if (a) (void)hipEventRecord(a, h->stream);
    }
  }
  ~Scope() {
    if (h->profiling && a && b) {
      (void)hipEventRecord(b, h->stream);
      h->prof[name].pending.emplace_back(a, b);
    }
  }
};

void resolve_profile(cylpic_ctx* h) {
  for (auto& kv : h->prof) {
    for (auto& ev : kv.second.pending) {
      float ms = 0;
      (void)hipEventSynchronize(ev.second);
      if (hipEventElapsedTime(&ms, ev.first, ev.second) == hipSuccess) {
        kv.second.ms += ms;
        kv.second.count += 1;
      }
      h->event_pool.push_back(ev.first);
      h->event_pool.push_back(ev.second);
    }
    kv.second.pending.clear();
  }
}

// -------------------------------------------------------------- helpers
int map_field_bc(int b) {  // boundary.F90:30-68
  if (b == BC_OTHER) b = BC_CLAMP;
  if (b == BC_REFLECT) b = BC_CLAMP;
  if (b == BC_OPEN) b = BC_SIMPLE_OUTFLOW;
  return b;
}
int map_particle_bc(int b) {  // boundary.F90:98-131
  if (b == BC_OTHER || b == BC_CONDUCT) b = BC_REFLECT;
  if (b == BC_SIMPLE_LASER || b == BC_SIMPLE_OUTFLOW || b == BC_CPML_LASER || b == BC_CPML_OUTFLOW) b = BC_OPEN;
  return b;
}

// Geometry tables of push_particles, particles.F90:190-217
void geometry(const Grid& g, std::vector<double>& t) {
  const int ny = g.ny, jng = g.ng, n = ny + 2 * jng + 1;
  t.assign(4 * (size_t)n, 0.0);
  double* rt = t.data();
  double* xt = rt + n;
  double* vol = xt + n;
  double* rat = vol + n;
  const double dx = g.dx, dy = g.dy;
  double r_low = g.y_grid_min - (double)jng * dy;
  xt[0] = 1.0 / (((2.0 * PI) * std::fabs(r_low)) * dx);
  for (int iy = 1 - jng; iy <= ny + jng; ++iy) {
    if (std::lround((2.0 * r_low) / dy) == -1) {
      const double hh = 0.5 * dy;
      rt[iy + jng] = 1.0 / (PI * (hh * hh));
    } else {
      const double rh = r_low + dy;
      rt[iy + jng] = 1.0 / (PI * std::fabs(rh * rh - r_low * r_low));
    }
    xt[iy + jng] = 1.0 / (((2.0 * PI) * std::fabs(r_low + dy)) * dx);
    r_low = r_low + dy;
  }
  for (int iy = 1 - jng; iy <= ny + jng; ++iy) {
    vol[iy + jng] = rt[iy + jng] / dx;
    rat[iy + jng] = xt[iy + jng] / xt[iy - 1 + jng];
  }
}

Geo geo_of(cylpic_ctx* h) {
  Geo g;
  g.inv_area_rt = h->geo;
  g.inv_area_xt = h->geo + h->ntab;
  g.inv_volume = h->geo + 2 * h->ntab;
  g.ratio_area_xt = h->geo + 3 * h->ntab;
  return g;
}

long nfield(const cylpic_ctx* h) { return h->g.plane * h->g.M; }

// run fn on the field set of the handle's precision
template <class Fn>
void with_fields(cylpic_ctx* h, Fn fn) {
  if (h->prec == CYLPIC_FP64) fn(h->f64);
  else fn(h->f32);
}
size_t cplx_bytes(const cylpic_ctx* h) { return h->prec == CYLPIC_FP64 ? sizeof(cplx<double>) : sizeof(cplx<float>); }
void* field_ptr(cylpic_ctx* h, int k) { return h->prec == CYLPIC_FP64 ? (void*)h->f64.f[k] : (void*)h->f32.f[k]; }
void swap_fields(cylpic_ctx* h, int a, int b) {
  std::swap(h->f64.f[a], h->f64.f[b]);
  std::swap(h->f32.f[a], h->f32.f[b]);
}

int alloc_fields(cylpic_ctx* h) {
  const long n = nfield(h);
  const long nb = (long)h->g.ny2 * h->g.M;
  if (h->prec == CYLPIC_FP64) {
    for (int k = 0; k < 15; ++k) {
      HIPCHK(hipMalloc((void**)&h->f64.f[k], n * sizeof(cplx<double>)));
      launch_fill_zero(h->f64.f[k], n, h->stream);
    }
    for (int k = 0; k < 6; ++k) {
      HIPCHK(hipMalloc((void**)&h->f64.bnd_xmin[k], nb * sizeof(cplx<double>)));
      HIPCHK(hipMalloc((void**)&h->f64.bnd_xmax[k], nb * sizeof(cplx<double>)));
      launch_fill_zero(h->f64.bnd_xmin[k], nb, h->stream);
      launch_fill_zero(h->f64.bnd_xmax[k], nb, h->stream);
    }
  } else {
    for (int k = 0; k < 15; ++k) {
      HIPCHK(hipMalloc((void**)&h->f32.f[k], n * sizeof(cplx<float>)));
      launch_fill_zero(h->f32.f[k], n, h->stream);
    }
    for (int k = 0; k < 6; ++k) {
      HIPCHK(hipMalloc((void**)&h->f32.bnd_xmin[k], nb * sizeof(cplx<float>)));
      HIPCHK(hipMalloc((void**)&h->f32.bnd_xmax[k], nb * sizeof(cplx<float>)));
      launch_fill_zero(h->f32.bnd_xmin[k], nb, h->stream);
      launch_fill_zero(h->f32.bnd_xmax[k], nb, h->stream);
    }
  }
  HIPCHK(hipGetLastError());
  return CYLPIC_OK;
}

// Tile shape of the push; the handle keeps its previous shape on failure.
int set_tiles(cylpic_ctx* h, int tx, int tr) {
  if (tx < 1 || tr < 1) return fail(CYLPIC_EINVAL, "tile sizes must be >= 1");
  Tiles t;
  t.tx = tx;
  t.tr = tr;
  // the moving window's tile offset survives a change that keeps the width
  // (cylpic_tune); a new width starts from offset 0 (the callers re-key)
  t.shift = tx == h->tl.tx ? h->tl.shift : 0;
  // a spare column for the moving window's tile offset (non-periodic x only)
  const bool periodic_x = h->g.bc_field[XMIN] == BC_PERIODIC || h->g.bc_field[XMAX] == BC_PERIODIC ||
                          h->g.bc_particle[XMIN] == BC_PERIODIC || h->g.bc_particle[XMAX] == BC_PERIODIC;
  t.ntx = (h->g.nx + tx - 1) / tx + (periodic_x ? 0 : 1);
  t.ntr = (h->g.ny + tr - 1) / tr;
  t.ntiles = t.ntx * t.ntr;
  if ((long)t.ntx * t.ntr >= (long)KEY_DEFERRED) return fail(CYLPIC_EINVAL, "too many tiles");
  // the upload's sort keys: 4 x the tile-major cell number + half-cell quadrant
  if ((long)t.ntiles * tx * tr * 4 >= (1L << 31)) return fail(CYLPIC_EINVAL, "grid too large for 32-bit cell keys");
  size_t lds = h->prec == CYLPIC_MIXED ? push_lds_bytes_mixed(h->g, t, h->acc_mode, h->gather_lds)
                                      : push_lds_bytes(h->g, t, h->gather_lds);
  // the device's LDS per workgroup, less the push kernels' 64 bytes of static LDS
  const size_t limit = (size_t)h->lds_max - 64;
  if (lds > limit)
    return fail(CYLPIC_EINVAL, "tile too large: the push needs " + std::to_string(lds) + " bytes of LDS (limit " +
                                   std::to_string(limit) + ")");
  h->tl = t;
  return CYLPIC_OK;
}

int ensure_sort_scratch(cylpic_ctx* h, long n) {
  if (n <= h->sc_cap) return CYLPIC_OK;
  if (h->sc_mem) (void)hipFree(h->sc_mem);
  h->sc_mem = nullptr;
  const size_t words = sort_scratch_words(n);
  HIPCHK(hipMalloc((void**)&h->sc_mem, words * sizeof(uint32_t)));
  uint32_t* p = h->sc_mem;
  h->sc.k[0] = p;
  p += n;
  h->sc.k[1] = p;
  p += n;
  h->sc.v[0] = p;
  p += n;
  h->sc.v[1] = p;
  p += n;
  h->sc.counts = p;
  const long nb = (n + 4095) / 4096;
  p += nb * 256;
  h->sc.scan_tmp = p;
  h->sc.cap = n;
  h->sc_cap = n;
  return CYLPIC_OK;
}

// FP64 / mixed particle views
// the species' current arrays from particle `off` on
PartMixed viewm(Species& s, long off = 0) {
  PartMixed p;
  p.cell = (int*)s.ma[0] + off;
  p.ox = (float*)s.ma[1] + off;
  p.y = (float*)s.ma[2] + off;
  p.z = (float*)s.ma[3] + off;
  p.ux = (float*)s.ma[4] + off;
  p.uy = (float*)s.ma[5] + off;
  p.uz = (float*)s.ma[6] + off;
  p.w = (float*)s.ma[7] + off;
  p.key = (int*)s.key[0] + off;
  return p;
}

PartF64 view64(Species& s, long off = 0) {
  PartF64 p;
  p.x = s.a[0] + off;
  p.y = s.a[1] + off;
  p.z = s.a[2] + off;
  p.px = s.a[3] + off;
  p.py = s.a[4] + off;
  p.pz = s.a[5] + off;
  p.w = s.a[6] + off;
  p.key = (int*)s.key[0] + off;
  return p;
}

// Host particles (SI, FP64) into the species' arrays from index base on.
int stage_in(cylpic_ctx* h, Species& s, long base, long n, const double* const src[7]) {
  if (n <= 0) return CYLPIC_OK;
  if (h->prec == CYLPIC_FP64) {
    for (int k = 0; k < 7; ++k)
      HIPCHK(hipMemcpyAsync(s.a[k] + base, src[k], n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    return CYLPIC_OK;
  }
  // the mixed push forms 1 + |u|^2 in FP32 (u = p / (m c)): each component
  // must stay below 1e18 so that the sum cannot overflow
  const double umax = 1.0e18 * s.m * C_LIGHT;
  for (int k = 3; k < 6; ++k)
    for (long i = 0; i < n; ++i)
      if (!(std::fabs(src[k][i]) < umax))
        return fail(CYLPIC_EINVAL, "momentum beyond the mixed path's range (|p| / (m c) must be below 1e18 and finite)");
  if (7 * n > h->stage_cap) {
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->stage) HIPCHK(hipFree(h->stage));
    h->stage = nullptr;
    HIPCHK(hipMalloc((void**)&h->stage, 7 * n * sizeof(double)));
    h->stage_cap = 7 * n;
  }
  double* t7[7];
  for (int k = 0; k < 7; ++k) {
    t7[k] = h->stage + k * n;
    HIPCHK(hipMemcpyAsync(t7[k], src[k], n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  }
  launch_to_mixed(h->g, s.m, t7, viewm(s, base), n, h->stream);
  return CYLPIC_OK;
}

ExtraRange extra_of(const Species& s) {
  return ExtraRange{s.n_extra > 0 ? s.extra_start : nullptr, (int)(s.n - s.n_extra)};
}

// Sort of a species whose keys were tile-sorted before the last push, decided
// on the device: the neighbour migration sort (counts, offsets, one scatter)
// and the radix sort are queued together, guarded by the species' migration
// flag, so that exactly one of them runs (the radix sort only when some
// particle moved beyond the 3 x 3 tile neighbourhood), and the host waits
// for nothing.  Both leave the sorted payload in the alternate buffers and
// the ranges in tile_start_alt, which are then swapped in.  The radix sort
// runs over the host's upper bound of the count, the slots a migration sort
// left behind its kept particles first marked removed; the kept count stays
// on the device (tile_start[ntiles]) until refresh_count reads it.
int sort_species_device(cylpic_ctx* h, Species& s, bool counted) {
  const MigCount mc{s.mig_counts, s.mig_counts_x, s.mig_flag};
  if (!counted) {
    launch_zero_words(s.mig_flag, 2, h->stream);
    launch_mig_count(s.key[0], s.tile_start, extra_of(s), h->tl, mc, h->stream);
  }
  const Guard mig{s.mig_flag, 0}, rad{s.mig_flag, 1};
  const bool f64 = h->prec == CYLPIC_FP64;
  const int narr = f64 ? 7 : 8;
  const void* src[8];
  void* dst[8];
  for (int k = 0; k < narr; ++k) {
    src[k] = f64 ? (const void*)s.a[k] : (const void*)s.ma[k];
    dst[k] = f64 ? (void*)s.b[k] : (void*)s.mb[k];
  }
  const uint64_t* id_in = s.ids ? (const uint64_t*)s.id[0] : nullptr;
  uint64_t* id_out = s.ids ? (uint64_t*)s.id[1] : nullptr;
  launch_mig_scatter(s.key[0], s.tile_start, extra_of(s), h->tl, mc, h->mig, s.tile_start_alt, f64 ? 8 : 4, src,
                     dst, narr, s.key[1], id_in, id_out, h->stream, mig);
  launch_guard_count(h->d_sortstat + 1, h->stream, mig);
  // the fallback
  uint32_t* perm = nullptr;
  launch_mark_stale(s.key[0], s.tile_start, h->tl.ntiles, s.n - s.n_extra, h->stream, rad);
  radix_sort(s.key[0], s.n, h->tl.ntiles, h->slot_bits, h->sc, s.key[1], &perm, h->stream, rad);
  launch_tile_bounds(s.key[1], s.n, h->tl.ntiles, s.tile_start_alt, h->stream, rad);
  if (f64) launch_gather_f64(perm, s.n, s.a, s.b, 7, h->stream, rad);
  else launch_gather_32(perm, s.n, s.ma, s.mb, 8, h->stream, rad);
  if (s.ids) launch_gather_64(perm, s.n, id_in, id_out, h->stream, rad);
  launch_guard_count(h->d_sortstat + 0, h->stream, rad);
  HIPCHK(hipGetLastError());
  for (int k = 0; k < narr; ++k) {
    if (f64) std::swap(s.a[k], s.b[k]);
    else std::swap(s.ma[k], s.mb[k]);
  }
  std::swap(s.key[0], s.key[1]);
  if (s.ids) std::swap(s.id[0], s.id[1]);
  std::swap(s.tile_start, s.tile_start_alt);
  s.d_count = s.tile_start + h->tl.ntiles;
  s.n_extra = 0;
  s.n_exact = false;
  return CYLPIC_OK;
}

// Stable sort of a species by tile key; removed particles (key ntiles) go
// last.  With `interleave` (on upload) the keys are tile-major cell keys with
// the half-cell quadrant below them: the sort groups each tile's particles by
// cell and within a cell by quadrant, and the gather then lays each tile out
// in rounds (the r-th particle of every cell, cell after cell), so that
// consecutive particles sit in consecutive cells.  A wavefront then holds
// particles of 64 cells in mostly the same quadrant, whose deposit stencils
// start on distinct cells (no two lanes adding to one LDS word).  Later
// stable sorts by tile keep that order while particles drift; the cell sort
// method (sort_cells) re-keys and re-interleaves at every sort.
int sort_species(cylpic_ctx* h, Species& s, int interleave = 0) {
  Scope sc(h, "sort");
  s.dirty = false;
  const bool counted = s.counted;
  s.counted = false;  // the counts describe the keys of the tile ranges before this sort
  if (s.n <= 0) {
    launch_tile_bounds(nullptr, 0, h->tl.ntiles, s.tile_start, h->stream);
    s.ranged = true;
    return CYLPIC_OK;
  }
  int rc = ensure_sort_scratch(h, s.n);
  if (rc) return rc;
  if (s.ranged && !interleave && h->migrate) {
    if (!h->mig) HIPCHK(hipMalloc((void**)&h->mig, mig_scratch_words(h->tl.ntiles) * sizeof(uint32_t)));
    return sort_species_device(h, s, counted);
  }
  h->sort_radix += 1;
  // slots a device-decided sort left behind the kept particles sort last
  if (!s.n_exact) launch_mark_stale(s.key[0], s.tile_start, h->tl.ntiles, s.n - s.n_extra, h->stream);
  if (!interleave && h->sort_cells && s.ranged) {
    // tile keys -> cell keys x 4 + quadrant; removed particles (tile key
    // ntiles) -> ntiles * cells * 4
    if (h->prec == CYLPIC_FP64) launch_keys_f64(h->g, h->tl, view64(s), s.n, 2, h->stream);
    else launch_keys_mixed(h->g, h->tl, viewm(s), s.n, 2, h->stream);
    interleave = 1;
  }
  uint32_t* perm = nullptr;
  // with `interleave` the keys are 4 x tile-major cell keys + half-cell
  // quadrant (upload): each cell's particles grouped by quadrant
  const int cells = interleave ? h->tl.tx * h->tl.tr : 0;
  radix_sort(s.key[0], s.n, interleave ? h->tl.ntiles * cells * 4 : h->tl.ntiles, h->slot_bits, h->sc, s.key[1],
             &perm, h->stream);
  std::swap(s.key[0], s.key[1]);
  if (interleave) {
    // the sources in rounds layout go to a radix key buffer, free after the sort
    launch_round_map(s.key[0], perm, s.n, h->tl.ntiles, cells, h->sc.k[0], h->stream);
    perm = h->sc.k[0];
    launch_div_keys(s.key[0], s.n, (uint32_t)(cells * 4), h->stream);
  }
  launch_tile_bounds(s.key[0], s.n, h->tl.ntiles, s.tile_start, h->stream);
  if (h->prec == CYLPIC_FP64) {
    launch_gather_f64(perm, s.n, s.a, s.b, 7, h->stream);
    for (int k = 0; k < 7; ++k) std::swap(s.a[k], s.b[k]);
  } else {
    launch_gather_32(perm, s.n, s.ma, s.mb, 8, h->stream);
    for (int k = 0; k < 8; ++k) std::swap(s.ma[k], s.mb[k]);
  }
  if (s.ids) {
    launch_gather_64(perm, s.n, (const uint64_t*)s.id[0], (uint64_t*)s.id[1], h->stream);
    std::swap(s.id[0], s.id[1]);
  }
  HIPCHK(hipGetLastError());
  s.ranged = true;
  s.n_extra = 0;
  return CYLPIC_OK;
}

// read back the alive count after a sort (synchronising)
int refresh_count(cylpic_ctx* h, Species& s) {
  if (!s.tile_start) return CYLPIC_OK;
  int c = 0;
  HIPCHK(hipMemcpyAsync(&c, s.tile_start + h->tl.ntiles, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  s.n = c;
  s.n_exact = true;
  return CYLPIC_OK;
}

int check_device_errors(cylpic_ctx* h) {
  int e = 0;
  HIPCHK(hipMemcpyAsync(&e, h->d_err, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (e) {
    HIPCHK(hipMemsetAsync(h->d_err, 0, sizeof(int), h->stream));
    std::string m = "device reported:";
    if (e & ERR_CELL_RANGE) m += " a particle outside the grid's guard cells (skipped);";
    if (e & ERR_DCELL) m += " a particle crossing more than one cell in a step (CFL violated, deposit skipped);";
    return fail(CYLPIC_EPHYS, m);
  }
  return CYLPIC_OK;
}

void free_all(cylpic_ctx* h) {
  for (int k = 0; k < 15; ++k) {
    if (h->f64.f[k]) (void)hipFree(h->f64.f[k]);
    if (h->f32.f[k]) (void)hipFree(h->f32.f[k]);
  }
  for (int k = 0; k < 6; ++k) {
    if (h->f64.bnd_xmin[k]) (void)hipFree(h->f64.bnd_xmin[k]);
    if (h->f64.bnd_xmax[k]) (void)hipFree(h->f64.bnd_xmax[k]);
    if (h->f32.bnd_xmin[k]) (void)hipFree(h->f32.bnd_xmin[k]);
    if (h->f32.bnd_xmax[k]) (void)hipFree(h->f32.bnd_xmax[k]);
  }
  for (auto& s : h->sp) {
    for (int k = 0; k < 7; ++k) {
      if (s.a[k]) (void)hipFree(s.a[k]);
      if (s.b[k]) (void)hipFree(s.b[k]);
    }
    for (int k = 0; k < 8; ++k) {
      if (s.ma[k]) (void)hipFree(s.ma[k]);
      if (s.mb[k]) (void)hipFree(s.mb[k]);
    }
    if (s.id[0]) (void)hipFree(s.id[0]);
    if (s.id[1]) (void)hipFree(s.id[1]);
    if (s.key[0]) (void)hipFree(s.key[0]);
    if (s.key[1]) (void)hipFree(s.key[1]);
    if (s.tile_start) (void)hipFree(s.tile_start);
    if (s.tile_start_alt) (void)hipFree(s.tile_start_alt);
    if (s.mig_counts) (void)hipFree(s.mig_counts);
    if (s.mig_counts_x) (void)hipFree(s.mig_counts_x);
    if (s.extra_start) (void)hipFree(s.extra_start);
    if (s.mig_flag) (void)hipFree(s.mig_flag);
  }
  if (h->geo) (void)hipFree(h->geo);
  if (h->geof) (void)hipFree(h->geof);
  for (int k = 0; k < 2; ++k) {
    if (h->src_dev[k]) (void)hipFree(h->src_dev[k]);
    if (h->src_host[k]) (void)hipHostFree(h->src_host[k]);
    for (int slot = 0; slot < 2; ++slot)
      if (h->src_ev[k][slot]) (void)hipEventDestroy(h->src_ev[k][slot]);
  }
  if (h->d_err) (void)hipFree(h->d_err);
  if (h->mig) (void)hipFree(h->mig);
  if (h->stage) (void)hipFree(h->stage);
  if (h->spare64) (void)hipFree(h->spare64);
  if (h->spare32) (void)hipFree(h->spare32);
  if (h->diag64) (void)hipFree(h->diag64);
  if (h->diag_part) (void)hipFree(h->diag_part);
  if (h->diag_hist) (void)hipFree(h->diag_hist);
  if (h->cont_q) (void)hipFree(h->cont_q);
  if (h->cm_planes) (void)hipFree(h->cm_planes);
  if (h->cm_out) (void)hipFree(h->cm_out);
  if (h->gauss_q) (void)hipFree(h->gauss_q);
  if (h->cont_out) (void)hipFree(h->cont_out);
  if (h->d_sortstat) (void)hipFree(h->d_sortstat);
  if (h->d_energy) (void)hipFree(h->d_energy);
  if (h->sc_mem) (void)hipFree(h->sc_mem);
  for (auto& kv : h->prof)
    for (auto& ev : kv.second.pending) {
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
  for (auto e : h->event_pool) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
}

#define CHECK_H(h) \
  if (!(h)) return fail(CYLPIC_EINVAL, "null handle")

int alloc_species_arrays(cylpic_ctx* h, Species& s) {
  const long cap = s.cap > 0 ? s.cap : 1;
  if (h->prec == CYLPIC_FP64) {
    for (int k = 0; k < 7; ++k) {
      HIPCHK(hipMalloc((void**)&s.a[k], cap * sizeof(double)));
      HIPCHK(hipMalloc((void**)&s.b[k], cap * sizeof(double)));
    }
  } else {
    for (int k = 0; k < 8; ++k) {
      HIPCHK(hipMalloc((void**)&s.ma[k], cap * sizeof(uint32_t)));
      HIPCHK(hipMalloc((void**)&s.mb[k], cap * sizeof(uint32_t)));
    }
  }
  HIPCHK(hipMalloc((void**)&s.key[0], cap * sizeof(uint32_t)));
  HIPCHK(hipMalloc((void**)&s.key[1], cap * sizeof(uint32_t)));
  HIPCHK(hipMalloc((void**)&s.tile_start, (h->tl.ntiles + 2) * sizeof(int)));
  HIPCHK(hipMemsetAsync(s.tile_start, 0, (h->tl.ntiles + 2) * sizeof(int), h->stream));
  HIPCHK(hipMalloc((void**)&s.tile_start_alt, (h->tl.ntiles + 2) * sizeof(int)));
  HIPCHK(hipMalloc((void**)&s.mig_counts, (size_t)h->tl.ntiles * MIG_DIRS * sizeof(uint32_t)));
  HIPCHK(hipMalloc((void**)&s.mig_counts_x, (size_t)h->tl.ntiles * MIG_DIRS * sizeof(uint32_t)));
  HIPCHK(hipMalloc((void**)&s.extra_start, (h->tl.ntiles + 2) * sizeof(int)));
  HIPCHK(hipMalloc((void**)&s.mig_flag, 2 * sizeof(int)));
  s.d_count = s.tile_start + h->tl.ntiles;
  return CYLPIC_OK;
}

// Change the tile shape with species present: each species is first
// settled under the old tiles (removed particles dropped, injected batches
// merged), then its per-tile tables are reallocated and it is re-sorted by
// tile-major cell key as on upload.
int retile(cylpic_ctx* h, int tx, int tr) {
  for (auto& s : h->sp)
    if (s.cap > 0 && s.n > 0 && (s.dirty || s.n_extra > 0)) {
      int rc = sort_species(h, s);
      if (rc) return rc;
      rc = refresh_count(h, s);
      if (rc) return rc;
    }
  int rc = set_tiles(h, tx, tr);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  if (h->mig) {
    HIPCHK(hipFree(h->mig));
    h->mig = nullptr;
  }
  for (auto& s : h->sp) {
    if (s.cap <= 0) continue;
    int* bufs[3] = {s.tile_start, s.tile_start_alt, s.extra_start};
    for (int* b : bufs)
      if (b) HIPCHK(hipFree(b));
    if (s.mig_counts) HIPCHK(hipFree(s.mig_counts));
    if (s.mig_counts_x) HIPCHK(hipFree(s.mig_counts_x));
    s.tile_start = s.tile_start_alt = s.extra_start = nullptr;
    s.mig_counts = s.mig_counts_x = nullptr;
    const int nt = h->tl.ntiles;
    HIPCHK(hipMalloc((void**)&s.tile_start, (nt + 2) * sizeof(int)));
    HIPCHK(hipMemsetAsync(s.tile_start, 0, (nt + 2) * sizeof(int), h->stream));
    HIPCHK(hipMalloc((void**)&s.tile_start_alt, (nt + 2) * sizeof(int)));
    HIPCHK(hipMalloc((void**)&s.mig_counts, (size_t)nt * MIG_DIRS * sizeof(uint32_t)));
    HIPCHK(hipMalloc((void**)&s.mig_counts_x, (size_t)nt * MIG_DIRS * sizeof(uint32_t)));
    HIPCHK(hipMalloc((void**)&s.extra_start, (nt + 2) * sizeof(int)));
    s.d_count = s.tile_start + nt;
    s.n_extra = 0;
    s.ranged = false;
    s.counted = false;
    s.dirty = false;
    if (s.n <= 0) {
      launch_tile_bounds(nullptr, 0, nt, s.tile_start, h->stream);
      s.ranged = true;
      continue;
    }
    if (h->prec == CYLPIC_FP64) launch_keys_f64(h->g, h->tl, view64(s), s.n, 1, h->stream);
    else launch_keys_mixed(h->g, h->tl, viewm(s), s.n, 1, h->stream);
    rc = sort_species(h, s, 1);
    if (rc) return rc;
  }
  HIPCHK(hipGetLastError());
  return CYLPIC_OK;
}


// update_eb_fields_final (fields.f90:341-353) with the laser sources
// already on the device (null: no laser on that side)
int fields_final(cylpic_ctx* h, const double* dsrc_xmin, const double* dsrc_xmax) {
  {
    Scope s(h, "update_b");
    with_fields(h, [&](auto& F) { launch_update_b(h->g, F, 0, h->stream); });
  }
  {
    Scope s(h, "field_bcs");
    with_fields(h, [&](auto& F) { launch_bfield_final_bcs(h->g, F, dsrc_xmin, dsrc_xmax, h->stream); });
  }
  {
    Scope s(h, "update_e");
    with_fields(h, [&](auto& F) { launch_update_e(h->g, F, h->stream); });
  }
  {
    Scope s(h, "field_bcs");
    with_fields(h, [&](auto& F) { launch_efield_bcs(h->g, F, h->stream); });
  }
  HIPCHK(hipGetLastError());
  return CYLPIC_OK;
}

// push_particles (particles.F90:159-731) in three parts, so that a batch can
// put the pushes of many runs into one launch between them:
//   push_begin   jxm_old = jxm; jxm = 0 (particles.F90:163-169): the arrays
//                swap roles and the new J is cleared;
//   push_species the push and deposit of one species, launched, or (desc
//                non-null, mixed path) written as a batched-push descriptor;
//   push_end     current_bcs_r_min_final and the scheduled sort.
int push_begin(cylpic_ctx* h) {
  for (int k = 0; k < 3; ++k) swap_fields(h, 6 + k, 12 + k);
  Scope s(h, "zero_j");
  with_fields(h, [&](auto& F) { launch_zero_j(F, nfield(h), h->stream); });
  return CYLPIC_OK;
}

PushConst push_const(const cylpic_ctx* h, const Species& s) {
  PushConst pc;
  const double dt = h->g.dt;
  pc.part_q = s.q;
  pc.part_mc = C_LIGHT * s.m;
  pc.ipart_mc = 1.0 / pc.part_mc;
  const double dtfac = (0.5 * dt) * 0.25;
  pc.cmratio = (pc.part_q * dtfac) * pc.ipart_mc;
  pc.ccmratio = C_LIGHT * pc.cmratio;
  pc.idx = 1.0 / h->g.dx;
  pc.idy = 1.0 / h->g.dy;
  pc.idt = 1.0 / dt;
  pc.dtco2 = C_LIGHT * (dt / 2.0);
  return pc;
}

int push_species(cylpic_ctx* h, Species& s, void* desc) {
  const PushConst pc = push_const(h, s);
  // the push counts its tiles' new keys by direction for the next
  // migration sort (the tile ranges are those of the last sort)
  MigCount migc{nullptr, nullptr, nullptr};
  if (h->migrate && s.ranged) {
    launch_zero_words(s.mig_flag, 2, h->stream);
    migc = MigCount{s.mig_counts, s.mig_counts_x, s.mig_flag};
  }
  s.dirty = true;
  s.counted = migc.counts != nullptr;
  if (desc) {
    fill_push_desc_mixed(desc, h->g, h->tl, pc, s.q, viewm(s), s.tile_start, extra_of(s), h->f32, h->geof,
                         h->ntab, h->d_err, migc);
    return CYLPIC_OK;
  }
  Scope sc(h, "push_deposit");
  if (h->prec == CYLPIC_FP64) {
    launch_push_deposit_f64(h->g, h->tl, pc, view64(s), s.tile_start, extra_of(s), h->f64, geo_of(h),
                            h->gather_lds, h->d_err, migc, h->stream);
  } else if (launch_push_deposit_mixed(h->g, h->tl, pc, s.q, viewm(s), s.tile_start, extra_of(s), h->f32,
                                       h->geof, h->ntab, h->acc_mode, h->gather_lds, h->d_err, migc, h->stream)) {
    return fail(CYLPIC_EINVAL, "mixed push: unsupported mode count");
  }
  return CYLPIC_OK;
}

int push_end(cylpic_ctx* h) {
  {
    Scope s(h, "current_axis");
    with_fields(h, [&](auto& F) { launch_current_axis(h->g, F, h->stream); });
  }
  HIPCHK(hipGetLastError());
  h->push_count += 1;
  if (h->push_count % h->sort_interval == 0) {
    for (auto& s : h->sp) {
      if (s.immobile || s.n <= 0) continue;
      int rc = sort_species(h, s);
      if (rc) return rc;
    }
  }
  return CYLPIC_OK;
}

}  // namespace

// ----------------------------------------------------------------------
// Scratch of cylpic_gauss_residual_m0_noalloc: two m = 0 charge planes of at
// most NOALLOC_MAX_PLANE cells, then the three result words.  16 MiB held by
// every device the library runs on until the process ends.  Every call on a
// device uses its one copy, so such calls that overlap on one device (other
// threads, other handles' streams) race.  Host code uses the per-device
// address from hipGetSymbolAddress, never the symbol itself.
// ----------------------------------------------------------------------
constexpr long NOALLOC_MAX_PLANE = 1L << 20;
static __device__ __align__(256) double scratch[2 * NOALLOC_MAX_PLANE + 4];

namespace {
constexpr int SCRATCH_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[SCRATCH_DEVICES];

// The scratch array's address on `device`, resolved once per device with
// that device current (cylpic_init does it first).
int scratch_address(int device, double** out) {
  if (device < 0 || device >= SCRATCH_DEVICES)
    return fail(CYLPIC_EINVAL, "device index beyond the no-allocation scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(CYLPIC_EHIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (double*)p;
  }
  *out = g_scratch_ptr[device];
  return CYLPIC_OK;
}
}  // namespace

// ======================================================================
// C ABI
// ======================================================================
extern "C" {

const char* cylpic_last_error(void) { return g_last_error.c_str(); }

int cylpic_init(cylpic_ctx** out, int nx, int ny, int n_mode, int ng, double dx, double dy, double dt,
                double x_min, double x_max, double y_min, double y_max, const int bc_field[4],
                const int bc_particle[4], int precision, int device) {
  if (!out) return fail(CYLPIC_EINVAL, "null output handle");
  *out = nullptr;
  if (ng != 5) return fail(CYLPIC_EINVAL, "ng must be 5 (triangle shape: png = 3, ng = png + 2)");
  if (nx < 2 * ng) return fail(CYLPIC_EINVAL, "nx must be >= 2*ng");
  if (ny < 3) return fail(CYLPIC_EINVAL, "ny must be >= 3");
  // the r_max current reflection (boundary.F90:986-1013) scales ghost row i
  // by radius ratios whose denominator r_max - i dy is zero at i = ny
  if (bc_particle && bc_particle[YMAX] == BC_REFLECT && ny <= ng)
    return fail(CYLPIC_EINVAL, "a reflecting r_max needs ny > ng");
  if (n_mode < 1 || n_mode > 8) return fail(CYLPIC_EINVAL, "n_mode must be in 1..8");
  if (!(dx > 0) || !(dy > 0) || !(dt > 0)) return fail(CYLPIC_EINVAL, "dx, dy, dt must be positive");
  if (precision != CYLPIC_FP64 && precision != CYLPIC_MIXED)
    return fail(CYLPIC_EINVAL, "precision must be CYLPIC_FP64 or CYLPIC_MIXED");
  if (precision == CYLPIC_MIXED && n_mode > 4) return fail(CYLPIC_EINVAL, "CYLPIC_MIXED supports n_mode <= 4");
  if (!bc_field || !bc_particle) return fail(CYLPIC_EINVAL, "null boundary arrays");
  // CPML grows the grid by cpml_thickness cells per side and changes the
  // particle extents and reflection at that edge (mpi_routines.F90:280-290,
  // boundary.F90:1561-1569, 1756-1766); none of that is built here
  for (int i = 0; i < 4; ++i)
    if (bc_field[i] == BC_CPML_LASER || bc_field[i] == BC_CPML_OUTFLOW)
      return fail(CYLPIC_EINVAL, "CPML field boundaries (cpml_laser, cpml_outflow) are not supported");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(CYLPIC_EHIP, "no HIP device");
  if (device < 0 || device >= ndev) return fail(CYLPIC_EINVAL, "device index out of range");

  cylpic_ctx* h = new (std::nothrow) cylpic_ctx();
  if (!h) return fail(CYLPIC_ENOMEM, "host allocation failed");
  h->device = device;
  h->prec = precision;
  Grid& g = h->g;
  g.nx = nx;
  g.ny = ny;
  g.M = n_mode;
  g.ng = ng;
  g.nx2 = nx + 2 * ng;
  g.ny2 = ny + 2 * ng;
  g.plane = (long)g.nx2 * g.ny2;
  g.dx = dx;
  g.dy = dy;
  g.dt = dt;
  g.x_min = x_min;
  g.x_max = x_max;
  g.y_min = y_min;
  g.y_max = y_max;
  g.x_grid_min = x_min + dx / 2.0;  // setup.F90:177-180
  g.y_grid_min = y_min + dy / 2.0;
  for (int i = 0; i < 4; ++i) {
    g.bc_field[i] = map_field_bc(bc_field[i]);
    g.bc_particle[i] = map_particle_bc(bc_particle[i]);
  }
  for (int i = 0; i < 4; ++i) {
    if (i == YMIN) continue;
    int b = g.bc_particle[i];
    if (b == BC_THERMAL || b == BC_HEAT_BATH || b == BC_MIXED) {
      delete h;
      return fail(CYLPIC_EINVAL, "thermal / heat-bath / mixed particle boundaries are not supported");
    }
  }
  if (g.bc_field[YMAX] == BC_PERIODIC || g.bc_particle[YMAX] == BC_PERIODIC) {
    delete h;
    return fail(CYLPIC_EINVAL, "r_max cannot be periodic");
  }
  if ((g.bc_particle[XMIN] == BC_PERIODIC) != (g.bc_particle[XMAX] == BC_PERIODIC)) {
    delete h;
    return fail(CYLPIC_EINVAL, "periodic particle boundaries must be set on both x ends");
  }

  auto bail = [&](int rc) {
    free_all(h);
    delete h;
    return rc;
  };
  if (hipSetDevice(device) != hipSuccess) return bail(fail(CYLPIC_EHIP, "hipSetDevice failed"));
  {
    double* unused = nullptr;
    int rc0 = scratch_address(device, &unused);
    if (rc0) return bail(rc0);
  }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(CYLPIC_EHIP, "stream creation failed"));
  if (hipDeviceGetAttribute(&h->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
    return bail(fail(CYLPIC_EHIP, "LDS size query failed"));
  int rc = alloc_fields(h);
  if (rc) return bail(rc);
  {
    // measured defaults: the mixed path gathers from LDS-staged fields, the
    // VALU-bound FP64 path from L1/L2
    h->gather_lds = precision == CYLPIC_MIXED ? 1 : 0;
    // default tile: the first of these whose LDS needs fit (the fastest
    // measured first on the 8192 x 512 x 3 wakefield: 16 x 4 for both paths;
    // FP64 16 x 8 next, 1.7 % slower)
    const int cand_m[][2] = {{16, 4}, {8, 8}, {8, 4}, {4, 4}, {4, 2}, {2, 2}, {1, 1}, {1, 1}, {1, 1}};
    const int cand_d[][2] = {{16, 4}, {16, 8}, {8, 8}, {8, 4}, {4, 4}, {4, 2}, {2, 2}, {1, 1}, {1, 1}};
    rc = CYLPIC_EINVAL;
    for (auto& c : (precision == CYLPIC_MIXED ? cand_m : cand_d)) {
      rc = set_tiles(h, std::min(c[0], nx), std::min(c[1], ny));
      if (rc == CYLPIC_OK) break;
    }
    if (rc) return bail(rc);
  }
  std::vector<double> t;
  geometry(g, t);
  h->ntab = ny + 2 * ng + 1;
  if (hipMalloc((void**)&h->geo, t.size() * sizeof(double)) != hipSuccess)
    return bail(fail(CYLPIC_ENOMEM, "geometry allocation failed"));
  if (hipMemcpy(h->geo, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(CYLPIC_EHIP, "geometry upload failed"));
  {
    std::vector<float> tf(t.begin(), t.end());
    if (hipMalloc((void**)&h->geof, tf.size() * sizeof(float)) != hipSuccess ||
        hipMemcpy(h->geof, tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(CYLPIC_ENOMEM, "geometry upload failed"));
  }
  for (int k = 0; k < 2; ++k) {
    if (hipMalloc((void**)&h->src_dev[k], 2 * (ny + 1) * sizeof(double)) != hipSuccess)
      return bail(fail(CYLPIC_ENOMEM, "source allocation failed"));
    if (hipHostMalloc((void**)&h->src_host[k], 2 * 2 * (ny + 1) * sizeof(double), 0) != hipSuccess)
      return bail(fail(CYLPIC_ENOMEM, "pinned source allocation failed"));
    for (int slot = 0; slot < 2; ++slot)
      if (hipEventCreateWithFlags(&h->src_ev[k][slot], hipEventDisableTiming) != hipSuccess)
        return bail(fail(CYLPIC_EHIP, "event creation failed"));
  }
  if (hipMalloc((void**)&h->d_err, sizeof(int)) != hipSuccess ||
      hipMalloc((void**)&h->d_energy, 2 * n_mode * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&h->d_sortstat, 2 * sizeof(int)) != hipSuccess)
    return bail(fail(CYLPIC_ENOMEM, "allocation failed"));
  if (hipMemset(h->d_err, 0, sizeof(int)) != hipSuccess || hipMemset(h->d_sortstat, 0, 2 * sizeof(int)) != hipSuccess)
    return bail(fail(CYLPIC_EHIP, "memset failed"));
  if (hipStreamSynchronize(h->stream) != hipSuccess) return bail(fail(CYLPIC_EHIP, "init sync failed"));
  *out = h;
  return CYLPIC_OK;
}

int cylpic_finalize(cylpic_ctx* h) {
  CHECK_H(h);
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_all(h);
  delete h;
  return CYLPIC_OK;
}

int cylpic_set_dt(cylpic_ctx* h, double dt) {
  CHECK_H(h);
  if (!(dt > 0)) return fail(CYLPIC_EINVAL, "dt must be positive");
  h->g.dt = dt;
  return CYLPIC_OK;
}

int cylpic_set_time(cylpic_ctx* h, double time) {
  CHECK_H(h);
  h->time = time;
  return CYLPIC_OK;
}

int cylpic_get_restart_state(cylpic_ctx* h, double* time, double* x_min, double* x_grid_min, int64_t* push_count,
                             int* tile_shift) {
  CHECK_H(h);
  if (time) *time = h->time;
  if (x_min) *x_min = h->g.x_min;
  if (x_grid_min) *x_grid_min = h->g.x_grid_min;
  if (push_count) *push_count = h->push_count;
  if (tile_shift) *tile_shift = h->tl.shift;
  return CYLPIC_OK;
}

int cylpic_set_restart_state(cylpic_ctx* h, double time, double x_min, double x_grid_min, int64_t push_count,
                             int tile_shift) {
  CHECK_H(h);
  for (auto& s : h->sp)
    if (s.n > 0) return fail(CYLPIC_ESTATE, "set the restart state before uploading particles");
  if (push_count < 0) return fail(CYLPIC_EINVAL, "push_count must be >= 0");
  const bool periodic_x = h->g.bc_field[XMIN] == BC_PERIODIC || h->g.bc_field[XMAX] == BC_PERIODIC ||
                          h->g.bc_particle[XMIN] == BC_PERIODIC || h->g.bc_particle[XMAX] == BC_PERIODIC;
  if (tile_shift < 0 || tile_shift >= h->tl.tx || (periodic_x && tile_shift != 0))
    return fail(CYLPIC_EINVAL, "tile_shift must be in [0, tile_x), and 0 for periodic x");
  if (!(std::isfinite(x_min) && std::isfinite(x_grid_min) && std::isfinite(time)))
    return fail(CYLPIC_EINVAL, "non-finite restart state");
  h->time = time;
  // the window's position as shift_window leaves it (window.F90:70-85)
  h->g.x_min = x_min;
  h->g.x_grid_min = x_grid_min;
  h->g.x_max = x_min + (double)h->g.nx * h->g.dx;
  h->push_count = push_count;
  h->tl.shift = tile_shift;
  return CYLPIC_OK;
}

int cylpic_set_tile(cylpic_ctx* h, int tile_x, int tile_r) {
  CHECK_H(h);
  for (auto& s : h->sp)
    if (s.cap > 0) return fail(CYLPIC_ESTATE, "set the tile shape before registering species");
  return set_tiles(h, tile_x, tile_r);
}

int cylpic_tune(cylpic_ctx* h, int acc_mode, int gather_lds) {
  CHECK_H(h);
  // both arguments are checked before any state changes; a tile shape that
  // no longer fits in LDS restores both
  if (acc_mode < 0 || acc_mode > 4) return fail(CYLPIC_EINVAL, "acc_mode must be 0..4");
  if (gather_lds < 0 || gather_lds > 2 || (gather_lds == 2 && (h->prec != CYLPIC_MIXED || acc_mode != 3)))
    return fail(CYLPIC_EINVAL, "gather_lds must be 0 or 1 (2: E from LDS, mixed path with acc_mode 3 only)");
  const int old_acc = h->acc_mode, old_g = h->gather_lds;
  h->acc_mode = acc_mode;
  h->gather_lds = gather_lds;
  int rc = set_tiles(h, h->tl.tx, h->tl.tr);
  if (rc) {
    h->acc_mode = old_acc;
    h->gather_lds = old_g;
  }
  return rc;
}

int cylpic_set_sort_method(cylpic_ctx* h, int method) {
  CHECK_H(h);
  if (method != CYLPIC_SORT_RADIX && method != CYLPIC_SORT_MIGRATE && method != CYLPIC_SORT_CELLS)
    return fail(CYLPIC_EINVAL, "sort method must be CYLPIC_SORT_RADIX, CYLPIC_SORT_MIGRATE or CYLPIC_SORT_CELLS");
  h->migrate = method == CYLPIC_SORT_MIGRATE;
  h->sort_cells = method == CYLPIC_SORT_CELLS;
  return CYLPIC_OK;
}

int cylpic_sort_counts(cylpic_ctx* h, int64_t* radix, int64_t* migrate) {
  CHECK_H(h);
  HIPCHK(hipSetDevice(h->device));
  int dev[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(dev, h->d_sortstat, sizeof(dev), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (radix) *radix = h->sort_radix + dev[0];
  if (migrate) *migrate = h->sort_migrate + dev[1];
  return CYLPIC_OK;
}

int cylpic_set_sort_interval(cylpic_ctx* h, int interval) {
  CHECK_H(h);
  if (interval < 1) return fail(CYLPIC_EINVAL, "sort interval must be >= 1");
  h->sort_interval = interval;
  return CYLPIC_OK;
}

int cylpic_set_species(cylpic_ctx* h, int id, double charge, double mass, int immobile, int64_t capacity) {
  CHECK_H(h);
  if (id < 0 || id > (int)h->sp.size()) return fail(CYLPIC_EINVAL, "species ids must be dense from 0");
  if (!(mass > 0)) return fail(CYLPIC_EINVAL, "mass must be positive");
  if (capacity < 0 || capacity > 0x7fffffffL) return fail(CYLPIC_EINVAL, "capacity out of range");
  HIPCHK(hipSetDevice(h->device));
  if (id == (int)h->sp.size()) h->sp.emplace_back();
  Species& s = h->sp[id];
  if (s.cap > 0) return fail(CYLPIC_ESTATE, "species already registered");
  s.q = charge;
  s.m = mass;
  s.immobile = immobile;
  s.cap = capacity;
  return alloc_species_arrays(h, s);
}

int cylpic_upload_particles(cylpic_ctx* h, int species, int64_t n, const double* x, const double* y,
                            const double* z, const double* px, const double* py, const double* pz,
                            const double* w, const int64_t* ids) {
  CHECK_H(h);
  if (species < 0 || species >= (int)h->sp.size()) return fail(CYLPIC_EINVAL, "unknown species");
  Species& s = h->sp[species];
  if (n < 0 || n > s.cap) return fail(CYLPIC_EINVAL, "particle count exceeds the species capacity");
  if (n > 0 && (!x || !y || !z || !px || !py || !pz || !w)) return fail(CYLPIC_EINVAL, "null particle array");
  HIPCHK(hipSetDevice(h->device));
  const double* src[7] = {x, y, z, px, py, pz, w};
  int rc = stage_in(h, s, 0, n, src);
  if (rc) return rc;
  if (ids) {
    if (!s.id[0]) {
      HIPCHK(hipMalloc((void**)&s.id[0], (s.cap > 0 ? s.cap : 1) * sizeof(int64_t)));
      HIPCHK(hipMalloc((void**)&s.id[1], (s.cap > 0 ? s.cap : 1) * sizeof(int64_t)));
    }
    if (n > 0) HIPCHK(hipMemcpyAsync(s.id[0], ids, n * sizeof(int64_t), hipMemcpyHostToDevice, h->stream));
    s.ids = true;
  } else {
    s.ids = false;
  }
  s.n = n;
  s.n_exact = true;
  s.n_extra = 0;
  s.ranged = false;
  s.counted = false;
  // tile-major cell keys and half-cell quadrants: the upload sort groups
  // each tile's particles by cell, and each cell's by quadrant
  if (h->prec == CYLPIC_FP64) launch_keys_f64(h->g, h->tl, view64(s), n, 1, h->stream);
  else launch_keys_mixed(h->g, h->tl, viewm(s), n, 1, h->stream);
  rc = sort_species(h, s, 1);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  return refresh_count(h, s);
}

int cylpic_inject_particles(cylpic_ctx* h, int species, int64_t n, const double* x, const double* y,
                            const double* z, const double* px, const double* py, const double* pz,
                            const double* w, const int64_t* ids) {
  CHECK_H(h);
  if (species < 0 || species >= (int)h->sp.size()) return fail(CYLPIC_EINVAL, "unknown species");
  Species& s = h->sp[species];
  if (n < 0) return fail(CYLPIC_EINVAL, "negative particle count");
  if (n == 0) return CYLPIC_OK;
  if (!x || !y || !z || !px || !py || !pz || !w) return fail(CYLPIC_EINVAL, "null particle array");
  if (!s.ranged) return fail(CYLPIC_ESTATE, "upload the species' particles before injecting");
  if ((ids != nullptr) != s.ids) return fail(CYLPIC_EINVAL, "ids must be given exactly when the upload gave them");
  HIPCHK(hipSetDevice(h->device));
  if (!s.n_exact) {
    int rc0 = refresh_count(h, s);
    if (rc0) return rc0;
  }
  Scope scope(h, "inject");
  // the extra region holds every batch injected since the last sort; this
  // batch joins it and the region is stable-sorted again by current key
  // (earlier batches in their current order first, removed particles last)
  const long xbase = s.n - s.n_extra;
  const long base = s.n;
  if (base + n > s.cap) return fail(CYLPIC_EINVAL, "injection exceeds the species capacity");
  const double* src[7] = {x, y, z, px, py, pz, w};
  int rc = stage_in(h, s, base, n, src);
  if (rc) return rc;
  if (ids) HIPCHK(hipMemcpyAsync(s.id[0] + base, ids, n * sizeof(int64_t), hipMemcpyHostToDevice, h->stream));
  // tile keys of the batch, a stable sort of the extra region by key in
  // place, and its tile ranges: the push treats it as each tile's extra range
  if (h->prec == CYLPIC_FP64) launch_keys_f64(h->g, h->tl, view64(s, base), n, 0, h->stream);
  else launch_keys_mixed(h->g, h->tl, viewm(s, base), n, 0, h->stream);
  const long nr = base + n - xbase;
  rc = ensure_sort_scratch(h, nr);
  if (rc) return rc;
  uint32_t* perm = nullptr;
  radix_sort(s.key[0] + xbase, nr, h->tl.ntiles, 0, h->sc, s.key[1] + xbase, &perm, h->stream);
  HIPCHK(hipMemcpyAsync(s.key[0] + xbase, s.key[1] + xbase, nr * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                        h->stream));
  const bool f64 = h->prec == CYLPIC_FP64;
  const int narr = f64 ? 7 : 8;
  const size_t wb = f64 ? sizeof(double) : sizeof(uint32_t);
  if (f64) {
    double* sa[7];
    double* sb[7];
    for (int k = 0; k < 7; ++k) {
      sa[k] = s.a[k] + xbase;
      sb[k] = s.b[k] + xbase;
    }
    launch_gather_f64(perm, nr, sa, sb, 7, h->stream);
  } else {
    uint32_t* sa[8];
    uint32_t* sb[8];
    for (int k = 0; k < 8; ++k) {
      sa[k] = s.ma[k] + xbase;
      sb[k] = s.mb[k] + xbase;
    }
    launch_gather_32(perm, nr, sa, sb, 8, h->stream);
  }
  for (int k = 0; k < narr; ++k) {
    void* dst = f64 ? (void*)(s.a[k] + xbase) : (void*)(s.ma[k] + xbase);
    const void* from = f64 ? (const void*)(s.b[k] + xbase) : (const void*)(s.mb[k] + xbase);
    HIPCHK(hipMemcpyAsync(dst, from, nr * wb, hipMemcpyDeviceToDevice, h->stream));
  }
  if (s.ids) {
    launch_gather_64(perm, nr, (const uint64_t*)s.id[0] + xbase, (uint64_t*)s.id[1] + xbase,
                     h->stream);
    HIPCHK(hipMemcpyAsync(s.id[0] + xbase, s.id[1] + xbase, nr * sizeof(int64_t), hipMemcpyDeviceToDevice,
                          h->stream));
  }
  launch_tile_bounds(s.key[0] + xbase, nr, h->tl.ntiles, s.extra_start, h->stream);
  HIPCHK(hipGetLastError());
  s.n = base + n;
  s.n_extra = nr;
  s.dirty = true;
  s.counted = false;
  return CYLPIC_OK;
}

int cylpic_move_window(cylpic_ctx* h, int ncells) {
  CHECK_H(h);
  if (ncells < 0) return fail(CYLPIC_EINVAL, "ncells must be >= 0");
  if (ncells == 0) return CYLPIC_OK;
  if (h->g.bc_field[XMIN] == BC_PERIODIC || h->g.bc_field[XMAX] == BC_PERIODIC ||
      h->g.bc_particle[XMIN] == BC_PERIODIC || h->g.bc_particle[XMAX] == BC_PERIODIC)
    return fail(CYLPIC_EINVAL, "a moving window needs non-periodic x boundaries");
  HIPCHK(hipSetDevice(h->device));
  Scope sc(h, "window");
  const long nf = nfield(h);
  for (int c = 0; c < ncells; ++c) {
    // window.F90:70-85 without CPML: x_grid_min, xb_min += dx; x_max = xb_global(nx + 1)
    h->g.x_grid_min = h->g.x_grid_min + h->g.dx;
    h->g.x_min = h->g.x_min + h->g.dx;
    h->g.x_max = h->g.x_min + (double)h->g.nx * h->g.dx;
    if (h->prec == CYLPIC_FP64) {
      if (!h->spare64) HIPCHK(hipMalloc((void**)&h->spare64, nf * sizeof(cplx<double>)));
      launch_shift_window_fields(h->g, h->f64, h->spare64, h->stream);
    } else {
      if (!h->spare32) HIPCHK(hipMalloc((void**)&h->spare32, nf * sizeof(cplx<float>)));
      launch_shift_window_fields(h->g, h->f32, h->spare32, h->stream);
    }
  }
  for (auto& s : h->sp) {
    if (s.n <= 0) continue;
    if (h->prec == CYLPIC_FP64) launch_window_f64(h->g, view64(s), s.n, h->tl.ntiles, h->stream);
    else launch_window_mixed(viewm(s), s.n, ncells, h->tl.ntiles, h->stream);
    s.dirty = true;
    s.counted = false;
  }
  // the tiles move with the plasma: raise their offset; each time it reaches
  // a whole tile, relabel the columns one down (order and ranges unchanged;
  // the vacated first column held only particles that just left the grid)
  for (int c = 0; c < ncells; ++c) {
    if (++h->tl.shift < h->tl.tx) continue;
    h->tl.shift = 0;
    for (auto& s : h->sp) {
      if (!s.tile_start) continue;
      launch_relabel_tiles(s.key[0], s.n, h->tl.ntiles, s.tile_start, s.tile_start_alt, h->stream);
      std::swap(s.tile_start, s.tile_start_alt);
      s.d_count = s.tile_start + h->tl.ntiles;
      if (s.n_extra > 0) {
        if (!h->mig) HIPCHK(hipMalloc((void**)&h->mig, mig_scratch_words(h->tl.ntiles) * sizeof(uint32_t)));
        launch_relabel_tiles(nullptr, 0, h->tl.ntiles, s.extra_start, (int*)h->mig, h->stream);
        HIPCHK(hipMemcpyAsync(s.extra_start, h->mig, (h->tl.ntiles + 2) * sizeof(int), hipMemcpyDeviceToDevice,
                              h->stream));
      }
    }
  }
  HIPCHK(hipGetLastError());
  return CYLPIC_OK;
}

int cylpic_particle_count(cylpic_ctx* h, int species, int64_t* n_out) {
  CHECK_H(h);
  if (species < 0 || species >= (int)h->sp.size() || !n_out) return fail(CYLPIC_EINVAL, "bad argument");
  HIPCHK(hipSetDevice(h->device));
  Species& s = h->sp[species];
  int rc = s.dirty ? sort_species(h, s) : CYLPIC_OK;  // drop removed particles first
  if (rc) return rc;
  rc = refresh_count(h, s);
  if (rc) return rc;
  *n_out = s.n;
  return CYLPIC_OK;
}

int cylpic_download_particles(cylpic_ctx* h, int species, int64_t capacity, int64_t* n_out, double* x,
                              double* y, double* z, double* px, double* py, double* pz, double* w,
                              int64_t* ids) {
  CHECK_H(h);
  if (species < 0 || species >= (int)h->sp.size()) return fail(CYLPIC_EINVAL, "unknown species");
  HIPCHK(hipSetDevice(h->device));
  Species& s = h->sp[species];
  int rc = s.dirty ? sort_species(h, s) : CYLPIC_OK;  // drop removed particles first
  if (rc) return rc;
  rc = refresh_count(h, s);
  if (rc) return rc;
  if (n_out) *n_out = s.n;
  if (s.n > capacity) return fail(CYLPIC_EINVAL, "output capacity smaller than the particle count");
  double* dst[7] = {x, y, z, px, py, pz, w};
  if (h->prec == CYLPIC_FP64) {
    for (int k = 0; k < 7; ++k)
      if (dst[k] && s.n > 0)
        HIPCHK(hipMemcpyAsync(dst[k], s.a[k], s.n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  } else if (s.n > 0) {
    double* tmp = nullptr;
    HIPCHK(hipMalloc((void**)&tmp, 7 * s.n * sizeof(double)));
    double* t7[7];
    for (int k = 0; k < 7; ++k) t7[k] = tmp + k * s.n;
    launch_from_mixed(h->g, s.m, viewm(s), t7, s.n, h->stream);
    for (int k = 0; k < 7; ++k)
      if (dst[k]) HIPCHK(hipMemcpyAsync(dst[k], t7[k], s.n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipFree(tmp));
  }
  if (ids && s.ids && s.n > 0)
    HIPCHK(hipMemcpyAsync(ids, s.id[0], s.n * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return check_device_errors(h);
}

int cylpic_upload_field(cylpic_ctx* h, int which, const double* f) {
  CHECK_H(h);
  if (which < 0 || which >= CYLPIC_NFIELDS || !f) return fail(CYLPIC_EINVAL, "bad field id or pointer");
  HIPCHK(hipSetDevice(h->device));
  if (h->prec == CYLPIC_FP64) {
    HIPCHK(hipMemcpyAsync(h->f64.f[which], f, nfield(h) * sizeof(cplx<double>), hipMemcpyHostToDevice, h->stream));
  } else {
    cplx<double>* tmp = nullptr;
    HIPCHK(hipMalloc((void**)&tmp, nfield(h) * sizeof(cplx<double>)));
    HIPCHK(hipMemcpyAsync(tmp, f, nfield(h) * sizeof(cplx<double>), hipMemcpyHostToDevice, h->stream));
    launch_d2f(tmp, h->f32.f[which], nfield(h), h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipFree(tmp));
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return CYLPIC_OK;
}

int cylpic_download_field(cylpic_ctx* h, int which, double* f) {
  CHECK_H(h);
  if (which < 0 || which >= CYLPIC_NFIELDS || !f) return fail(CYLPIC_EINVAL, "bad field id or pointer");
  HIPCHK(hipSetDevice(h->device));
  if (h->prec == CYLPIC_FP64) {
    HIPCHK(hipMemcpyAsync(f, h->f64.f[which], nfield(h) * sizeof(cplx<double>), hipMemcpyDeviceToHost, h->stream));
  } else {
    cplx<double>* tmp = nullptr;
    HIPCHK(hipMalloc((void**)&tmp, nfield(h) * sizeof(cplx<double>)));
    launch_f2d(h->f32.f[which], tmp, nfield(h), h->stream);
    HIPCHK(hipMemcpyAsync(f, tmp, nfield(h) * sizeof(cplx<double>), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipFree(tmp));
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return check_device_errors(h);
}

int cylpic_setup_field_boundaries(cylpic_ctx* h) {
  CHECK_H(h);
  HIPCHK(hipSetDevice(h->device));
  with_fields(h, [&](auto& F) { launch_setup_bnd(h->g, F, h->stream); });
  HIPCHK(hipGetLastError());
  return CYLPIC_OK;
}

int cylpic_update_eb_fields_half(cylpic_ctx* h) {
  CHECK_H(h);
  HIPCHK(hipSetDevice(h->device));
  {
    Scope s(h, "update_e");
    with_fields(h, [&](auto& F) { launch_update_e(h->g, F, h->stream); });
  }
  {
    Scope s(h, "field_bcs");
    with_fields(h, [&](auto& F) { launch_efield_bcs(h->g, F, h->stream); });
  }
  {
    // bxm_old = bxm (fields.f90:326-328) in the update's own pass
    Scope s(h, "update_b");
    with_fields(h, [&](auto& F) { launch_update_b(h->g, F, 1, h->stream); });
  }
  {
    Scope s(h, "field_bcs");
    with_fields(h, [&](auto& F) { launch_bfield_bcs(h->g, F, 1, h->stream); });
  }
  HIPCHK(hipGetLastError());
  return CYLPIC_OK;
}

int cylpic_push_particles(cylpic_ctx* h) {
  CHECK_H(h);
  HIPCHK(hipSetDevice(h->device));
  int rc = push_begin(h);
  for (auto& s : h->sp) {
    if (rc) return rc;
    if (s.immobile || s.n <= 0) continue;
    rc = push_species(h, s, nullptr);
  }
  if (rc) return rc;
  return push_end(h);
}

int cylpic_sort_particles(cylpic_ctx* h) {
  CHECK_H(h);
  HIPCHK(hipSetDevice(h->device));
  for (auto& s : h->sp) {
    int rc = sort_species(h, s);
    if (rc) return rc;
  }
  return CYLPIC_OK;
}

int cylpic_current_finish(cylpic_ctx* h) {
  CHECK_H(h);
  HIPCHK(hipSetDevice(h->device));
  Scope s(h, "current_finish");
  with_fields(h, [&](auto& F) { launch_current_finish(h->g, F, h->stream); });
  if (h->smooth_its + h->smooth_comp_its > 0) {
    const long nf = nfield(h);
    if (h->prec == CYLPIC_FP64) {
      if (!h->spare64) HIPCHK(hipMalloc((void**)&h->spare64, nf * sizeof(cplx<double>)));
      launch_smooth_current(h->g, h->f64, h->spare64, h->smooth_its, h->smooth_comp_its, h->smooth_strides.data(),
                            (int)h->smooth_strides.size(), h->stream);
    } else {
      if (!h->spare32) HIPCHK(hipMalloc((void**)&h->spare32, nf * sizeof(cplx<float>)));
      launch_smooth_current(h->g, h->f32, h->spare32, h->smooth_its, h->smooth_comp_its, h->smooth_strides.data(),
                            (int)h->smooth_strides.size(), h->stream);
    }
  }
  HIPCHK(hipGetLastError());
  return CYLPIC_OK;
}

int cylpic_set_smoothing(cylpic_ctx* h, int its, int comp_its, const int* strides, int nstrides) {
  CHECK_H(h);
  if (its < 0 || comp_its < 0) return fail(CYLPIC_EINVAL, "iteration counts must be >= 0");
  std::vector<int> st;
  if (strides && nstrides > 0) st.assign(strides, strides + nstrides);
  else st.assign(1, 1);
  for (int v : st)
    if (v < 1 || v > h->g.ng) return fail(CYLPIC_EINVAL, "smoothing strides must be in 1..ng");
  h->smooth_its = its;
  h->smooth_comp_its = comp_its;
  h->smooth_strides = st;
  return CYLPIC_OK;
}

int cylpic_update_eb_fields_final(cylpic_ctx* h, const double* src_xmin, const double* src_xmax) {
  CHECK_H(h);
  HIPCHK(hipSetDevice(h->device));
  const size_t sb = 2 * (size_t)(h->g.ny + 1) * sizeof(double);
  const double* dsrc[2] = {nullptr, nullptr};
  const double* hsrc[2] = {src_xmin, src_xmax};
  // stage through pinned memory in two slots used in turn: the slot written
  // now was last read by the copy two calls ago, which its event covers, so
  // the host does not wait for the stream
  const int slot = h->src_slot;
  h->src_slot ^= 1;
  for (int k = 0; k < 2; ++k) {
    if (!hsrc[k]) continue;
    double* pin = h->src_host[k] + slot * 2 * (size_t)(h->g.ny + 1);
    HIPCHK(hipEventSynchronize(h->src_ev[k][slot]));
    std::memcpy(pin, hsrc[k], sb);
    HIPCHK(hipMemcpyAsync(h->src_dev[k], pin, sb, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipEventRecord(h->src_ev[k][slot], h->stream));
    dsrc[k] = h->src_dev[k];
  }
  return fields_final(h, dsrc[0], dsrc[1]);
}

int cylpic_step(cylpic_ctx* h, const double* src_xmin, const double* src_xmax) {
  int rc = cylpic_update_eb_fields_half(h);
  if (rc) return rc;
  rc = cylpic_push_particles(h);
  if (rc) return rc;
  rc = cylpic_current_finish(h);
  if (rc) return rc;
  h->time += h->g.dt;
  return cylpic_update_eb_fields_final(h, src_xmin, src_xmax);
}

int cylpic_mode_energy(cylpic_ctx* h, double* w_e, double* w_b) {
  CHECK_H(h);
  if (!w_e || !w_b) return fail(CYLPIC_EINVAL, "null output");
  HIPCHK(hipSetDevice(h->device));
  const int M = h->g.M;
  HIPCHK(hipMemsetAsync(h->d_energy, 0, 2 * M * sizeof(double), h->stream));
  with_fields(h, [&](auto& F) { launch_energy(h->g, F, h->d_energy, h->stream); });
  std::vector<double> out(2 * M);
  HIPCHK(hipMemcpyAsync(out.data(), h->d_energy, 2 * M * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  const double mu0 = 4.0e-7 * PI;
  for (int m = 0; m < M; ++m) {
    const double kappa = m == 0 ? 2.0 * PI : PI;
    w_e[m] = kappa * 0.5 * EPSILON0 * out[2 * m] * h->g.dx * h->g.dy;
    w_b[m] = kappa * 0.5 / mu0 * out[2 * m + 1] * h->g.dx * h->g.dy;
  }
  return CYLPIC_OK;
}

// A species' alive particles at the front of its arrays (sorting first when
// a push may have left removed ones, as a download does).
static int settle_species(cylpic_ctx* h, Species& s) {
  int rc = s.dirty ? sort_species(h, s) : CYLPIC_OK;
  if (rc) return rc;
  return refresh_count(h, s);
}

int cylpic_number_density_modes(cylpic_ctx* h, int species, double* out) {
  CHECK_H(h);
  if (!out) return fail(CYLPIC_EINVAL, "null output");
  if (species >= (int)h->sp.size()) return fail(CYLPIC_EINVAL, "unknown species");
  HIPCHK(hipSetDevice(h->device));
  const long nf = nfield(h);
  if (!h->diag64) HIPCHK(hipMalloc((void**)&h->diag64, nf * sizeof(cplx<double>)));
  launch_fill_zero(h->diag64, nf, h->stream);
  const int s0 = species < 0 ? 0 : species, s1 = species < 0 ? (int)h->sp.size() - 1 : species;
  for (int is = s0; is <= s1; ++is) {
    Species& s = h->sp[is];
    int rc = settle_species(h, s);
    if (rc) return rc;
    const int* ts = s.ranged ? s.tile_start : nullptr;
    if (h->prec == CYLPIC_FP64)
      launch_number_density_f64(h->g, h->tl, view64(s), ts, extra_of(s), s.n, h->diag64, h->stream);
    else launch_number_density_mixed(h->g, h->tl, viewm(s), ts, extra_of(s), s.n, h->diag64, h->stream);
  }
  launch_density_bcs(h->g, h->diag64, h->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, h->diag64, nf * sizeof(cplx<double>), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return check_device_errors(h);
}

int cylpic_dist_fn(cylpic_ctx* h, int species, int ndims, const int* direction, double* ranges, int* resolution,
                   const double* restrictions, const int* use_restrictions, double* out, int64_t capacity,
                   int64_t* nbins) {
  CHECK_H(h);
  if (species < 0 || species >= (int)h->sp.size()) return fail(CYLPIC_EINVAL, "unknown species");
  if (ndims < 1 || ndims > DF_MAXDIMS || !direction || !ranges || !resolution || !nbins)
    return fail(CYLPIC_EINVAL, "ndims must be 1..3 with direction, ranges, resolution and nbins given");
  HIPCHK(hipSetDevice(h->device));
  *nbins = 0;
  const Grid& g = h->g;
  DistFn d{};
  d.ndims = ndims;
  bool parallel[3] = {false, false, false}, any_calc = false;
  for (int k = 0; k < ndims; ++k) {
    const int dir = direction[k];
    d.direction[k] = dir;
    double r0 = ranges[2 * k], r1 = ranges[2 * k + 1];
    if (dir == DF_X || dir == DF_Y) {  // dist_fn.F90:197-286
      const double hh = dir == DF_X ? g.dx : g.dy;
      const double lo = dir == DF_X ? g.x_min : g.y_min, hi = dir == DF_X ? g.x_max : g.y_max;
      const double gmin = dir == DF_X ? g.x_grid_min : g.y_grid_min;
      const double gmax = gmin + (double)((dir == DF_X ? g.nx : g.ny) - 1) * hh;
      if (std::fabs(r0 - r1) <= 2.2250738585072014e-308) {
        r0 = gmin - 0.5 * hh;
        r1 = gmax + 0.5 * hh;
      } else {
        r0 = std::fmax(r0, lo);
        r1 = std::fmin(r1, hi);
        r0 = lo + std::floor((r0 - lo) / hh) * hh;
        r1 = lo + std::ceil((r1 - lo) / hh) * hh;
        r0 = std::fmax(r0, gmin - 0.5 * hh);
        r1 = std::fmin(r1, gmax + 0.5 * hh);
      }
      d.resolution[k] = (int)std::max(0L, std::lround((r1 - r0) / hh));
      d.dgrid[k] = hh;
      parallel[k] = true;
    } else {
      if (dir < DF_PX || dir > DF_MOD_P) return fail(CYLPIC_EINVAL, "unknown distribution function direction");
      if (resolution[k] < 1) return fail(CYLPIC_EINVAL, "momentum-space resolution must be >= 1");
      d.resolution[k] = resolution[k];
      if (std::fabs(r0 - r1) <= 2.2250738585072014e-308) {
        d.calc_range[k] = 1;
        any_calc = true;
      }
      if (dir >= DF_XY_ANGLE && dir <= DF_ZX_ANGLE) d.need_angle[dir - DF_XY_ANGLE] = 1;
    }
    d.range[k][0] = r0;
    d.range[k][1] = r1;
  }
  for (int k = ndims; k < 3; ++k) d.resolution[k] = 1;
  for (int k = 0; k < DF_MAXDIRS; ++k) {
    d.use_restriction[k] = use_restrictions ? use_restrictions[k] : 0;
    if (d.use_restriction[k]) {
      d.restriction[k][0] = restrictions[2 * k];
      d.restriction[k][1] = restrictions[2 * k + 1];
    }
  }
  for (int k = 0; k < 3; ++k)  // an angle that is only restricted is computed too
    if (d.use_restriction[DF_XY_ANGLE - 1 + k]) d.need_angle[k] = 1;
  Species& s = h->sp[species];
  int rc = settle_species(h, s);
  if (rc) return rc;
  if (s.n < 1) return CYLPIC_OK;  // the reference writes nothing for an empty species
  const double mc = s.m * C_LIGHT;
  const int nblocks = (int)std::min<long>(1024, (s.n + 255) / 256);
  if (any_calc) {  // dist_fn.F90:341-453
    if (!h->diag_part) HIPCHK(hipMalloc((void**)&h->diag_part, 1024 * 6 * sizeof(double)));
    if (h->prec == CYLPIC_FP64) launch_df_range_f64(g, view64(s), s.n, mc, d, h->diag_part, nblocks, h->stream);
    else launch_df_range_mixed(g, viewm(s), s.n, mc, d, h->diag_part, nblocks, h->stream);
    HIPCHK(hipGetLastError());
    std::vector<double> part((size_t)nblocks * 6);
    HIPCHK(hipMemcpyAsync(part.data(), h->diag_part, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                          h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int k = 0; k < ndims; ++k)
      if (d.calc_range[k]) {
        double lo = 1.7976931348623157e308 * 0.4, hi = -lo;
        for (int b = 0; b < nblocks; ++b) {
          lo = std::fmin(lo, part[(size_t)b * 6 + 2 * k]);
          hi = std::fmax(hi, part[(size_t)b * 6 + 2 * k + 1]);
        }
        d.range[k][0] = lo;
        d.range[k][1] = hi;
      }
  }
  for (int k = 0; k < ndims; ++k) {  // dist_fn.F90:455-472
    if (std::fabs(d.range[k][0] - d.range[k][1]) <= 2.2250738585072014e-308) {
      d.range[k][0] = -1.0;
      d.range[k][1] = 1.0;
    }
    if (d.direction[k] >= DF_XY_ANGLE && d.direction[k] <= DF_ZX_ANGLE)
      d.angle_max[d.direction[k] - DF_XY_ANGLE] = d.range[k][1];
    if (!parallel[k]) d.dgrid[k] = (d.range[k][1] - d.range[k][0]) / (double)d.resolution[k];
    ranges[2 * k] = d.range[k][0];
    ranges[2 * k + 1] = d.range[k][1];
    resolution[k] = d.resolution[k];
  }
  const long nb = (long)d.resolution[0] * d.resolution[1] * d.resolution[2];
  *nbins = nb;
  if (!out || nb == 0) return CYLPIC_OK;
  if (nb > capacity) return fail(CYLPIC_EINVAL, "output capacity smaller than the number of bins");
  if (nb > h->diag_hist_cap) {
    if (h->diag_hist) HIPCHK(hipFree(h->diag_hist));
    h->diag_hist = nullptr;
    h->diag_hist_cap = 0;
    HIPCHK(hipMalloc((void**)&h->diag_hist, nb * sizeof(double)));
    h->diag_hist_cap = nb;
  }
  HIPCHK(hipMemsetAsync(h->diag_hist, 0, nb * sizeof(double), h->stream));
  if (h->prec == CYLPIC_FP64) launch_df_bin_f64(g, view64(s), s.n, mc, d, h->diag_hist, nblocks, h->stream);
  else launch_df_bin_mixed(g, viewm(s), s.n, mc, d, h->diag_hist, nblocks, h->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, h->diag_hist, nb * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return check_device_errors(h);
}

// The m = 0 charge of the mobile species at x + vs v dt/2 into q; with
// `immobile`, immobile species too, at x.  q is cleared first, by a fill
// kernel when `clear_by_kernel` (the no-allocation entry) or hipMemsetAsync.
static int charge_half(cylpic_ctx* h, double* q, bool immobile = false, double vs = 1.0,
                       bool clear_by_kernel = false) {
  if (clear_by_kernel) launch_zero_f64(q, h->g.plane, h->stream);
  else HIPCHK(hipMemsetAsync(q, 0, h->g.plane * sizeof(double), h->stream));
  for (auto& s : h->sp) {
    if ((s.immobile && !immobile) || s.n <= 0) continue;
    const int* ts = s.ranged ? s.tile_start : nullptr;
    const double v = s.immobile ? 0.0 : vs;
    if (h->prec == CYLPIC_FP64)
      launch_charge_half_f64(h->g, h->tl, view64(s), ts, extra_of(s), s.n, s.m, s.q, v, q, h->stream);
    else launch_charge_half_mixed(h->g, h->tl, viewm(s), ts, extra_of(s), s.n, s.q, v, q, h->stream);
  }
  HIPCHK(hipGetLastError());
  return CYLPIC_OK;
}

// words of one set of M complex planes of the mode check
static long cm_words(const cylpic_ctx* h) { return 2 * (long)h->g.M * h->g.plane; }

int cylpic_continuity_begin(cylpic_ctx* h) {
  CHECK_H(h);
  HIPCHK(hipSetDevice(h->device));
  if (!h->cont_q) HIPCHK(hipMalloc((void**)&h->cont_q, 2 * h->g.plane * sizeof(double)));
  if (!h->cont_out) HIPCHK(hipMalloc((void**)&h->cont_out, 3 * sizeof(unsigned long long)));
  if (!h->cm_planes) HIPCHK(hipMalloc((void**)&h->cm_planes, 4 * cm_words(h) * sizeof(double)));
  if (!h->cm_out) HIPCHK(hipMalloc((void**)&h->cm_out, 4 * h->g.M * sizeof(unsigned long long)));
  int rc = charge_half(h, h->cont_q);
  if (rc) return rc;
  // every mode: the charge at the push's half step (mobile species; an
  // immobile one adds the same charge before and after)
  launch_zero_f64(h->cm_planes, cm_words(h), h->stream);
  for (auto& s : h->sp) {
    if (s.immobile || s.n <= 0) continue;
    const int* ts = s.ranged ? s.tile_start : nullptr;
    if (h->prec == CYLPIC_FP64)
      launch_cont_modes_begin_f64(h->g, h->tl, view64(s), s.n, ts, extra_of(s), s.m, s.q, h->cm_planes, h->stream);
    else launch_cont_modes_begin_mixed(h->g, h->tl, viewm(s), s.n, ts, extra_of(s), s.q, h->cm_planes, h->stream);
  }
  HIPCHK(hipGetLastError());
  h->cont_armed = true;
  return CYLPIC_OK;
}

int cylpic_continuity_end(cylpic_ctx* h, double* residual, double* dq, double* q) {
  CHECK_H(h);
  if (!residual || !dq || !q) return fail(CYLPIC_EINVAL, "null output");
  if (!h->cont_armed) return fail(CYLPIC_ESTATE, "cylpic_continuity_begin was not called before this push");
  HIPCHK(hipSetDevice(h->device));
  h->cont_armed = false;
  double* q1 = h->cont_q + h->g.plane;
  int rc = charge_half(h, q1);
  if (rc) return rc;
  HIPCHK(hipMemsetAsync(h->cont_out, 0, 3 * sizeof(unsigned long long), h->stream));
  with_fields(h, [&](auto& F) {
    launch_continuity_m0(h->g, h->cont_q, q1, F.jx(), F.jr(), geo_of(h), h->cont_out, h->stream);
  });
  // every mode: charge at t + 3dt/2 and the azimuthal terms, then the maxima
  const long w = cm_words(h);
  double* cq0 = h->cm_planes;
  double* cq1 = cq0 + w;
  double* cz = cq1 + w;
  double* czp = cz + w;
  launch_zero_f64(cq1, 3 * w, h->stream);
  for (auto& s : h->sp) {
    if (s.immobile || s.n <= 0) continue;
    const int* ts = s.ranged ? s.tile_start : nullptr;
    // cq1, cz, czp are consecutive sets of w words
    if (h->prec == CYLPIC_FP64)
      launch_cont_modes_end_f64(h->g, h->tl, view64(s), s.n, ts, extra_of(s), s.m, s.q, cq1, w, h->stream);
    else launch_cont_modes_end_mixed(h->g, h->tl, viewm(s), s.n, ts, extra_of(s), s.q, cq1, w, h->stream);
  }
  launch_zero_words((uint32_t*)h->cm_out, 8L * h->g.M, h->stream);
  with_fields(h, [&](auto& F) {
    launch_continuity_modes(h->g, cq0, cq1, cz, czp, F.jx(), F.jr(), geo_of(h), h->cm_out, h->stream);
  });
  HIPCHK(hipGetLastError());
  unsigned long long bits[3];
  std::vector<unsigned long long> mbits(4 * h->g.M);
  HIPCHK(hipMemcpyAsync(bits, h->cont_out, sizeof(bits), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(mbits.data(), h->cm_out, mbits.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  std::memcpy(residual, &bits[0], sizeof(double));
  std::memcpy(dq, &bits[1], sizeof(double));
  std::memcpy(q, &bits[2], sizeof(double));
  h->cm_last.assign(mbits.size(), 0.0);
  std::memcpy(h->cm_last.data(), mbits.data(), mbits.size() * sizeof(double));
  return check_device_errors(h);
}

int cylpic_continuity_modes(cylpic_ctx* h, double* residual, double* dq, double* jtheta, double* z) {
  CHECK_H(h);
  if (h->cm_last.empty()) return fail(CYLPIC_ESTATE, "no continuity check has run on this handle");
  for (int m = 0; m < h->g.M; ++m) {
    if (residual) residual[m] = h->cm_last[4 * m];
    if (dq) dq[m] = h->cm_last[4 * m + 1];
    if (jtheta) jtheta[m] = h->cm_last[4 * m + 2];
    if (z) z[m] = h->cm_last[4 * m + 3];
  }
  return CYLPIC_OK;
}

int cylpic_gauss_residual_m0(cylpic_ctx* h, double* residual, double* q, double* flux) {
  CHECK_H(h);
  if (!residual || !q || !flux) return fail(CYLPIC_EINVAL, "null output");
  HIPCHK(hipSetDevice(h->device));
  if (!h->gauss_q) HIPCHK(hipMalloc((void**)&h->gauss_q, 2 * h->g.plane * sizeof(double)));
  if (!h->cont_out) HIPCHK(hipMalloc((void**)&h->cont_out, 3 * sizeof(unsigned long long)));
  // the charges half a step behind and ahead (their mean is the charge E
  // is consistent with between steps); immobile species at x in both
  double* qa = h->gauss_q;
  double* qb = h->gauss_q + h->g.plane;
  int rc = charge_half(h, qa, true, -1.0);
  if (!rc) rc = charge_half(h, qb, true, 1.0);
  if (rc) return rc;
  HIPCHK(hipMemsetAsync(h->cont_out, 0, 3 * sizeof(unsigned long long), h->stream));
  with_fields(h, [&](auto& F) {
    launch_gauss_m0(h->g, qa, qb, F.ex(), F.er(), geo_of(h), h->cont_out, h->stream);
  });
  HIPCHK(hipGetLastError());
  unsigned long long bits[3];
  HIPCHK(hipMemcpyAsync(bits, h->cont_out, sizeof(bits), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  std::memcpy(residual, &bits[0], sizeof(double));
  std::memcpy(q, &bits[1], sizeof(double));
  std::memcpy(flux, &bits[2], sizeof(double));
  return check_device_errors(h);
}

int cylpic_gauss_residual_m0_noalloc(cylpic_ctx* h, double* residual, double* q, double* flux) {
  CHECK_H(h);
  if (!residual || !q || !flux) return fail(CYLPIC_EINVAL, "null output");
  if (h->g.plane > NOALLOC_MAX_PLANE)
    return fail(CYLPIC_EINVAL, "cylpic_gauss_residual_m0_noalloc: the grid's (nx + 2 ng) (ny + 2 ng) = " +
                                   std::to_string(h->g.plane) + " cells exceed its scratch of " +
                                   std::to_string(NOALLOC_MAX_PLANE) + " cells per charge plane");
  HIPCHK(hipSetDevice(h->device));
  double* base = nullptr;
  int rc = scratch_address(h->device, &base);
  if (rc) return rc;
  double* qa = base;
  double* qb = base + NOALLOC_MAX_PLANE;
  unsigned long long* words = (unsigned long long*)(base + 2 * NOALLOC_MAX_PLANE);
  rc = charge_half(h, qa, true, -1.0, true);
  if (!rc) rc = charge_half(h, qb, true, 1.0, true);
  if (rc) return rc;
  launch_zero_f64((double*)words, 3, h->stream);
  with_fields(h, [&](auto& F) { launch_gauss_m0(h->g, qa, qb, F.ex(), F.er(), geo_of(h), words, h->stream); });
  HIPCHK(hipGetLastError());
  unsigned long long bits[3];
  HIPCHK(hipMemcpyAsync(bits, words, sizeof(bits), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  std::memcpy(residual, &bits[0], sizeof(double));
  std::memcpy(q, &bits[1], sizeof(double));
  std::memcpy(flux, &bits[2], sizeof(double));
  return check_device_errors(h);
}

int cylpic_autotune(cylpic_ctx* h, int pushes, int* tile_x, int* tile_r, int* gather_lds, double* ms_best) {
  CHECK_H(h);
  if (pushes < 1) return fail(CYLPIC_EINVAL, "pushes must be >= 1");
  HIPCHK(hipSetDevice(h->device));
  // settle every species so that the snapshot is the sorted particle set
  for (auto& s : h->sp) {
    if (s.cap <= 0 || !s.tile_start) continue;
    int rc = settle_species(h, s);
    if (rc) return rc;
  }
  // snapshot: each species' particles (and ids), the currents and their
  // *_old copies, and the step bookkeeping the pushes advance
  const bool f64 = h->prec == CYLPIC_FP64;
  const int narr = f64 ? 7 : 8;
  const size_t wb = f64 ? sizeof(double) : sizeof(uint32_t);
  size_t bytes = 6 * nfield(h) * cplx_bytes(h);
  for (auto& s : h->sp) bytes += (size_t)s.n * (narr * wb + (s.ids ? sizeof(int64_t) : 0));
  char* snap = nullptr;
  HIPCHK(hipMalloc((void**)&snap, bytes));
  std::vector<long> n0;
  auto save = [&](bool restore) -> int {
    char* p = snap;
    auto mv = [&](void* live, size_t nb) -> int {
      if (nb == 0) return CYLPIC_OK;
      HIPCHK(hipMemcpyAsync(restore ? live : (void*)p, restore ? (const void*)p : (const void*)live, nb,
                            hipMemcpyDeviceToDevice, h->stream));
      p += nb;
      return CYLPIC_OK;
    };
    for (int k = 0; k < 6; ++k) {
      int rc = mv(field_ptr(h, k < 3 ? 6 + k : 12 + (k - 3)), nfield(h) * cplx_bytes(h));
      if (rc) return rc;
    }
    for (size_t i = 0; i < h->sp.size(); ++i) {
      Species& s = h->sp[i];
      const long n = restore ? n0[i] : s.n;
      for (int k = 0; k < narr; ++k) {
        int rc = mv(f64 ? (void*)s.a[k] : (void*)s.ma[k], n * wb);
        if (rc) return rc;
      }
      if (s.ids) {
        int rc = mv(s.id[0], n * sizeof(int64_t));
        if (rc) return rc;
      }
      if (restore) {
        s.n = n;
        s.n_exact = true;
        s.n_extra = 0;
        s.dirty = false;
      }
    }
    return CYLPIC_OK;
  };
  for (auto& s : h->sp) n0.push_back(s.n);
  const long push_count0 = h->push_count, radix0 = h->sort_radix, migrate0 = h->sort_migrate;
  const int interval0 = h->sort_interval;
  h->sort_interval = 1 << 30;  // time the pushes alone
  const int tx0 = h->tl.tx, tr0 = h->tl.tr, g0 = h->gather_lds, shift0 = h->tl.shift;
  int rc = save(false);
  if (rc) {
    (void)hipFree(snap);
    return rc;
  }
  struct Cand {
    int tx, tr, g;
  };
  std::vector<Cand> cands;
  if (f64) cands = {{16, 8, 0}, {16, 4, 0}, {8, 8, 0}, {8, 4, 0}, {32, 4, 0}, {16, 8, 1}};
  else cands = {{16, 4, 1}, {8, 8, 1}, {8, 4, 1}, {16, 8, 1}, {32, 4, 1}, {16, 4, 0}, {16, 4, 2}, {32, 4, 2}};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  double best = -1.0;
  Cand pick{tx0, tr0, g0};
  for (const Cand& c : cands) {
    if (c.tx > h->g.nx || c.tr > h->g.ny) continue;
    if (c.g == 2 && h->acc_mode != 3) continue;  // the split gather exists for the int64 box only
    rc = save(true);
    if (rc) break;
    h->gather_lds = c.g;
    if (retile(h, c.tx, c.tr) != CYLPIC_OK) continue;  // too much LDS for this shape
    rc = cylpic_push_particles(h);  // warm-up
    if (rc) break;
    HIPCHK(hipEventRecord(e0, h->stream));
    for (int i = 0; i < pushes && !rc; ++i) rc = cylpic_push_particles(h);
    if (rc) break;
    HIPCHK(hipEventRecord(e1, h->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    if (best < 0 || ms < best) {
      best = ms;
      pick = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // back to the snapshot, under the chosen shape
  int rc2 = save(true);
  if (!rc2) {
    h->gather_lds = pick.g;
    rc2 = retile(h, pick.tx, pick.tr);
    if (rc2) {  // cannot happen for a shape that ran; fall back to the old one
      h->gather_lds = g0;
      rc2 = retile(h, tx0, tr0);
    }
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  (void)hipFree(snap);
  h->push_count = push_count0;
  h->sort_interval = interval0;
  h->sort_radix = radix0;
  h->sort_migrate = migrate0;
  (void)shift0;  // retile starts the tile offset at 0: keys are recomputed from the positions
  if (rc) return rc;
  if (rc2) return rc2;
  for (auto& s : h->sp) {
    rc = refresh_count(h, s);
    if (rc) return rc;
  }
  if (tile_x) *tile_x = pick.tx;
  if (tile_r) *tile_r = pick.tr;
  if (gather_lds) *gather_lds = pick.g;
  if (ms_best) *ms_best = best / pushes;
  return check_device_errors(h);
}

int cylpic_synchronize(cylpic_ctx* h) {
  CHECK_H(h);
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipStreamSynchronize(h->stream));
  return check_device_errors(h);
}

int cylpic_profile(cylpic_ctx* h, int enable) {
  CHECK_H(h);
  if (!enable) resolve_profile(h);
  h->profiling = enable != 0;
  if (enable) h->prof.clear();
  return CYLPIC_OK;
}

int cylpic_kernel_time(cylpic_ctx* h, const char* name, double* ms, int64_t* count) {
  CHECK_H(h);
  if (!name || !ms || !count) return fail(CYLPIC_EINVAL, "null argument");
  resolve_profile(h);
  auto it = h->prof.find(name);
  if (it == h->prof.end()) {
    *ms = 0;
    *count = 0;
    return CYLPIC_OK;
  }
  *ms = it->second.ms;
  *count = it->second.count;
  return CYLPIC_OK;
}

}  // extern "C"

// ======================================================================
// Batches of runs stepped through HIP graphs
// ======================================================================
// A batch step is every run's timestep on its own stream, forked from and
// joined to the batch's stream, with the mixed pushes of all runs that share
// a mode count (with the default tile, accumulator and gather) in one
// persistent launch between the runs' field half updates and the rest of
// their steps.  It runs the same host code as an eager step, so every run's
// host-side state (buffer swaps, push count, sort schedule) advances as it
// would.
//
// cylpic_batch_steps goes in chunks of P steps, P the period of the runs'
// buffer parities (the least common multiple of twice their sort intervals).
// A chunk that starts from the state the previous chunk started and ended in
// (a periodic run: no window move, injection or download in between) is
// captured into one HIP graph, which is kept under that state's signature
// and relaunched whenever a chunk starts from the same state, with only the
// push counts and times advanced on the host.  Other chunks, and single
// steps, run eagerly.  Laser sources go through a device ring of P steps that
// the chunk's kernels read, filled by one copy on the batch's stream before
// the chunk from one of two pinned buffers used in turn, each guarded by the
// event of the chunk that last used it; push descriptors likewise through
// pinned blocks (the graphs own theirs).
struct cylpic_batch {
  int device = 0;
  std::vector<cylpic_ctx*> runs;
  hipStream_t origin = nullptr;
  hipEvent_t fork = nullptr, pushed = nullptr;
  std::vector<hipEvent_t> ready, done;
  hipEvent_t launched[2] = {nullptr, nullptr};
  int slot = 0;
  struct Group {
    int M = 0;
    size_t lds = 0;
    int nblocks = 0;
    std::vector<std::pair<int, int>> members;  // (run, species)
    void* d_desc = nullptr;
    int* d_item_start = nullptr;
    int* d_counters = nullptr;
  };
  std::vector<Group> groups;
  std::vector<size_t> src_off;  // run r's sources start at src_off[r] within a step's block
  size_t src_len = 0;           // doubles per step (both sides, every run)
  size_t chunk_cap = 0;         // steps the buffers below hold
  double* src_pin[2] = {nullptr, nullptr};
  double* src_ring = nullptr;   // device
  char* eager_desc[2] = {nullptr, nullptr};
  struct Entry {
    hipGraphExec_t exec = nullptr;
    char* desc = nullptr;     // pinned descriptors the graph reads
    std::vector<long> radix;  // host-decided radix sorts per run in the chunk
    long chunk = 0;           // the chunk that last launched it
  };
  std::map<std::string, Entry> cache;  // periodic chunks, by start state
  std::vector<Entry> retired;          // launched once, freed once that launch is done
  std::string periodic;                // start state of the last chunk that ended where it started
  long steps = 0, chunks = 0, instantiations = 0, replays = 0;
};

namespace {

size_t desc_block_bytes(const cylpic_batch* b) {
  size_t n = 0;
  for (auto& gp : b->groups) n += gp.members.size() * push_desc_mixed_bytes() + (gp.members.size() + 1) * sizeof(int);
  return n + 64;
}

void free_entry(cylpic_batch::Entry& e) {
  if (e.exec) (void)hipGraphExecDestroy(e.exec);
  if (e.desc) (void)hipHostFree(e.desc);
  e.exec = nullptr;
  e.desc = nullptr;
}

void free_chunk_buffers(cylpic_batch* b) {
  for (auto*& p : b->src_pin) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
  }
  for (auto*& p : b->eager_desc) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
  }
  if (b->src_ring) (void)hipFree(b->src_ring);
  b->src_ring = nullptr;
  b->chunk_cap = 0;
}

void batch_free(cylpic_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->origin) (void)hipStreamSynchronize(b->origin);
  for (auto& kv : b->cache) free_entry(kv.second);
  for (auto& e : b->retired) free_entry(e);
  for (auto& gp : b->groups) {
    if (gp.d_desc) (void)hipFree(gp.d_desc);
    if (gp.d_item_start) (void)hipFree(gp.d_item_start);
    if (gp.d_counters) (void)hipFree(gp.d_counters);
  }
  free_chunk_buffers(b);
  for (auto e : b->ready) (void)hipEventDestroy(e);
  for (auto e : b->done) (void)hipEventDestroy(e);
  for (auto e : {b->fork, b->pushed, b->launched[0], b->launched[1]})
    if (e) (void)hipEventDestroy(e);
  if (b->origin) (void)hipStreamDestroy(b->origin);
  delete b;
}

// The work of one batch step on the streams (eager, or captured by the
// caller).  desc: pinned space for this step's push descriptors; src: this
// step's sources in device memory (src_len doubles: per run, x_min then
// x_max), null for none.
int batch_enqueue(cylpic_batch* b, char* desc, const double* src) {
  const int nr = (int)b->runs.size();
  HIPCHK(hipEventRecord(b->fork, b->origin));
  for (int r = 0; r < nr; ++r) HIPCHK(hipStreamWaitEvent(b->runs[r]->stream, b->fork, 0));
  std::vector<std::vector<int>> grouped(nr);
  for (int r = 0; r < nr; ++r) grouped[r].assign(b->runs[r]->sp.size(), -1);
  for (int gi = 0; gi < (int)b->groups.size(); ++gi)
    for (auto& m : b->groups[gi].members) grouped[m.first][m.second] = gi;
  // descriptor space of group gi: [members][desc], then the item starts
  std::vector<char*> gdesc(b->groups.size());
  {
    char* p = desc;
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
      gdesc[gi] = p;
      const size_t nm = b->groups[gi].members.size();
      p += nm * push_desc_mixed_bytes() + (nm + 1) * sizeof(int);
    }
  }
  std::vector<int> ndesc(b->groups.size(), 0);
  std::vector<std::vector<int>> items(b->groups.size());
  for (int r = 0; r < nr; ++r) {
    cylpic_ctx* h = b->runs[r];
    int rc = cylpic_update_eb_fields_half(h);
    if (!rc) rc = push_begin(h);
    if (rc) return rc;
    for (int k = 0; k < (int)h->sp.size(); ++k) {
      Species& s = h->sp[k];
      if (s.immobile || s.n <= 0) continue;
      const int gi = grouped[r][k];
      if (gi >= 0) {
        rc = push_species(h, s, gdesc[gi] + (size_t)ndesc[gi] * push_desc_mixed_bytes());
        ndesc[gi] += 1;
        items[gi].push_back(h->tl.ntiles);
      } else {
        rc = push_species(h, s, nullptr);
      }
      if (rc) return rc;
    }
    HIPCHK(hipEventRecord(b->ready[r], h->stream));
    HIPCHK(hipStreamWaitEvent(b->origin, b->ready[r], 0));
  }
  for (int gi = 0; gi < (int)b->groups.size(); ++gi) {
    auto& gp = b->groups[gi];
    if (!ndesc[gi]) continue;
    // descriptor d owns work items [start[d], start[d + 1])
    int* start = (int*)(gdesc[gi] + gp.members.size() * push_desc_mixed_bytes());
    start[0] = 0;
    for (int d = 0; d < ndesc[gi]; ++d) start[d + 1] = start[d] + items[gi][d];
    HIPCHK(hipMemcpyAsync(gp.d_desc, gdesc[gi], (size_t)ndesc[gi] * push_desc_mixed_bytes(), hipMemcpyHostToDevice,
                          b->origin));
    HIPCHK(hipMemcpyAsync(gp.d_item_start, start, (ndesc[gi] + 1) * sizeof(int), hipMemcpyHostToDevice, b->origin));
    launch_push_deposit_mixed_batch(gp.d_desc, gp.d_item_start, ndesc[gi], gp.M, gp.lds, gp.d_counters, gp.nblocks,
                                    b->origin);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(b->pushed, b->origin));
  for (int r = 0; r < nr; ++r) {
    cylpic_ctx* h = b->runs[r];
    HIPCHK(hipStreamWaitEvent(h->stream, b->pushed, 0));
    int rc = push_end(h);
    if (!rc) rc = cylpic_current_finish(h);
    if (rc) return rc;
    h->time += h->g.dt;
    const size_t sb = 2 * (size_t)(h->g.ny + 1);
    const double* lo = src ? src + b->src_off[r] : nullptr;
    rc = fields_final(h, lo, src ? lo + sb : nullptr);
    if (rc) return rc;
    HIPCHK(hipEventRecord(b->done[r], h->stream));
    HIPCHK(hipStreamWaitEvent(b->origin, b->done[r], 0));
  }
  return CYLPIC_OK;
}

// Captures `nsteps` steps into an instantiated graph (descriptors in `desc`,
// the sources of step j at src + j * src_len in device memory, or none).
int batch_capture(cylpic_batch* b, int nsteps, char* desc, const double* src, hipGraphExec_t* exec) {
  std::vector<bool> prof;
  for (cylpic_ctx* h : b->runs) {
    prof.push_back(h->profiling);
    h->profiling = false;  // no profiling events inside a graph
  }
  HIPCHK(hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
  int rc = CYLPIC_OK;
  const size_t db = desc_block_bytes(b);
  for (int j = 0; j < nsteps && !rc; ++j) rc = batch_enqueue(b, desc + j * db, src ? src + j * b->src_len : nullptr);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(b->origin, &graph);
  for (size_t r = 0; r < b->runs.size(); ++r) b->runs[r]->profiling = prof[r];
  if (rc || ec != hipSuccess || !graph) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc ? rc : fail(CYLPIC_EHIP, std::string("graph capture failed: ") + hipGetErrorString(ec));
  }
  const hipError_t ei = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return fail(CYLPIC_EHIP, std::string("graph instantiation failed: ") + hipGetErrorString(ei));
  b->instantiations += 1;
  return CYLPIC_OK;
}

template <class T>
void put(std::string& s, const T& v) {
  s.append(reinterpret_cast<const char*>(&v), sizeof(v));
}

// Everything of a run that a step's launches or their parameters depend on.
void run_signature(std::string& s, const cylpic_ctx* h, int period) {
  const Grid& g = h->g;
  for (int v : {g.nx, g.ny, g.M, g.ng}) put(s, v);
  for (double v : {g.dx, g.dy, g.dt, g.x_min, g.x_max, g.y_min, g.y_max, g.x_grid_min, g.y_grid_min}) put(s, v);
  for (int k = 0; k < 4; ++k) {
    put(s, g.bc_field[k]);
    put(s, g.bc_particle[k]);
  }
  for (int v : {h->tl.tx, h->tl.tr, h->tl.ntx, h->tl.ntr, h->tl.ntiles, h->tl.shift}) put(s, v);
  for (int k = 0; k < 15; ++k) {
    put(s, (const void*)h->f64.f[k]);
    put(s, (const void*)h->f32.f[k]);
  }
  for (int k = 0; k < 6; ++k) {
    put(s, (const void*)h->f64.bnd_xmin[k]);
    put(s, (const void*)h->f64.bnd_xmax[k]);
    put(s, (const void*)h->f32.bnd_xmin[k]);
    put(s, (const void*)h->f32.bnd_xmax[k]);
  }
  for (const void* p : {(const void*)h->spare64, (const void*)h->spare32, (const void*)h->mig, (const void*)h->sc_mem,
                        (const void*)h->src_dev[0], (const void*)h->src_dev[1], (const void*)h->geo,
                        (const void*)h->geof, (const void*)h->d_err, (const void*)h->d_sortstat,
                        (const void*)h->sc.k[0], (const void*)h->sc.k[1], (const void*)h->sc.v[0],
                        (const void*)h->sc.v[1], (const void*)h->sc.counts, (const void*)h->sc.scan_tmp})
    put(s, p);
  put(s, h->sc.cap);
  for (int v : {h->prec, h->ntab, h->sort_interval, h->slot_bits, (int)h->migrate, (int)h->sort_cells, h->acc_mode, h->gather_lds,
                h->smooth_its, h->smooth_comp_its, (int)h->cont_armed, (int)(h->push_count % period)})
    put(s, v);
  for (int v : h->smooth_strides) put(s, v);
  for (const Species& sp : h->sp) {
    put(s, sp.q);
    put(s, sp.m);
    for (long v : {sp.cap, sp.n, sp.n_extra}) put(s, v);
    for (int v : {sp.immobile, (int)sp.n_exact, (int)sp.ids, (int)sp.dirty, (int)sp.ranged, (int)sp.counted}) put(s, v);
    for (int k = 0; k < 7; ++k) {
      put(s, (const void*)sp.a[k]);
      put(s, (const void*)sp.b[k]);
    }
    for (int k = 0; k < 8; ++k) {
      put(s, (const void*)sp.ma[k]);
      put(s, (const void*)sp.mb[k]);
    }
    for (const void* p : {(const void*)sp.id[0], (const void*)sp.id[1], (const void*)sp.key[0], (const void*)sp.key[1],
                          (const void*)sp.tile_start, (const void*)sp.tile_start_alt, (const void*)sp.extra_start,
                          (const void*)sp.d_count, (const void*)sp.mig_counts, (const void*)sp.mig_counts_x,
                          (const void*)sp.mig_flag})
      put(s, p);
  }
}

// The least common multiple of the runs' doubled sort intervals, the period
// of every run's buffer parities, capped at 64 steps (a longer period only
// means chunks that never repeat: each is captured afresh).
int batch_period(const cylpic_batch* b) {
  long p = 2;
  for (const cylpic_ctx* h : b->runs) p = std::min(64L, std::lcm(p, 2L * std::max(1, h->sort_interval)));
  return (int)p;
}

}  // namespace

extern "C" {

int cylpic_batch_create(cylpic_batch** out, cylpic_ctx* const* runs, int nruns) {
  if (!out) return fail(CYLPIC_EINVAL, "null output");
  *out = nullptr;
  if (!runs || nruns < 1) return fail(CYLPIC_EINVAL, "a batch needs at least one run");
  for (int r = 0; r < nruns; ++r) {
    if (!runs[r]) return fail(CYLPIC_EINVAL, "null run handle");
    if (runs[r]->device != runs[0]->device) return fail(CYLPIC_EINVAL, "the runs of a batch share one device");
    for (int q = 0; q < r; ++q)
      if (runs[q] == runs[r]) return fail(CYLPIC_EINVAL, "a run appears twice in the batch");
  }
  cylpic_batch* b = new (std::nothrow) cylpic_batch();
  if (!b) return fail(CYLPIC_ENOMEM, "host allocation failed");
  b->device = runs[0]->device;
  b->runs.assign(runs, runs + nruns);
  auto bail = [&](int rc) {
    batch_free(b);
    return rc;
  };
  if (hipSetDevice(b->device) != hipSuccess) return bail(fail(CYLPIC_EHIP, "hipSetDevice failed"));
  if (hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(CYLPIC_EHIP, "stream creation failed"));
  for (hipEvent_t* e : {&b->fork, &b->pushed, &b->launched[0], &b->launched[1]})
    if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess)
      return bail(fail(CYLPIC_EHIP, "event creation failed"));
  b->ready.assign(nruns, nullptr);
  b->done.assign(nruns, nullptr);
  for (int r = 0; r < nruns; ++r)
    if (hipEventCreateWithFlags(&b->ready[r], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->done[r], hipEventDisableTiming) != hipSuccess)
      return bail(fail(CYLPIC_EHIP, "event creation failed"));
  // everything a step could allocate lazily is allocated now: a capture
  // cannot allocate
  for (cylpic_ctx* h : b->runs) {
    if (!h->mig && hipMalloc((void**)&h->mig, mig_scratch_words(h->tl.ntiles) * sizeof(uint32_t)) != hipSuccess)
      return bail(fail(CYLPIC_ENOMEM, "migration scratch allocation failed"));
    long cap = 0;
    for (auto& sp : h->sp) cap = std::max(cap, sp.cap);
    if (int rc = ensure_sort_scratch(h, cap)) return bail(rc);
    if (h->smooth_its + h->smooth_comp_its > 0) {
      const long nf = nfield(h);
      if (h->prec == CYLPIC_FP64 && !h->spare64 &&
          hipMalloc((void**)&h->spare64, nf * sizeof(cplx<double>)) != hipSuccess)
        return bail(fail(CYLPIC_ENOMEM, "smoothing buffer allocation failed"));
      if (h->prec == CYLPIC_MIXED && !h->spare32 &&
          hipMalloc((void**)&h->spare32, nf * sizeof(cplx<float>)) != hipSuccess)
        return bail(fail(CYLPIC_ENOMEM, "smoothing buffer allocation failed"));
    }
    b->src_off.push_back(b->src_len);
    b->src_len += 4 * (size_t)(h->g.ny + 1);
  }
  // the groups of the batched push: mixed species with the default tile,
  // accumulator and gather, by mode count
  for (int r = 0; r < nruns; ++r) {
    cylpic_ctx* h = b->runs[r];
    if (h->prec != CYLPIC_MIXED || !batch_push_supported(h->tl, h->g.M, h->acc_mode, h->gather_lds)) continue;
    for (int k = 0; k < (int)h->sp.size(); ++k) {
      if (h->sp[k].immobile || h->sp[k].cap <= 0) continue;
      cylpic_batch::Group* gp = nullptr;
      for (auto& gg : b->groups)
        if (gg.M == h->g.M) gp = &gg;
      if (!gp) {
        b->groups.emplace_back();
        gp = &b->groups.back();
        gp->M = h->g.M;
        gp->lds = push_lds_bytes_mixed(h->g, h->tl, h->acc_mode, h->gather_lds);
      }
      gp->members.emplace_back(r, k);
    }
  }
  for (auto& gp : b->groups) {
    const size_t nd = gp.members.size();
    gp.nblocks = push_batch_blocks(gp.M, gp.lds);
    if (hipMalloc(&gp.d_desc, nd * push_desc_mixed_bytes()) != hipSuccess ||
        hipMalloc((void**)&gp.d_item_start, (nd + 1) * sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&gp.d_counters, NUM_XCD * sizeof(int)) != hipSuccess)
      return bail(fail(CYLPIC_ENOMEM, "batch allocation failed"));
  }
  for (cylpic_ctx* h : b->runs)
    if (hipStreamSynchronize(h->stream) != hipSuccess) return bail(fail(CYLPIC_EHIP, "synchronize failed"));
  *out = b;
  return CYLPIC_OK;
}

// Chunk buffers for `steps` steps (outside any capture).  The kept graphs
// read the ring being replaced: they go too.
static int ensure_chunk_buffers(cylpic_batch* b, size_t steps) {
  if (steps <= b->chunk_cap) return CYLPIC_OK;
  HIPCHK(hipStreamSynchronize(b->origin));
  for (auto& kv : b->cache) free_entry(kv.second);
  b->cache.clear();
  for (auto& e : b->retired) free_entry(e);
  b->retired.clear();
  free_chunk_buffers(b);
  const size_t sbytes = steps * std::max<size_t>(b->src_len, 1) * sizeof(double);
  for (int k = 0; k < 2; ++k) {
    if (hipHostMalloc((void**)&b->src_pin[k], sbytes, 0) != hipSuccess ||
        hipHostMalloc((void**)&b->eager_desc[k], steps * desc_block_bytes(b), 0) != hipSuccess)
      return fail(CYLPIC_ENOMEM, "pinned batch buffer allocation failed");
  }
  if (hipMalloc((void**)&b->src_ring, sbytes) != hipSuccess)
    return fail(CYLPIC_ENOMEM, "batch source ring allocation failed");
  b->chunk_cap = steps;
  return CYLPIC_OK;
}

static std::string batch_signature(const cylpic_batch* b, int nsteps, bool src) {
  const int period = batch_period(b);
  std::string key;
  put(key, nsteps);
  put(key, (int)src);
  for (const cylpic_ctx* h : b->runs) run_signature(key, h, period);
  return key;
}

// One chunk of `nsteps` steps: a kept graph when the runs start from a state
// seen before and periodic; a new graph (kept if the chunk turns out
// periodic) when the previous chunk started and ended in this start state;
// else eagerly.  src: host sources, nsteps blocks of src_len doubles, or null.
static int batch_chunk(cylpic_batch* b, int nsteps, const double* src, bool graphs) {
  const int parity = b->slot;
  b->slot ^= 1;
  const long chunk = b->chunks++;
  // the buffers of this parity were last used by the chunk before last;
  // chunks complete in order, so every graph retired before it is free
  HIPCHK(hipEventSynchronize(b->launched[parity]));
  {
    size_t keep = 0;
    for (size_t i = 0; i < b->retired.size(); ++i) {
      if (b->retired[i].chunk <= chunk - 2) free_entry(b->retired[i]);
      else b->retired[keep++] = b->retired[i];
    }
    b->retired.resize(keep);
  }
  if (int rc = ensure_chunk_buffers(b, (size_t)nsteps)) return rc;
  if (src) {
    std::memcpy(b->src_pin[parity], src, (size_t)nsteps * b->src_len * sizeof(double));
    HIPCHK(hipMemcpyAsync(b->src_ring, b->src_pin[parity], (size_t)nsteps * b->src_len * sizeof(double),
                          hipMemcpyHostToDevice, b->origin));
  }
  const double* ring = src ? b->src_ring : nullptr;
  const std::string key = graphs ? batch_signature(b, nsteps, src != nullptr) : std::string();
  if (graphs) {
    auto it = b->cache.find(key);
    if (it != b->cache.end()) {
      HIPCHK(hipGraphLaunch(it->second.exec, b->origin));
      HIPCHK(hipEventRecord(b->launched[parity], b->origin));
      it->second.chunk = chunk;
      for (size_t r = 0; r < b->runs.size(); ++r) {
        cylpic_ctx* h = b->runs[r];
        h->push_count += nsteps;
        h->time += nsteps * h->g.dt;
        h->sort_radix += it->second.radix[r];
      }
      b->steps += nsteps;
      b->replays += 1;
      return CYLPIC_OK;
    }
  }
  if (graphs && key == b->periodic) {
    cylpic_batch::Entry e;
    e.chunk = chunk;
    for (const cylpic_ctx* h : b->runs) e.radix.push_back(-h->sort_radix);
    HIPCHK(hipHostMalloc((void**)&e.desc, (size_t)nsteps * desc_block_bytes(b), 0));
    int rc = batch_capture(b, nsteps, e.desc, ring, &e.exec);
    if (rc) {
      free_entry(e);
      return rc;
    }
    HIPCHK(hipGraphLaunch(e.exec, b->origin));
    HIPCHK(hipEventRecord(b->launched[parity], b->origin));
    b->steps += nsteps;
    for (size_t r = 0; r < b->runs.size(); ++r) e.radix[r] += b->runs[r]->sort_radix;
    if (batch_signature(b, nsteps, src != nullptr) == key && b->cache.size() < 8) b->cache[key] = e;
    else b->retired.push_back(e);
    return CYLPIC_OK;
  }
  const size_t db = desc_block_bytes(b);
  for (int j = 0; j < nsteps; ++j) {
    int rc = batch_enqueue(b, b->eager_desc[parity] + j * db, ring ? ring + j * b->src_len : nullptr);
    if (rc) return rc;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(b->launched[parity], b->origin));
  b->steps += nsteps;
  if (graphs && batch_signature(b, nsteps, src != nullptr) == key) b->periodic = key;
  return CYLPIC_OK;
}

int cylpic_batch_step(cylpic_batch* b, const double* const* src_xmin, const double* const* src_xmax) {
  if (!b) return fail(CYLPIC_EINVAL, "null batch");
  HIPCHK(hipSetDevice(b->device));
  std::vector<double> src;
  if (src_xmin || src_xmax) {
    src.assign(b->src_len, 0.0);
    for (size_t r = 0; r < b->runs.size(); ++r) {
      const size_t nb = 2 * (size_t)(b->runs[r]->g.ny + 1);
      if (src_xmin && src_xmin[r]) std::memcpy(&src[b->src_off[r]], src_xmin[r], nb * sizeof(double));
      if (src_xmax && src_xmax[r]) std::memcpy(&src[b->src_off[r] + nb], src_xmax[r], nb * sizeof(double));
    }
  }
  return batch_chunk(b, 1, src.empty() ? nullptr : src.data(), false);
}

int cylpic_batch_steps(cylpic_batch* b, int nsteps, const double* src) {
  if (!b) return fail(CYLPIC_EINVAL, "null batch");
  if (nsteps < 0) return fail(CYLPIC_EINVAL, "negative step count");
  HIPCHK(hipSetDevice(b->device));
  const int period = batch_period(b);
  int done = 0;
  while (done < nsteps) {
    const int n = std::min(period, nsteps - done);
    int rc = batch_chunk(b, n, src ? src + (size_t)done * b->src_len : nullptr, n == period);
    if (rc) return rc;
    done += n;
  }
  return CYLPIC_OK;
}

int cylpic_batch_synchronize(cylpic_batch* b) {
  if (!b) return fail(CYLPIC_EINVAL, "null batch");
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->origin));
  for (auto& e : b->retired) free_entry(e);
  b->retired.clear();
  for (cylpic_ctx* h : b->runs) {
    int rc = check_device_errors(h);
    if (rc) return rc;
  }
  return CYLPIC_OK;
}

int cylpic_batch_stats(cylpic_batch* b, int64_t* steps, int64_t* instantiations, int64_t* replays, int* push_groups) {
  if (!b) return fail(CYLPIC_EINVAL, "null batch");
  if (steps) *steps = b->steps;
  if (instantiations) *instantiations = b->instantiations;
  if (replays) *replays = b->replays;
  if (push_groups) *push_groups = (int)b->groups.size();
  return CYLPIC_OK;
}

int cylpic_batch_destroy(cylpic_batch* b) {
  if (!b) return fail(CYLPIC_EINVAL, "null batch");
  batch_free(b);
  return CYLPIC_OK;
}

}  // extern "C"
