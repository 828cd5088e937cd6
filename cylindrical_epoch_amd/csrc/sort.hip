// Stable LSD radix sort of the particles' integer tile keys on gfx950, and
// the payload permutation.  The sorted order equals a stable CPU sort by key
// (the particles with equal keys keep their previous relative order), which
// is what the parity tests check bit-exactly.
//
// One pass sorts 8 key bits:
//   count   - per-block digit histogram (4096 keys per 256-thread block)
//   scan    - exclusive scan of the [digit][block] histogram
//   scatter - stable in-block rank per digit, computed one 64-key wavefront
//             row at a time: the lanes holding the same digit are found with
//             one ballot per distinct digit in the row (few for nearly sorted
//             keys), and the per-digit running counts live in registers
//             spread across the wavefront's lanes.
#include <algorithm>

#include "launch.h"
#include "migrate.h"

namespace cylpic {

constexpr int RADIX = 256;
constexpr int SORT_BLOCK = 256;
constexpr int IPT = 16;                       // keys per thread
constexpr int SORT_TILE = SORT_BLOCK * IPT;   // 4096 keys per block
constexpr int WAVES = SORT_BLOCK / 64;
constexpr int WAVE_ITEMS = SORT_TILE / WAVES; // 1024 keys per wavefront

// composite key of element i: pass 0 builds it from the tile key
__device__ __forceinline__ uint32_t comp_key(const uint32_t* __restrict__ keys, long i, int compose_bits) {
  const uint32_t k = keys[i];
  return compose_bits >= 0 ? ((k << compose_bits) | ((uint32_t)i & ((1u << compose_bits) - 1u))) : k;
}

__global__ __launch_bounds__(SORT_BLOCK) void k_radix_count(const uint32_t* __restrict__ keys, long n,
                                                             int shift, int compose_bits,
                                                             uint32_t* __restrict__ counts, int nb, Guard gd) {
  if (guard_skip(gd)) return;
  __shared__ uint32_t hist[RADIX];
  for (int d = threadIdx.x; d < RADIX; d += SORT_BLOCK) hist[d] = 0;
  __syncthreads();
  const long base = (long)blockIdx.x * SORT_TILE;
  for (int r = 0; r < IPT; ++r) {
    const long i = base + (long)r * SORT_BLOCK + threadIdx.x;
    if (i < n) atomicAdd(&hist[(comp_key(keys, i, compose_bits) >> shift) & (RADIX - 1)], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < RADIX; d += SORT_BLOCK) counts[(long)d * nb + blockIdx.x] = hist[d];
}

// compose_bits >= 0: first pass (keys are tile keys, values the index);
// out_shift >= 0: last pass (write key >> out_shift, the tile key).
__global__ __launch_bounds__(SORT_BLOCK) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                               const uint32_t* __restrict__ vin, long n,
                                                               int shift, int compose_bits, int out_shift,
                                                               const uint32_t* __restrict__ offs,
                                                               int nb, uint32_t* __restrict__ kout,
                                                               uint32_t* __restrict__ vout, Guard gd) {
  if (guard_skip(gd)) return;
  __shared__ uint32_t wcnt[WAVES][RADIX];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  const long base = (long)blockIdx.x * SORT_TILE + (long)w * WAVE_ITEMS;
  // this wavefront's running count per digit (its own row of wcnt)
  for (int d = lane; d < RADIX; d += 64) wcnt[w][d] = 0;
  uint32_t key[IPT], val[IPT], rank[IPT];
#pragma unroll
  for (int r = 0; r < IPT; ++r) {
    const long i = base + (long)r * 64 + lane;
    const bool valid = i < n;
    const uint32_t k = valid ? comp_key(kin, i, compose_bits) : 0u;
    const uint32_t v = valid ? (vin ? vin[i] : (uint32_t)i) : 0u;
    const uint32_t d = (k >> shift) & (RADIX - 1);
    // the lanes holding my digit: the AND over its 8 bits of the ballot of
    // that bit (or of its complement), 9 ballots whatever the digits
    unsigned long long same = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long m = __ballot(bit);
      same &= bit ? m : ~m;
    }
    const uint32_t before = (uint32_t)__popcll(same & lt_mask);
    uint32_t myrank = 0;
    if (valid) {
      const uint32_t cnt = wcnt[w][d];
      myrank = cnt + before;
      // the digit's first lane advances the count (after every lane read it)
      if (before == 0) wcnt[w][d] = cnt + (uint32_t)__popcll(same);
    }
    key[r] = k;
    val[r] = v;
    rank[r] = myrank;
  }
  __syncthreads();
  // exclusive prefix over the wavefronts of this block, per digit
  for (int d = threadIdx.x; d < RADIX; d += SORT_BLOCK) {
    uint32_t s = 0;
    for (int ww = 0; ww < WAVES; ++ww) {
      uint32_t c = wcnt[ww][d];
      wcnt[ww][d] = s;
      s += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < IPT; ++r) {
    const long i = base + (long)r * 64 + lane;
    if (i < n) {
      const uint32_t d = (key[r] >> shift) & (RADIX - 1);
      const uint32_t pos = offs[(long)d * nb + blockIdx.x] + wcnt[w][d] + rank[r];
      kout[pos] = out_shift >= 0 ? (key[r] >> out_shift) : key[r];
      vout[pos] = val[r];
    }
  }
}

// ---------------- exclusive scan of uint32 (recursive three-phase) -------
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_IPT = 4;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_IPT;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[SCAN_BLOCK / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t wbase = 0, tot = 0;
  for (int k = 0; k < SCAN_BLOCK / 64; ++k) {
    uint32_t s = wsum[k];
    if (k < w) wbase += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return wbase + x - v;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const uint32_t* __restrict__ in, long n,
                                                            uint32_t* __restrict__ partial, Guard gd) {
  if (guard_skip(gd)) return;
  const long base = (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_IPT;
  uint32_t s = 0;
  for (int k = 0; k < SCAN_IPT; ++k)
    if (base + k < n) s += in[base + k];
  uint32_t tot;
  block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(uint32_t* __restrict__ data, long n,
                                                           const uint32_t* __restrict__ partial_scanned, Guard gd) {
  if (guard_skip(gd)) return;
  const long base = (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_IPT;
  uint32_t v[SCAN_IPT];
  uint32_t s = 0;
  for (int k = 0; k < SCAN_IPT; ++k) {
    v[k] = (base + k < n) ? data[base + k] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t pre = block_exclusive_scan(s, &tot) + (partial_scanned ? partial_scanned[blockIdx.x] : 0u);
  for (int k = 0; k < SCAN_IPT; ++k) {
    if (base + k < n) data[base + k] = pre;
    pre += v[k];
  }
}

static void exclusive_scan(uint32_t* data, long n, uint32_t* tmp, hipStream_t s, Guard gd) {
  const long nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb <= 1) {
    hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(SCAN_BLOCK), 0, s, data, n, (const uint32_t*)nullptr, gd);
    return;
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, data, n, tmp, gd);
  exclusive_scan(tmp, nb, tmp + nb, s, gd);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, data, n, tmp, gd);
}

size_t sort_scratch_words(long n) {
  const long nb = (n + SORT_TILE - 1) / SORT_TILE;
  const long ncount = nb * RADIX;
  long scan = 0, m = ncount;
  while (m > 1) {
    m = (m + SCAN_TILE - 1) / SCAN_TILE;
    scan += m;
  }
  return (size_t)(4 * n + ncount + scan + 64);
}

void radix_sort(const uint32_t* tiles_in, long n, int ntiles, int slot_bits, SortScratch& sc,
                uint32_t* tiles_out, uint32_t** perm_out, hipStream_t s, Guard gd) {
  int bits = slot_bits;
  while ((1l << (bits - slot_bits)) <= (long)ntiles) ++bits;
  const int passes = (bits + 7) / 8;
  const int nb = (int)((n + SORT_TILE - 1) / SORT_TILE);
  const uint32_t* kin = tiles_in;
  const uint32_t* vin = nullptr;
  int cur = 0;
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    const int compose = p == 0 ? slot_bits : -1;
    const bool last = p == passes - 1;
    uint32_t* kout = last ? tiles_out : sc.k[cur];
    uint32_t* vout = sc.v[cur];
    hipLaunchKernelGGL(k_radix_count, dim3(nb), dim3(SORT_BLOCK), 0, s, kin, n, shift, compose, sc.counts, nb, gd);
    exclusive_scan(sc.counts, (long)nb * RADIX, sc.scan_tmp, s, gd);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(SORT_BLOCK), 0, s, kin, vin, n, shift, compose,
                       last ? slot_bits : -1, sc.counts, nb, kout, vout, gd);
    kin = kout;
    vin = vout;
    cur ^= 1;
  }
  *perm_out = (uint32_t*)vin;
}

// tile_start[t] = first sorted index with key >= t, t in 0..ntiles, and
// tile_start[ntiles + 1] = n
__global__ void k_tile_bounds(const uint32_t* __restrict__ key, long n, int ntiles, int* __restrict__ tile_start,
                              Guard gd) {
  if (guard_skip(gd)) return;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const long lo = (i == 0) ? -1 : (long)key[i - 1];
  const long hi = (i == n) ? (long)ntiles : (long)key[i];
  for (long t = lo + 1; t <= hi && t <= ntiles; ++t) tile_start[t] = (int)i;
  if (i == n) tile_start[ntiles + 1] = (int)n;  // sentinel past the removed ones
}

void launch_tile_bounds(const uint32_t* sorted_keys, long n, int ntiles, int* tile_start, hipStream_t s, Guard gd) {
  hipLaunchKernelGGL(k_tile_bounds, dim3((unsigned)((n + 1 + 255) / 256)), dim3(256), 0, s, sorted_keys, n,
                     ntiles, tile_start, gd);
}

__global__ void k_mark_stale(uint32_t* __restrict__ key, const int* __restrict__ ts, int ntiles, long n_main,
                             Guard gd) {
  if (guard_skip(gd)) return;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_main && i >= ts[ntiles]) key[i] = (uint32_t)ntiles;
}

void launch_mark_stale(uint32_t* key, const int* ts, int ntiles, long n_main, hipStream_t s, Guard gd) {
  if (n_main > 0)
    hipLaunchKernelGGL(k_mark_stale, dim3((unsigned)((n_main + 255) / 256)), dim3(256), 0, s, key, ts, ntiles,
                       n_main, gd);
}

__global__ void k_guard_count(int* counter, Guard gd) {
  if (!guard_skip(gd)) counter[0] += 1;
}

void launch_guard_count(int* counter, hipStream_t s, Guard gd) {
  hipLaunchKernelGGL(k_guard_count, dim3(1), dim3(1), 0, s, counter, gd);
}

__global__ void k_zero_words(uint32_t* __restrict__ p, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0u;
}

void launch_zero_words(void* p, long n, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)std::min<long>((n + 255) / 256, 4096)), dim3(256), 0, s,
                       (uint32_t*)p, n);
}

// tile columns relabelled one down: keys k < ntiles -> k - 1, and range
// starts ts_out[k] = ts_in[k + 1] (k < ntiles), the two last entries kept
__global__ void k_relabel_keys(uint32_t* __restrict__ key, long n, uint32_t ntiles) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && key[i] < ntiles) key[i] -= 1;
}
__global__ void k_relabel_ranges(const int* __restrict__ ts_in, int* __restrict__ ts_out, int ntiles) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ntiles) ts_out[k] = ts_in[k + 1];
  else if (k <= ntiles + 1) ts_out[k] = ts_in[k];
}

void launch_relabel_tiles(uint32_t* key, long n, int ntiles, const int* ts_in, int* ts_out, hipStream_t s) {
  if (key && n > 0)
    hipLaunchKernelGGL(k_relabel_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, key, n, (uint32_t)ntiles);
  hipLaunchKernelGGL(k_relabel_ranges, dim3((ntiles + 2 + 255) / 256), dim3(256), 0, s, ts_in, ts_out, ntiles);
}

__global__ void k_div_keys(uint32_t* __restrict__ key, long n, uint32_t div) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) key[i] /= div;
}

void launch_div_keys(uint32_t* key, long n, uint32_t div, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_div_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, key, n, div);
}

struct Ptr8 {
  const void* src[8];
  void* dst[8];
};

// Interleave of each tile (rounds): the particles of a tile segment,
// grouped by cell (ascending cell key, stable within a cell), are laid out as
// rounds, round r holding the r-th particle of every cell that has more than
// r, in cell order.  While every cell still has particles a round is the
// tile's whole cell list, so 64 consecutive particles sit in 64 different
// cells (and, for a quadrant-sorted load, mostly in one half-cell quadrant),
// whose deposit stencils start on distinct current-box words, however uneven
// the cell counts.  With perm[i] the source of sorted index i and dest(i) its
// place in the rounds layout, map[dest(i)] = perm[i]: the payload gather
// then reads map in order and writes coalesced.  One workgroup per tile;
// workgroup ntiles keeps the removed particles (key ntiles * cells * 4,
// sorted last) in place.  Destination of offset j in cell c at rank r: the
// sum over cells c' of min(n_c', r) (the earlier rounds), plus the cells
// c' < c with n_c' > r.
__device__ __forceinline__ long lower_bound_u32(const uint32_t* __restrict__ a, long lo, long hi, uint32_t v) {
  while (lo < hi) {
    const long mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_round_dest(const uint32_t* __restrict__ key, const uint32_t* __restrict__ perm,
                                                    long n, int ntiles, int cells, uint32_t* __restrict__ map) {
  extern __shared__ int cnt[];  // cells counts, then cells starts within the segment
  int* first = cnt + cells;
  const int t = blockIdx.x;
  const uint32_t K = (uint32_t)cells * 4u;
  const long s = lower_bound_u32(key, 0, n, (uint32_t)t * K);
  const long e = t == ntiles ? n : lower_bound_u32(key, s, n, (uint32_t)(t + 1) * K);
  if (t == ntiles) {
    for (long i = s + threadIdx.x; i < e; i += blockDim.x) map[i] = perm[i];
    return;
  }
  for (int c = threadIdx.x; c < cells; c += blockDim.x)
    first[c] = (int)(lower_bound_u32(key, s, e, ((uint32_t)t * (uint32_t)cells + (uint32_t)c) * 4u) - s);
  __syncthreads();
  for (int c = threadIdx.x; c < cells; c += blockDim.x) cnt[c] = (c + 1 < cells ? first[c + 1] : (int)(e - s)) - first[c];
  __syncthreads();
  for (long i = s + threadIdx.x; i < e; i += blockDim.x) {
    const int j = (int)(i - s);
    const int c = (int)(key[i] / 4u - (uint32_t)t * (uint32_t)cells);
    const int r = j - first[c];
    int d = 0;
    for (int q = 0; q < cells; ++q) {
      const int nq = cnt[q];
      d += min(nq, r) + (q < c && nq > r ? 1 : 0);
    }
    map[s + d] = perm[i];
  }
}

void launch_round_map(const uint32_t* key, const uint32_t* perm, long n, int ntiles, int cells, uint32_t* map,
                      hipStream_t s) {
  if (n <= 0) return;
  const size_t lds = (size_t)(2 * cells) * sizeof(int);
  if (lds > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_round_dest), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
  hipLaunchKernelGGL(k_round_dest, dim3((unsigned)ntiles + 1), dim3(256), lds, s, key, perm, n, ntiles, cells, map);
}

template <class W>
__global__ void k_gather(const uint32_t* __restrict__ perm, long n, Ptr8 p, int narr, Guard gd) {
  if (guard_skip(gd)) return;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = perm[i];
  for (int a = 0; a < narr; ++a) ((W*)p.dst[a])[i] = ((const W*)p.src[a])[j];
}

void launch_gather_f64(const uint32_t* perm, long n, const double* const* src, double* const* dst, int narr,
                       hipStream_t s, Guard gd) {
  Ptr8 p;
  for (int a = 0; a < narr; ++a) {
    p.src[a] = src[a];
    p.dst[a] = dst[a];
  }
  if (n > 0)
    hipLaunchKernelGGL(k_gather<uint64_t>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, perm, n, p, narr,
                       gd);
}
void launch_gather_32(const uint32_t* perm, long n, const uint32_t* const* src, uint32_t* const* dst, int narr,
                      hipStream_t s, Guard gd) {
  Ptr8 p;
  for (int a = 0; a < narr; ++a) {
    p.src[a] = src[a];
    p.dst[a] = dst[a];
  }
  if (n > 0)
    hipLaunchKernelGGL(k_gather<uint32_t>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, perm, n, p, narr,
                       gd);
}
void launch_gather_64(const uint32_t* perm, long n, const uint64_t* src, uint64_t* dst, hipStream_t s,
                      Guard gd) {
  Ptr8 p;
  p.src[0] = src;
  p.dst[0] = dst;
  if (n > 0)
    hipLaunchKernelGGL(k_gather<uint64_t>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, perm, n, p, 1, gd);
}

// ---------------- neighbour migration sort -------------------------------
// Between two sorts a particle moves less than one cell in x and in r (the
// CFL limit), so its new tile is its old tile or one of the 8 around it (x
// taken modulo the tile columns, which covers periodic wraps).  If the keys
// were sorted by tile before the push, the stable order by new key is: for
// each destination tile, the particles arriving from its source tiles in
// ascending source index, each source's particles in their current order.
// That order is produced by one counting pass over the keys, an offsets pass
// over the tiles and one scatter pass of the payload, instead of the radix
// passes plus a gather.  A particle whose tile moved farther sets a flag and
// the caller runs the radix sort instead.
constexpr int MIG_BLOCK = 256;
constexpr int MIG_WAVES = MIG_BLOCK / 64;

// counts (and counts_x for the extra ranges) by direction; flag[0] |= 1 on a
// farther move, flag[1] += particles kept
__global__ __launch_bounds__(MIG_BLOCK) void k_mig_count(const uint32_t* __restrict__ key,
                                                         const int* __restrict__ ts, ExtraRange xr, int ntx,
                                                         int ntiles, MigCount mc) {
  __shared__ uint32_t c[MIG_DIRS];
  const int S = blockIdx.x;
  mig_count_tile(key, S, ts[S], ts[S + 1], ntx, ntiles, mc.counts, mc.flag, c);
  if (xr.start)
    mig_count_tile(key, S, xr.base + xr.start[S], xr.base + xr.start[S + 1], ntx, ntiles, mc.counts_x, mc.flag, c);
}

// per destination tile D: tot[D] = arrivals; pre[S * 9 + dir] = arrivals at
// D from main ranges of sources before S (ascending tile index); the extra
// ranges follow all main ranges: pre_x[S * 9 + dir] = all main arrivals +
// extra arrivals from sources before S.  tot[ntiles] = 0.
__global__ void k_mig_offsets(const uint32_t* __restrict__ counts, const uint32_t* __restrict__ counts_x, int ntx,
                              int ntr, int ntiles, uint32_t* __restrict__ tot, uint32_t* __restrict__ pre,
                              uint32_t* __restrict__ pre_x, Guard gd) {
  if (guard_skip(gd)) return;
  const int D = blockIdx.x * blockDim.x + threadIdx.x;
  if (D > ntiles) return;
  if (D == ntiles) {
    tot[D] = 0;
    return;
  }
  const int tx = D % ntx, ty = D / ntx;
  int src[9];
  uint32_t c[9], cx[9];
#pragma unroll
  for (int dir = 0; dir < 9; ++dir) {
    const int dtx = dir % 3 - 1, dty = dir / 3 - 1;
    const int sty = ty - dty;
    src[dir] = -1;
    c[dir] = 0;
    cx[dir] = 0;
    if (sty >= 0 && sty < ntr) {
      const int stx = ((tx - dtx) % ntx + ntx) % ntx;
      src[dir] = sty * ntx + stx;
      c[dir] = counts[(long)src[dir] * MIG_DIRS + dir];
      if (counts_x) cx[dir] = counts_x[(long)src[dir] * MIG_DIRS + dir];
    }
  }
  uint32_t total = 0, total_x = 0;
#pragma unroll
  for (int dir = 0; dir < 9; ++dir) {
    total += c[dir];
    total_x += cx[dir];
  }
#pragma unroll
  for (int dir = 0; dir < 9; ++dir) {
    if (src[dir] < 0) continue;
    uint32_t p = 0, px = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k)
      if (src[k] >= 0 && (src[k] < src[dir] || (src[k] == src[dir] && k < dir))) {
        p += c[k];
        px += cx[k];
      }
    pre[(long)src[dir] * 9 + dir] = p;
    if (counts_x) pre_x[(long)src[dir] * 9 + dir] = total + px;
  }
  tot[D] = total + total_x;
}

__global__ void k_mig_sentinel(int* ts, int ntiles, Guard gd) {
  if (!guard_skip(gd)) ts[ntiles + 1] = ts[ntiles];
}

// Stable rank of each particle of the range [p0, p1) among those with the
// same direction, then payload copy to its sorted position.
template <class W>
__device__ __forceinline__ void mig_scatter_range(const uint32_t* __restrict__ key, int S, int p0, int p1, int ntx,
                                                  int ntiles, const uint32_t* __restrict__ pre,
                                                  const int* __restrict__ ts_new, const Ptr8& p, int narr,
                                                  uint32_t* __restrict__ key_out, const uint64_t* __restrict__ id_in,
                                                  uint64_t* __restrict__ id_out, uint32_t* run,
                                                  uint32_t (*wc)[9]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long lt = (1ull << lane) - 1ull;
  if (threadIdx.x < 9) run[threadIdx.x] = 0;
  __syncthreads();
  for (int base = p0; base < p1; base += MIG_BLOCK) {
    const int i = base + (int)threadIdx.x;
    const bool valid = i < p1;
    const uint32_t d = valid ? key[i] : (uint32_t)ntiles;
    const int dir = valid ? mig_dir(d, S, ntx, (uint32_t)ntiles) : -2;
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const unsigned long long m = __ballot(dir == k);
      if (dir == k) r = (uint32_t)__popcll(m & lt);
      if (lane == 0) wc[w][k] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (dir >= 0 && dir < 9) {
      uint32_t off = run[dir];
      for (int ww = 0; ww < w; ++ww) off += wc[ww][dir];
      const long dst = (long)ts_new[d] + pre[(long)S * 9 + dir] + off + r;
      for (int a = 0; a < narr; ++a) ((W*)p.dst[a])[dst] = ((const W*)p.src[a])[i];
      key_out[dst] = d;
      if (id_in) id_out[dst] = id_in[i];
    }
    __syncthreads();
    if (threadIdx.x < 9) {
      uint32_t t = 0;
      for (int ww = 0; ww < MIG_WAVES; ++ww) t += wc[ww][threadIdx.x];
      run[threadIdx.x] += t;
    }
    __syncthreads();
  }
}

// per source tile: its main range, then its extra range
template <class W>
__global__ __launch_bounds__(MIG_BLOCK) void k_mig_scatter(const uint32_t* __restrict__ key,
                                                           const int* __restrict__ ts, ExtraRange xr, int ntx,
                                                           int ntiles, const uint32_t* __restrict__ pre,
                                                           const uint32_t* __restrict__ pre_x,
                                                           const int* __restrict__ ts_new, Ptr8 p, int narr,
                                                           uint32_t* __restrict__ key_out,
                                                           const uint64_t* __restrict__ id_in,
                                                           uint64_t* __restrict__ id_out, Guard gd) {
  if (guard_skip(gd)) return;
  __shared__ uint32_t run[9];
  __shared__ uint32_t wc[MIG_WAVES][9];
  const int S = blockIdx.x;
  mig_scatter_range<W>(key, S, ts[S], ts[S + 1], ntx, ntiles, pre, ts_new, p, narr, key_out, id_in, id_out, run, wc);
  if (xr.start)
    mig_scatter_range<W>(key, S, xr.base + xr.start[S], xr.base + xr.start[S + 1], ntx, ntiles, pre_x, ts_new, p,
                         narr, key_out, id_in, id_out, run, wc);
}

size_t mig_scratch_words(int ntiles) {
  long scan = 0, m = ntiles + 1;
  while (m > 1) {
    m = (m + SCAN_TILE - 1) / SCAN_TILE;
    scan += m;
  }
  return (size_t)ntiles * 18 + (size_t)scan + 64;
}

void launch_mig_count(const uint32_t* key, const int* ts, ExtraRange xr, const Tiles& tl, MigCount mc,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_mig_count, dim3(tl.ntiles), dim3(MIG_BLOCK), 0, s, key, ts, xr, tl.ntx, tl.ntiles, mc);
}

void launch_mig_scatter(const uint32_t* key, const int* ts, ExtraRange xr, const Tiles& tl, MigCount mc,
                        uint32_t* scratch, int* ts_new, int word_bytes, const void* const* src, void* const* dst,
                        int narr, uint32_t* key_out, const uint64_t* id_in, uint64_t* id_out, hipStream_t s,
                        Guard gd) {
  const int nt = tl.ntiles;
  uint32_t* pre = scratch;
  uint32_t* pre_x = pre + (size_t)nt * 9;
  uint32_t* tmp = pre_x + (size_t)nt * 9;
  hipLaunchKernelGGL(k_mig_offsets, dim3((nt + 1 + 255) / 256), dim3(256), 0, s, mc.counts,
                     xr.start ? (const uint32_t*)mc.counts_x : nullptr, tl.ntx, tl.ntr, nt, (uint32_t*)ts_new, pre,
                     pre_x, gd);
  exclusive_scan((uint32_t*)ts_new, nt + 1, tmp, s, gd);
  hipLaunchKernelGGL(k_mig_sentinel, dim3(1), dim3(1), 0, s, ts_new, nt, gd);
  Ptr8 p;
  for (int a = 0; a < narr; ++a) {
    p.src[a] = src[a];
    p.dst[a] = dst[a];
  }
  if (word_bytes == 8)
    hipLaunchKernelGGL(k_mig_scatter<uint64_t>, dim3(nt), dim3(MIG_BLOCK), 0, s, key, ts, xr, tl.ntx, nt, pre, pre_x,
                       ts_new, p, narr, key_out, id_in, id_out, gd);
  else
    hipLaunchKernelGGL(k_mig_scatter<uint32_t>, dim3(nt), dim3(MIG_BLOCK), 0, s, key, ts, xr, tl.ntx, nt, pre, pre_x,
                       ts_new, p, narr, key_out, id_in, id_out, gd);
}

}  // namespace cylpic
