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
#include "launch.h"

namespace cylpic {

constexpr int RADIX = 256;
constexpr int SORT_BLOCK = 256;
constexpr int IPT = 16;                       // keys per thread
constexpr int SORT_TILE = SORT_BLOCK * IPT;   // 4096 keys per block
constexpr int WAVES = SORT_BLOCK / 64;
constexpr int WAVE_ITEMS = SORT_TILE / WAVES; // 1024 keys per wavefront
constexpr long INTERLEAVE_COLUMNS = 64;

// composite key of element i: pass 0 builds it from the tile key
__device__ __forceinline__ uint32_t comp_key(const uint32_t* __restrict__ keys, long i, int compose_bits) {
  const uint32_t k = keys[i];
  return compose_bits >= 0 ? ((k << compose_bits) | ((uint32_t)i & ((1u << compose_bits) - 1u))) : k;
}

__global__ __launch_bounds__(SORT_BLOCK) void k_radix_count(const uint32_t* __restrict__ keys, long n,
                                                             int shift, int compose_bits,
                                                             uint32_t* __restrict__ counts, int nb) {
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

__device__ __forceinline__ uint32_t slot_get(int slot, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3) {
  return slot == 0 ? c0 : (slot == 1 ? c1 : (slot == 2 ? c2 : c3));
}

// compose_bits >= 0: first pass (keys are tile keys, values the index);
// out_shift >= 0: last pass (write key >> out_shift, the tile key).
__global__ __launch_bounds__(SORT_BLOCK) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                               const uint32_t* __restrict__ vin, long n,
                                                               int shift, int compose_bits, int out_shift,
                                                               const uint32_t* __restrict__ offs,
                                                               int nb, uint32_t* __restrict__ kout,
                                                               uint32_t* __restrict__ vout) {
  __shared__ uint32_t wcnt[WAVES][RADIX];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  const long base = (long)blockIdx.x * SORT_TILE + (long)w * WAVE_ITEMS;
  uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;  // running count of digit lane + 64 * slot
  uint32_t key[IPT], val[IPT], rank[IPT];
#pragma unroll
  for (int r = 0; r < IPT; ++r) {
    const long i = base + (long)r * 64 + lane;
    const bool valid = i < n;
    const uint32_t k = valid ? comp_key(kin, i, compose_bits) : 0u;
    const uint32_t v = valid ? (vin ? vin[i] : (uint32_t)i) : 0u;
    const uint32_t d = (k >> shift) & (RADIX - 1);
    unsigned long long active = __ballot(valid);
    uint32_t myrank = 0;
    while (active) {
      const int leader = __ffsll((long long)active) - 1;
      const uint32_t d0 = (uint32_t)__builtin_amdgcn_readlane((int)d, leader);
      const unsigned long long same = __ballot(valid && d == d0);
      const int slot = (int)(d0 >> 6), owner = (int)(d0 & 63);
      const uint32_t basecnt =
          (uint32_t)__builtin_amdgcn_readlane((int)slot_get(slot, c0, c1, c2, c3), owner);
      if (valid && d == d0) myrank = basecnt + (uint32_t)__popcll(same & lt_mask);
      if (lane == owner) {
        const uint32_t add = (uint32_t)__popcll(same);
        if (slot == 0) c0 += add;
        else if (slot == 1) c1 += add;
        else if (slot == 2) c2 += add;
        else c3 += add;
      }
      active &= ~same;
    }
    key[r] = k;
    val[r] = v;
    rank[r] = myrank;
  }
  wcnt[w][lane] = c0;
  wcnt[w][lane + 64] = c1;
  wcnt[w][lane + 128] = c2;
  wcnt[w][lane + 192] = c3;
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
                                                            uint32_t* __restrict__ partial) {
  const long base = (long)blockIdx.x * SCAN_TILE + (long)threadIdx.x * SCAN_IPT;
  uint32_t s = 0;
  for (int k = 0; k < SCAN_IPT; ++k)
    if (base + k < n) s += in[base + k];
  uint32_t tot;
  block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(uint32_t* __restrict__ data, long n,
                                                           const uint32_t* __restrict__ partial_scanned) {
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

static void exclusive_scan(uint32_t* data, long n, uint32_t* tmp, hipStream_t s) {
  const long nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb <= 1) {
    hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(SCAN_BLOCK), 0, s, data, n, (const uint32_t*)nullptr);
    return;
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, data, n, tmp);
  exclusive_scan(tmp, nb, tmp + nb, s);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, data, n, tmp);
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
                uint32_t* tiles_out, uint32_t** perm_out, hipStream_t s) {
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
    hipLaunchKernelGGL(k_radix_count, dim3(nb), dim3(SORT_BLOCK), 0, s, kin, n, shift, compose, sc.counts, nb);
    exclusive_scan(sc.counts, (long)nb * RADIX, sc.scan_tmp, s);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(SORT_BLOCK), 0, s, kin, vin, n, shift, compose,
                       last ? slot_bits : -1, sc.counts, nb, kout, vout);
    kin = kout;
    vin = vout;
    cur ^= 1;
  }
  *perm_out = (uint32_t*)vin;
}

// tile_start[t] = first sorted index with key >= t, t in 0..ntiles, and
// tile_start[ntiles + 1] = n
__global__ void k_tile_bounds(const uint32_t* __restrict__ key, long n, int ntiles, int* __restrict__ tile_start) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const long lo = (i == 0) ? -1 : (long)key[i - 1];
  const long hi = (i == n) ? (long)ntiles : (long)key[i];
  for (long t = lo + 1; t <= hi && t <= ntiles; ++t) tile_start[t] = (int)i;
  if (i == n) tile_start[ntiles + 1] = (int)n;  // sentinel past the removed ones
}

void launch_tile_bounds(const uint32_t* sorted_keys, long n, int ntiles, int* tile_start, hipStream_t s) {
  hipLaunchKernelGGL(k_tile_bounds, dim3((unsigned)((n + 1 + 255) / 256)), dim3(256), 0, s, sorted_keys, n,
                     ntiles, tile_start);
}

struct Ptr8 {
  const void* src[8];
  void* dst[8];
};

// Destination of sorted index i.  With `inter` set, the particles of each
// tile segment [s, s+L) are transposed as an S-column table: the one at
// segment offset j goes to column j mod S, row j / S, columns laid out one
// after another (ragged: the first L mod S columns hold one more row).
// Consecutive destinations then come from positions S apart in the stable
// order, i.e. (for a cell-grouped load) from different cells.
__device__ __forceinline__ long dest_of(long i, const uint32_t* __restrict__ key, const int* __restrict__ ts,
                                        int inter) {
  if (!inter) return i;
  const uint32_t t = key[i];
  const long s = ts[t], L = (long)ts[t + 1] - s, j = i - s;
  const long S = INTERLEAVE_COLUMNS, A = L / S, R = L % S;
  const long b = j % S, a = j / S;
  return s + b * A + (b < R ? b : R) + a;
}

template <class W>
__global__ void k_gather(const uint32_t* __restrict__ perm, long n, Ptr8 p, int narr,
                         const uint32_t* __restrict__ key, const int* __restrict__ ts, int inter) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = perm[i];
  const long d = dest_of(i, key, ts, inter);
  for (int a = 0; a < narr; ++a) ((W*)p.dst[a])[d] = ((const W*)p.src[a])[j];
}

void launch_gather_f64(const uint32_t* perm, long n, const double* const* src, double* const* dst, int narr,
                       const uint32_t* key, const int* ts, int inter, hipStream_t s) {
  Ptr8 p;
  for (int a = 0; a < narr; ++a) {
    p.src[a] = src[a];
    p.dst[a] = dst[a];
  }
  if (n > 0)
    hipLaunchKernelGGL(k_gather<uint64_t>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, perm, n, p, narr,
                       key, ts, inter);
}
void launch_gather_32(const uint32_t* perm, long n, const uint32_t* const* src, uint32_t* const* dst, int narr,
                      const uint32_t* key, const int* ts, int inter, hipStream_t s) {
  Ptr8 p;
  for (int a = 0; a < narr; ++a) {
    p.src[a] = src[a];
    p.dst[a] = dst[a];
  }
  if (n > 0)
    hipLaunchKernelGGL(k_gather<uint32_t>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, perm, n, p, narr,
                       key, ts, inter);
}
void launch_gather_64(const uint32_t* perm, long n, const uint64_t* src, uint64_t* dst, const uint32_t* key,
                      const int* ts, int inter, hipStream_t s) {
  Ptr8 p;
  p.src[0] = src;
  p.dst[0] = dst;
  if (n > 0)
    hipLaunchKernelGGL(k_gather<uint64_t>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, perm, n, p, 1, key,
                       ts, inter);
}

}  // namespace cylpic
