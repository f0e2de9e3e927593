// sort.hip — stable LSD radix sort of (cell key, particle index) pairs.
//
// Replaces thrust::sort_by_key over the uint cell ids
// (fluids/Simulator.cu:196-198).  Only the key and a 4-byte index are sorted;
// the payload is gathered once afterwards (k_permute), instead of dragging
// five float3/uint arrays through every pass.
//
// Per 8-bit digit pass (reduce-then-scan):
//   k_radix_hist      per-tile digit counts, digit-major table [digit][tile]
//   k_radix_scan_rows one workgroup per digit: exclusive scan of its row
//   k_radix_scatter   LDS-staged stable ranking of the tile, then coalesced
//                     scatter of runs of equal digits
// Stability: inside a tile, equal digits are ranked in index order (items
// processed in index order, wave64 ballot matching inside an item, per-wave
// prefix across the 4 waves); across tiles, by the row scan.
#include "pbf_device.h"
#include "pbf_launch.h"

namespace pbf {

namespace {

constexpr int kSortBlock = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortBlock * kSortItems;
constexpr int kRadix = 256;
constexpr int kWaves = kSortBlock / 64;

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return (1ull << lane) - 1ull;
}

// Lanes of this wave whose low `bits` digit bits equal d (wave64 ballots).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, int bits) {
  uint64_t m = ~0ull;
  for (int b = 0; b < bits; ++b) {
    const bool set = (d >> b) & 1u;
    const uint64_t v = __ballot(set);
    m &= set ? v : ~v;
  }
  return m;
}

// Exclusive scan over the 256 threads of the workgroup; *total gets the sum.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s, uint32_t* total) {
  const int tid = threadIdx.x;
  s[tid] = v;
  __syncthreads();
  for (int off = 1; off < kSortBlock; off <<= 1) {
    const uint32_t t = tid >= off ? s[tid - off] : 0u;
    __syncthreads();
    s[tid] += t;
    __syncthreads();
  }
  const uint32_t incl = s[tid];
  *total = s[kSortBlock - 1];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(kSortBlock) void k_radix_hist(const uint32_t* __restrict__ keys, int n,
                                                           int shift, int bits, int num_tiles,
                                                           uint32_t* __restrict__ hist) {
  __shared__ uint32_t cnt[kRadix];
  const int tid = threadIdx.x;
  cnt[tid] = 0u;
  __syncthreads();
  const int tile = blockIdx.x;
  const int base = tile * kSortTile;
  const uint32_t mask = (1u << bits) - 1u;
  const uint64_t lt = lanemask_lt();
  for (int k = 0; k < kSortItems; ++k) {
    const int i = base + k * kSortBlock + tid;
    const bool valid = i < n;
    const uint32_t d = valid ? (keys[i] >> shift) & mask : 0u;
    const uint64_t peers = match_digit(d, bits) & __ballot(valid);
    // one LDS atomic per distinct digit per wave
    if (valid && (peers & lt) == 0ull) atomicAdd(&cnt[d], (uint32_t)__popcll(peers));
  }
  __syncthreads();
  hist[(size_t)tid * num_tiles + tile] = cnt[tid];
}

__global__ __launch_bounds__(kSortBlock) void k_radix_scan_rows(uint32_t* __restrict__ hist,
                                                                int num_tiles,
                                                                uint32_t* __restrict__ totals) {
  __shared__ uint32_t s[kSortBlock];
  uint32_t* row = hist + (size_t)blockIdx.x * num_tiles;
  uint32_t running = 0u;
  for (int c0 = 0; c0 < num_tiles; c0 += kSortBlock) {
    const int t = c0 + (int)threadIdx.x;
    const uint32_t v = t < num_tiles ? row[t] : 0u;
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan(v, s, &tot);
    if (t < num_tiles) row[t] = running + ex;
    running += tot;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = running;
}

__global__ __launch_bounds__(kSortBlock) void k_radix_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int n, int shift, int bits, int num_tiles,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ totals) {
  __shared__ uint32_t s_key[kSortTile];
  __shared__ uint32_t s_val[kSortTile];
  __shared__ uint32_t wcnt[2][kWaves][kRadix];
  __shared__ uint32_t dtot[kRadix];
  __shared__ uint32_t dstart[kRadix];
  __shared__ uint32_t gbase[kRadix];
  __shared__ uint32_t scan_tmp[kRadix];

  const int tid = threadIdx.x;
  const int w = tid >> 6;
  const int tile = blockIdx.x;
  const int base = tile * kSortTile;
  const uint32_t mask = (1u << bits) - 1u;
  const uint64_t lt = lanemask_lt();

  {
    // global start of each digit for this tile: all smaller digits of every
    // tile + this digit in earlier tiles
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan(totals[tid], scan_tmp, &tot);
    gbase[tid] = ex + hist[(size_t)tid * num_tiles + tile];
  }
  dtot[tid] = 0u;
  for (int ww = 0; ww < kWaves; ++ww) {
    wcnt[0][ww][tid] = 0u;
    wcnt[1][ww][tid] = 0u;
  }

  uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    const int i = base + k * kSortBlock + tid;
    key[k] = i < n ? kin[i] : 0u;
    val[k] = i < n ? vin[i] : 0u;
  }
  __syncthreads();

#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    const int i = base + k * kSortBlock + tid;
    const bool valid = i < n;
    const uint32_t d = (key[k] >> shift) & mask;
    const uint64_t peers = match_digit(d, bits) & __ballot(valid);
    const uint32_t below = (uint32_t)__popcll(peers & lt);
    const int par = k & 1;
    if (valid && below == 0u) wcnt[par][w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    {
      // thread tid owns digit tid: prefix over the waves of this item on top
      // of the counts of earlier items; clears the other parity's table for
      // the next item
      uint32_t run = dtot[tid];
      for (int ww = 0; ww < kWaves; ++ww) {
        const uint32_t cc = wcnt[par][ww][tid];
        wcnt[par][ww][tid] = run;
        run += cc;
        wcnt[par ^ 1][ww][tid] = 0u;
      }
      dtot[tid] = run;
    }
    __syncthreads();
    rank[k] = valid ? wcnt[par][w][d] + below : 0u;
  }

  __syncthreads();
  {
    uint32_t tot;
    dstart[tid] = block_exclusive_scan(dtot[tid], scan_tmp, &tot);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    const int i = base + k * kSortBlock + tid;
    if (i < n) {
      const uint32_t d = (key[k] >> shift) & mask;
      const uint32_t p = dstart[d] + rank[k];
      if (p < (uint32_t)kSortTile) {
        s_key[p] = key[k];
        s_val[p] = val[k];
      }
    }
  }
  __syncthreads();
  const int cnt = n - base < kSortTile ? n - base : kSortTile;
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    const int p = k * kSortBlock + tid;
    if (p < cnt) {
      const uint32_t kk = s_key[p];
      const uint32_t d = (kk >> shift) & mask;
      const uint32_t dst = gbase[d] + ((uint32_t)p - dstart[d]);
      if (dst < (uint32_t)n) {  // always true for a consistent ranking
        kout[dst] = kk;
        vout[dst] = s_val[p];
      }
    }
  }
}

}  // namespace

int radix_tiles(int n) { return (n + kSortTile - 1) / kSortTile; }

int radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, int n, int key_bits,
                     uint32_t* hist, uint32_t* totals, hipStream_t st) {
  if (n <= 0) return 0;
  const int tiles = radix_tiles(n);
  uint32_t* kin = k0;
  uint32_t* vin = v0;
  uint32_t* kout = k1;
  uint32_t* vout = v1;
  int cur = 0;
  for (int shift = 0; shift < key_bits; shift += 8) {
    const int bits = key_bits - shift < 8 ? key_bits - shift : 8;
    k_radix_hist<<<tiles, kSortBlock, 0, st>>>(kin, n, shift, bits, tiles, hist);
    k_radix_scan_rows<<<kRadix, kSortBlock, 0, st>>>(hist, tiles, totals);
    k_radix_scatter<<<tiles, kSortBlock, 0, st>>>(kin, vin, kout, vout, n, shift, bits, tiles,
                                                  hist, totals);
    uint32_t* t = kin;
    kin = kout;
    kout = t;
    t = vin;
    vin = vout;
    vout = t;
    cur ^= 1;
  }
  return cur;
}

}  // namespace pbf
