// pbf_pass.h — the generic neighbour-pass kernels of the gfx950 PBF step,
// instantiated by each translation unit over its pass policies.
//
// A pass policy P defines, for one neighbour pass:
//   Args                 kernel arguments (device arrays)
//   Target, Acc          a target's own data; its running sum (<= 8 floats)
//   target(a, i, c)      load target i
//   screen(a)            the float4 array whose x, y, z the screening tests
//   load_f(a, j)         candidate data for the pair term
//   hit(t, d, c)         cheap test on the candidate's x, y, z d: false only
//                        if the pair adds exactly 0 (i itself has d == t.x)
//   add(acc, t, j, d, c, a) the pair term (the body of the reference's loop)
//   accumulate(tot, p)   tot += a chunk's partial sum
//   finish(a, t, s, c, g) store the result of target t
// Each neighbour pass is three kernels:
//   light   one work-item per target over the cell-sorted order; workgroups
//           remapped so each XCD's L2 serves one contiguous particle range; a
//           heavy target (pbf_device.h: kLightMax, heavy_level, kLevelShare)
//           is appended to the heavy list instead;
//   chunks  one work-item per kChunk-candidate chunk of the heavy targets;
//   combine one work-item per heavy target, adding its chunk sums in order.
// Both follow the canonical summation order of pbf_device.h.
#pragma once

#include "pbf_device.h"
#include "pbf_launch.h"

namespace pbf {
namespace {

constexpr int kBlock = 256;
constexpr int kRankBuckets = 64;  // one wavefront scans them
// Blocks of 64 candidates whose hits the light kernels evaluate in one loop.
#ifndef PBF_LIGHT_GROUP
#define PBF_LIGHT_GROUP 2
#endif
constexpr int kLightGroup = PBF_LIGHT_GROUP;
// A pass policy with kStage = true has its light kernel stage the
// candidates of each workgroup in LDS (see k_light).  Capacity: kStagePerLane
// float4 elements per work-item of the workgroup.
template <class P, class = void>
struct StageMode {
  static constexpr bool value = false;
};
template <class P>
struct StageMode<P, std::void_t<decltype(P::kStage)>> {
  static constexpr bool value = P::kStage;
};
// The stage holds x, y and z in three arrays of kStagePerLane * workgroup
// size + kStagePad floats (a screening batch reads up to 7 elements past the
// end of a target's last column, masked off).
#ifndef PBF_STAGE_PER_LANE
#define PBF_STAGE_PER_LANE 8
#endif
constexpr int kStagePerLane = PBF_STAGE_PER_LANE;
constexpr int kStagePad = 8;
inline size_t light_lds_bytes(int lb, bool staged) {
  return (size_t)10 * lb * sizeof(uint2) +
         (staged ? ((size_t)kStagePerLane * lb + kStagePad) * 12 : 0);
}

// Minimum wavefronts per SIMD the neighbour kernels are compiled for (caps
// their registers).
#ifndef PBF_PASS_WAVES
#define PBF_PASS_WAVES 1
#endif

inline int blocks_for(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// Floats per chunk partial of pass policy P: its Acc, packed (the partial
// buffer holds kPartialWords floats per slot, the largest Acc).
template <class P>
__host__ __device__ constexpr uint32_t partial_words() {
  static_assert(sizeof(typename P::Acc) % sizeof(float) == 0, "Acc must be floats");
  static_assert(sizeof(typename P::Acc) <= kPartialWords * sizeof(float), "partial too large");
  return (uint32_t)(sizeof(typename P::Acc) / sizeof(float));
}

// ---------------------------------------------------------------------------
// The hit-mask hooks of sweep_seq for target ti with m candidates under
// policy P (MaskMode, pbf_device.h).
template <class P>
struct MaskIO {
  const HeavyLists& hl;
  uint32_t ti, m;
  __device__ __forceinline__ bool kept() const {
    return hl.masks != nullptr && m <= kMaskBlocks * kChunk;
  }
  __device__ __forceinline__ bool get(uint32_t b, uint64_t& mask) const {
    if constexpr (MaskMode<P>::value == kMaskLoad) {
      if (kept()) {
        mask = hl.masks[(size_t)b * hl.mask_stride + ti];
        return true;
      }
    }
    return false;
  }
  __device__ __forceinline__ void put(uint32_t b, uint64_t mask) const {
    if constexpr (MaskMode<P>::value == kMaskStore) {
      if (kept()) hl.masks[(size_t)b * hl.mask_stride + ti] = mask;
    }
  }
};

// Generic neighbour-pass kernels over a pass policy.
// ---------------------------------------------------------------------------

// Light: one work-item per target; the workgroup size is a launch shape
// (64, 128 or 256), the column tables live in dynamic LDS.  A heavy target
// (more candidates than fixed_threshold, or when that is 0 than the
// threshold word the previous pass's combine kernel set) gets one chunk item per kChunk candidates and a
// record, appended with one atomic per wavefront each.  The counters of
// parity pass & 1 belong to this pass; the kernel clears the other parity's
// counters for the next pass (their previous users finished earlier on the
// stream).  If the chunk list is full the lane sums its target itself, still
// chunk by chunk, so the canonical order holds either way.
template <class P, class S>
__global__ __launch_bounds__(kBlock, PBF_PASS_WAVES) void k_light(typename P::Args a, int n,
                                                                  StepConsts c, S sc,
                                                  const uint32_t* __restrict__ cs, HeavyLists hl,
                                                  int pass, uint32_t fixed_threshold) {
  constexpr bool kStg = StageMode<P>::value;
  extern __shared__ uint2 seq_dyn[];
  __shared__ uint32_t key_sh[kBlock];
  __shared__ uint32_t order_sh[kBlock];
  __shared__ uint32_t bucket_sh[kRankBuckets];
  // staging: per wavefront and column k the smallest first slot and the
  // largest end slot of its light targets' column k; then per column the
  // workgroup's range [st_lo, st_lo + size) and its first stage element
  __shared__ uint32_t wlo_sh[kStg ? kBlock / 64 : 1][9], whi_sh[kStg ? kBlock / 64 : 1][9];
  __shared__ uint32_t st_lo[9], st_base[10];
  __shared__ int st_ok;
  uint32_t* rec_cnt = hl.counters + rec_counter(pass & 1);
  uint32_t* item_cnt = hl.counters + item_counter(pass & 1);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    hl.counters[rec_counter((pass + 1) & 1)] = 0u;
    hl.counters[item_counter((pass + 1) & 1)] = 0u;
  }
  const uint32_t word = hl.counters[kThresholdWord];
  const uint32_t thr = fixed_threshold ? fixed_threshold : (word ? word : kLightMax);
  const int tid = threadIdx.x;
  const int nb = (int)blockDim.x;
  const int block_base = (P::kXcdRanges ? xcd_block_index() : (int)blockIdx.x) * nb;
  const int i = block_base + tid;
  const bool active = i < n;
  const SeqTable st{&seq_dyn[tid], nb};
  typename P::Target t{};
  uint32_t m = 0;
  Cols cl;
  if (active) {
    t = P::target(a, (uint32_t)i, c);
    const SceneGeo g = sc.at((uint32_t)i, c);
    const Cell q = cell_of(t.x.x, t.x.y, t.x.z, g, c.h);
    if constexpr (kStg) {
      m = load_cols(cl, q, cs, g);
    } else {
      m = build_seq(st, q, cs, g);
    }
  } else if constexpr (kStg) {
#pragma unroll
    for (int k = 0; k < 9; ++k) cl.lo[k] = cl.len[k] = 0u;
  }
  {
    // this pass's statistics for the next pass's threshold: per level, how
    // many of the wavefront's targets have level < m <= kLightMax
    unsigned long long w[kLevelWords] = {};
#pragma unroll
    for (int k = 0; k < kLevels; ++k) {
      const uint64_t b = __ballot(m > heavy_level(k) && m <= kLightMax);
      w[k / 3] += (unsigned long long)__popcll(b) << (kLevelBits * (k % 3));
    }
    if ((threadIdx.x & 63) == 0) {
      const int wave = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
      unsigned long long* slot =
          reinterpret_cast<unsigned long long*>(hl.counters + stat_slot(wave % kStatSlots));
#pragma unroll
      for (int j = 0; j < kLevelWords; ++j)
        if (w[j]) atomicAdd(&slot[j], w[j]);
    }
  }
  const bool is_heavy = active && m > thr;
  const uint32_t nch = is_heavy ? (m + kChunk - 1) / kChunk : 0u;
  bool in_lane = false;  // chunk list full: this lane sums its chunks itself
  if (__ballot(is_heavy) != 0ull) {
    // Wave-aggregated allocation.  Partials are target-major (lane l owns
    // [first, first + nch)); the work items are chunk-major over the wave
    // (chunk 0 of every heavy lane, then chunk 1, ...), so the lanes of a
    // chunk kernel wavefront work on the same chunk of neighbouring targets,
    // which in a crowded cell share their candidate sequence: the candidate
    // loads coalesce instead of being gathered lane by lane.
    uint32_t item_total;
    const uint32_t item_off = wave_exclusive_sum(nch, &item_total);
    uint32_t item_base = 0, rec_base = 0;
    if ((threadIdx.x & 63) == 0) item_base = atomicAdd(item_cnt, item_total);
    item_base = __shfl(item_base, 0);
    const uint32_t first = item_base + item_off;
    const bool wave_fits = item_base + item_total <= hl.item_cap;
    uint32_t maxnch = nch;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const uint32_t o = __shfl_xor(maxnch, off);
      maxnch = o > maxnch ? o : maxnch;
    }
    const uint64_t lt = (1ull << (threadIdx.x & 63)) - 1ull;
    uint32_t pos = item_base;
    for (uint32_t ch = 0; ch < maxnch; ++ch) {
      const bool has = nch > ch;
      const uint64_t bm = __ballot(has);
      const uint32_t u = pos + (uint32_t)__popcll(bm & lt);
      if (has && u < hl.item_cap)
        hl.items[u] = wave_fits ? make_uint4((uint32_t)i, ch, first + ch, 0u)
                                : make_uint4(~0u, 0u, 0u, 0u);
      pos += (uint32_t)__popcll(bm);
    }
    const uint64_t hm = __ballot(is_heavy);
    if (wave_fits) {
      if ((threadIdx.x & 63) == 0) rec_base = atomicAdd(rec_cnt, (uint32_t)__popcll(hm));
      rec_base = __shfl(rec_base, 0);
      if (is_heavy)
        hl.recs[rec_base + (uint32_t)__popcll(hm & lt)] = make_uint4((uint32_t)i, first, nch, 0u);
    } else {
      in_lane = is_heavy;
    }
  }
  // Lanes take the block's remaining targets in order of candidate count, so
  // each wavefront gets targets of similar cost instead of waiting on its
  // longest lane: a counting sort over buckets of 8 candidates in LDS
  // (bucket 0: idle slots and targets handed to the chunk kernels).  The
  // targets stay in this block's particle range, whose neighbourhoods
  // overlap.  Each target's sum is unchanged: only the lane that computes it
  // moves, so the order within a bucket (set by LDS atomics) does not matter.
  const uint32_t key = active && (!is_heavy || in_lane) ? m + 1u : 0u;
  key_sh[tid] = key;
  if (tid < kRankBuckets) bucket_sh[tid] = 0u;
  if constexpr (kStg) {
    // the workgroup's candidate ranges, column by column, over its light
    // targets (a heavy target summed in its lane reads global memory)
    const bool light = active && !is_heavy;
    const int wave = tid >> 6;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const bool has = light && cl.len[k] != 0u;
      uint32_t lo = has ? cl.lo[k] : 0xFFFFFFFFu;
      uint32_t hi = has ? cl.lo[k] + cl.len[k] : 0u;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const uint32_t ol = __shfl_xor(lo, off), oh = __shfl_xor(hi, off);
        lo = ol < lo ? ol : lo;
        hi = oh > hi ? oh : hi;
      }
      if ((tid & 63) == 0) {
        wlo_sh[wave][k] = lo;
        whi_sh[wave][k] = hi;
      }
    }
  }
  __syncthreads();
  if constexpr (kStg) {
    if (tid == 0) {
      uint32_t total = 0;
      for (int k = 0; k < 9; ++k) {
        uint32_t lo = 0xFFFFFFFFu, hi = 0u;
        for (int w = 0; w < (nb >> 6); ++w) {
          lo = wlo_sh[w][k] < lo ? wlo_sh[w][k] : lo;
          hi = whi_sh[w][k] > hi ? whi_sh[w][k] : hi;
        }
        st_lo[k] = lo;
        st_base[k] = total;
        total += hi > lo ? hi - lo : 0u;
      }
      st_base[9] = total;
      st_ok = total <= (uint32_t)(kStagePerLane * nb);
    }
  }
  const uint32_t bucket =
      key == 0u ? 0u : 1u + min((key - 1u) >> 3, (uint32_t)kRankBuckets - 2u);
  const uint32_t within = atomicAdd(&bucket_sh[bucket], 1u);
  __syncthreads();
  const uint32_t stage_n = (uint32_t)(kStagePerLane * nb + kStagePad);
  [[maybe_unused]] float* const stage = reinterpret_cast<float*>(seq_dyn + 10 * nb);
  if constexpr (kStg) {
    // column tables: a light target's columns as stage elements when the
    // workgroup's ranges fit the stage, else (and for a heavy target summed
    // in its lane) as global slots; then the ranges are copied in,
    // consecutive work-items taking consecutive elements
    const bool ok = st_ok != 0;
    if (active) {
      const bool staged = ok && !is_heavy;
      write_seq(st, cl, [&](int k, uint32_t lo) {
        return staged ? st_base[k] + (lo - st_lo[k]) : lo;
      });
    }
    if (ok) {
      const uint32_t total = st_base[9];
      const float4* __restrict__ src_x = P::screen(a);
      for (uint32_t e = (uint32_t)tid; e < total; e += (uint32_t)nb) {
        int k = 0;
#pragma unroll
        for (int q = 1; q < 9; ++q) k += e >= st_base[q] ? 1 : 0;
        const float4 v = src_x[st_lo[k] + (e - st_base[k])];
        stage[e] = v.x;
        stage[stage_n + e] = v.y;
        stage[2 * stage_n + e] = v.z;
      }
    }
  }
  if (tid < kRankBuckets) {  // the first wavefront scans the bucket counts
    uint32_t total;
    const uint32_t cnt_b = bucket_sh[tid];
    bucket_sh[tid] = wave_exclusive_sum(cnt_b, &total);
  }
  __syncthreads();
  order_sh[bucket_sh[bucket] + within] = (uint32_t)tid;
  __syncthreads();
  const int src = (int)order_sh[tid];
  const uint32_t ks = key_sh[src];
  if (ks == 0u) return;  // idle slot, or a target handed to the chunk kernels
  const uint32_t ms = ks - 1u;
  const uint32_t ti = (uint32_t)(block_base + src);
  const SeqTable sts{&seq_dyn[src], nb};
  t = P::target(a, ti, c);
  auto hit = [&](const float3& d) { return P::hit(t, d, c); };
  // the sweep's blocks are the canonical chunks: each is summed from zero
  // and added to the total in order (also for a heavy target whose chunk
  // list was full)
  typename P::Acc s, part;
  const MaskIO<P> mio{hl, ti, ms};
  auto add = [&](uint32_t j, const auto& d) { P::add(part, t, j, d, c, a); };
  auto block_end = [&] {
    P::accumulate(s, part);
    part = typename P::Acc{};
  };
  auto get_mask = [&](uint32_t b, uint64_t& mk) { return mio.get(b, mk); };
  auto put_mask = [&](uint32_t b, uint64_t mk) { mio.put(b, mk); };
  if constexpr (kStg) {
    if (st_ok != 0 && ms <= thr) {
      // the candidates are stage elements (slot = element index); add()
      // sees the element index as j (pass policies with kStage use j only
      // for the self test, which their distance test already decides)
      sweep_seq<P::kHitBatch, kLightGroup>(
          sts, ms, 0u, ms,
          [&](uint32_t off) {
            const uint32_t e = off >> 4;
            return make_float3(stage[e], stage[stage_n + e], stage[2 * stage_n + e]);
          },
          [&](uint32_t j) {
            return make_float4(stage[j], stage[stage_n + j], stage[2 * stage_n + j], 0.0f);
          },
          hit, add, block_end, get_mask, put_mask);
      P::finish(a, t, s, c, sc.at(ti, c));
      return;
    }
  }
  const Float4Buf sb = float4_buf(P::screen(a), (uint32_t)n);
  // a pass on stored hit masks screens nothing: its hits of up to four
  // blocks run through one loop (kLightMax candidates), which balances the
  // lanes best
  constexpr int kGroup = MaskMode<P>::value == kMaskLoad ? (int)(kLightMax / kChunk) : kLightGroup;
  sweep_seq<P::kHitBatch, kGroup>(
      sts, ms, 0u, ms, [&](uint32_t off) { return load_xyz_at(sb, off); },
      [&](uint32_t j) { return P::load_f(a, j); }, hit, add, block_end, get_mask, put_mask);
  P::finish(a, t, s, c, sc.at(ti, c));
}

// One work-item per chunk item (i, chunk, partial slot): the partial sum of
// kChunk consecutive candidates of a heavy target.  Grid-stride over the
// items; the grid size is a launch shape.
template <class P, class S>
__global__ __launch_bounds__(kBlock, PBF_PASS_WAVES) void k_chunks(typename P::Args a, int n,
                                                                   StepConsts c, S sc,
                                                   const uint32_t* __restrict__ cs, HeavyLists hl,
                                                   int pass) {
  __shared__ uint2 seq[10][kBlock];
  const uint32_t items = hl.counters[item_counter(pass & 1)];
  const uint32_t total = items < hl.item_cap ? items : hl.item_cap;
  // only as many workgroups as there are items to go round, spread over all
  // XCDs, each XCD on one contiguous range of the (chunk-major) item list
  const uint32_t need = (total + kBlock - 1) / kBlock;
  const int active = need < gridDim.x ? (int)need : (int)gridDim.x;
  const int vb = P::kXcdRanges ? xcd_block_index_of(active)
                               : ((int)blockIdx.x < active ? (int)blockIdx.x : -1);
  if (vb < 0) return;
  const uint32_t stride = (uint32_t)active * kBlock;
  const SeqTable st{&seq[0][threadIdx.x], kBlock};
  const Float4Buf sb = float4_buf(P::screen(a), (uint32_t)n);
  for (uint32_t u = (uint32_t)vb * kBlock + threadIdx.x; u < total; u += stride) {
    const uint4 it = hl.items[u];
    if (it.x == ~0u) continue;
    const typename P::Target t = P::target(a, it.x, c);
    const SceneGeo g = sc.at(it.x, c);
    const uint32_t m = build_seq(st, cell_of(t.x.x, t.x.y, t.x.z, g, c.h), cs, g);
    typename P::Acc part;
    const uint32_t s0 = it.y * kChunk;
    const uint32_t s1 = s0 + kChunk < m ? s0 + kChunk : m;
    const MaskIO<P> mio{hl, it.x, m};
    sweep_seq<P::kHitBatch, 1>(
        st, m, s0, s1, [&](uint32_t off) { return load_xyz_at(sb, off); },
        [&](uint32_t j) { return P::load_f(a, j); },
        [&](const float3& d) { return P::hit(t, d, c); },
        [&](uint32_t j, const auto& d) { P::add(part, t, j, d, c, a); }, [] {},
        [&](uint32_t b, uint64_t& mk) { return mio.get(b, mk); },
        [&](uint32_t b, uint64_t mk) { mio.put(b, mk); });
    *reinterpret_cast<typename P::Acc*>(hl.partials + (size_t)it.z * partial_words<P>()) = part;
  }
}

// One work-item per heavy target: chunk partials added in chunk order.
template <class P, class S>
__global__ __launch_bounds__(kBlock) void k_combine(typename P::Args a, int n, StepConsts c, S sc,
                                                    HeavyLists hl, int pass) {
  const uint32_t nrec = hl.counters[rec_counter(pass & 1)];
  if (blockIdx.x == 0 && threadIdx.x < 64) {
    // the next pass's heavy threshold from this pass's statistics, which are
    // cleared for the next pass
    unsigned long long* slot =
        reinterpret_cast<unsigned long long*>(hl.counters + stat_slot((int)threadIdx.x));
    unsigned long long cnt[kLevels];
#pragma unroll
    for (int j = 0; j < kLevelWords; ++j) {
      const unsigned long long w = slot[j];
      slot[j] = 0ull;
#pragma unroll
      for (int f = 0; f < 3 && 3 * j + f < kLevels; ++f)
        cnt[3 * j + f] = (w >> (kLevelBits * f)) & ((1ull << kLevelBits) - 1ull);
    }
#pragma unroll
    for (int k = 0; k < kLevels; ++k)
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) cnt[k] += __shfl_xor(cnt[k], off);
    if (threadIdx.x == 0) {
      uint32_t thr = kLightMax;
#pragma unroll
      for (int k = kLevels - 1; k >= 0; --k)
        if (cnt[k] * kLevelShare <= (unsigned long long)n) thr = heavy_level(k);
      hl.counters[kThresholdWord] = thr;
    }
  }
  const uint32_t stride = gridDim.x * kBlock;
  for (uint32_t r = blockIdx.x * kBlock + threadIdx.x; r < nrec; r += stride) {
    const uint4 rec = hl.recs[r];
    const typename P::Target t = P::target(a, rec.x, c);
    typename P::Acc s;
    using PAcc = typename P::Acc;
    const float* base = hl.partials + (size_t)rec.y * partial_words<P>();
    auto part = [&](uint32_t k) {
      return *reinterpret_cast<const PAcc*>(base + (size_t)k * partial_words<P>());
    };
    // eight loads in flight, then the adds in chunk order
    uint32_t k = 0;
    for (; k + 8u <= rec.z; k += 8u) {
      PAcc p8[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) p8[q] = part(k + q);
#pragma unroll
      for (int q = 0; q < 8; ++q) P::accumulate(s, p8[q]);
    }
    for (; k < rec.z; ++k) P::accumulate(s, part(k));
    P::finish(a, t, s, c, sc.at(rec.x, c));
  }
}

int grid_for_items(long long items, int max_blocks) {
  const long long b = (items + kBlock - 1) / kBlock;
  return (int)(b < 1 ? 1 : (b > max_blocks ? max_blocks : b));
}

template <class P, class S>
void launch_pass_s(const typename P::Args& a, int n, const StepConsts& c, const S& sc,
                   const uint32_t* cs, const HeavyLists& hl, int pass, const Shapes& sh,
                   hipStream_t st) {
  const int lb = sh.light_block;
  k_light<P, S><<<(n + lb - 1) / lb, lb, light_lds_bytes(lb, StageMode<P>::value), st>>>(
      a, n, c, sc, cs, hl, pass, (uint32_t)sh.heavy);
  k_chunks<P, S><<<grid_for_items(hl.item_cap, sh.chunk_blocks), kBlock, 0, st>>>(a, n, c, sc,
                                                                                    cs, hl, pass);
  k_combine<P, S><<<grid_for_items(n, 512), kBlock, 0, st>>>(a, n, c, sc, hl, pass);
}

template <class P>
void launch_pass(const typename P::Args& a, int n, const StepConsts& c, const Scenes& scn,
                 const uint32_t* cs, const HeavyLists& hl, int pass, const Shapes& sh,
                 hipStream_t st) {
  if (scn.geo)
    launch_pass_s<P>(a, n, c, SceneTable{scn.geo, scn.sid}, cs, hl, pass, sh, st);
  else
    launch_pass_s<P>(a, n, c, OneScene{}, cs, hl, pass, sh, st);
}

}  // namespace
}  // namespace pbf
