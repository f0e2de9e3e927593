// pbf_device.h — shared device-side definitions of the gfx950 PBF step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace pbf {

// Geometry of one scene's box and grid.  A batch of scenes shares one global
// cell-id space: scene s owns ids [base, base + dx*dy*dz), so one stable sort
// of all particles by cell id leaves every scene's particles in its own
// contiguous range, in the order sorting that scene alone would give.
struct alignas(16) SceneGeo {
  float llx, lly, llz;  // box lower corner (llim)
  float ulx, uly, ulz;  // box upper corner (ulim)
  float lox, loy, loz;  // box clamp lower: llim + LIM_EPS, rounded to FP32
  float hix, hiy, hiz;  // box clamp upper: ulim - LIM_EPS, rounded to FP32
  int dx, dy, dz;       // grid dims, (int)ceilf((ulim-llim)/h)
  uint32_t base;        // first global cell id of the scene
};

// Per-step constants, built on the host once per step and passed by value
// (kernarg segment).  Field meanings follow the reference's functors
// (fluids/Simulator.cu:22-161) and constants (fluids/helper.h:4-9).
struct StepConsts {
  SceneGeo g;              // the scene of a single-scene step
  int cells;               // cells of all scenes
  float h, h2;             // kernel radius, h*h
  float h2_outer;          // prefilter bound: r2 >= h2_outer => both kernels are 0
  float keps;              // KERNAL_EPS
  float keps2_lo;          // 0.99 keps^2: r2 < keps2_lo => sqrtf(r2) < keps (no gradient)
  float poly6;             // 315/(64 pi h^9)
  float spiky;             // -45/(pi h^6)
  float spiky_rho0;        // spiky / pho0 (the lambda gradient, Simulator_kernel.cuh:92)
  float pho0;
  float lambda_eps;
  float coef_corr;         // -k_corr / W(delta_q)^n_corr
  float n_corr;
  int n_corr_int;          // n_corr when it is an integer in [0,16], else -1
  float k_boundary;        // k_boundaryDensity
  float wall_c;            // 2 pi / 3 (DensityBoundary)
  float max_dp;            // MAX_DP
  float dt, inv_dt;        // dt, 1/dt (FP32, as h_updateVelocity)
  float gz;                // -g
  float c_xsph;
};

// Scene sources: where a kernel finds the geometry of particle i's scene.
// OneScene: the single scene in the kernarg constants.  SceneTable: a batch;
// sid[i] is the scene of index i, valid in input and in sorted order alike
// (both keep scene s in its own index range).
struct OneScene {
  __device__ __forceinline__ SceneGeo at(uint32_t, const StepConsts& c) const { return c.g; }
};
struct SceneTable {
  const SceneGeo* __restrict__ geo;
  const uint16_t* __restrict__ sid;
  __device__ __forceinline__ SceneGeo at(uint32_t i, const StepConsts&) const {
    return geo[sid[i]];
  }
};

// Workgroup index remapped so that each of the 8 XCDs (which receive
// workgroups round-robin and each own an L2) processes one contiguous range
// of particles: neighbouring particles, which share neighbour data, then
// share an L2.
__device__ __forceinline__ int xcd_block_index() {
  const int G = (int)gridDim.x;
  const int p = (int)blockIdx.x;
  const int per = G >> 3, rem = G & 7, x = p & 7;
  return x * per + (x < rem ? x : rem) + (p >> 3);
}

// The same remap over the first `active` blocks of the grid only (blocks
// from `active` on get -1).  For a grid-stride kernel whose work fills only a
// prefix of the grid: remapping over the whole grid would hand that prefix to
// the first XCDs and leave the others idle.
__device__ __forceinline__ int xcd_block_index_of(int active) {
  const int p = (int)blockIdx.x;
  if (p >= active) return -1;
  const int per = active >> 3, rem = active & 7, x = p & 7;
  return x * per + (x < rem ? x : rem) + (p >> 3);
}

// One axis of the cell coordinate, (int)((p - llim)/h) clamped to
// [0, dim-1] in FP32 (fluids/Simulator.cu:30-36).  Clamping the quotient
// before truncating gives the same integer for every finite quotient and
// maps NaN to 0.  The division is IEEE (built with
// -fhip-fp32-correctly-rounded-divide-sqrt) so keys match the host oracle.
__device__ __forceinline__ int cell_coord(float p, float ll, float h, int dim) {
  const float q = (p - ll) / h;
  return (int)fminf(fmaxf(q, 0.0f), (float)(dim - 1));
}

struct Cell {
  int x, y, z;
};

__device__ __forceinline__ Cell cell_of(float x, float y, float z, const SceneGeo& g, float h) {
  return Cell{cell_coord(x, g.llx, h, g.dx), cell_coord(y, g.lly, h, g.dy),
              cell_coord(z, g.llz, h, g.dz)};
}

// Cell id, x-major and z fastest (xyzToId, fluids/Simulator.cu:52), offset
// by the scene's base.
__device__ __forceinline__ uint32_t cell_id(const Cell& q, const SceneGeo& g) {
  return g.base + (uint32_t)((q.x * g.dy + q.y) * g.dz + q.z);
}

// Canonical summation order, shared with the host oracle's FP32 mirror: a
// target's candidate sequence is summed in consecutive chunks of kChunk
// candidates, each summed sequentially from zero, the chunk sums added in
// chunk order (one chunk: the plain sequential sum).  The order does not
// depend on who computes the sum, so the kernels may split the work freely:
// a light target's work-item walks all its chunks; a heavy target's chunks
// go to one work-item each (the chunk kernel) and are added up by the
// combine kernel.  Crowded cells (particles piled into a corner) then cost a
// few chunk lengths instead of one lane thousands of dependent gathers.
constexpr uint32_t kChunk = 64;
// A target is heavy when it has more than kLightMax candidates, or more than
// the threshold the previous neighbour pass chose: the smallest level L of
// heavy_level() such that at most n / kLevelShare targets had L < m <=
// kLightMax (else kLightMax).  A light wavefront lasts as long as its longest
// lane, so in sparse water the few long lanes set the light kernel's tail and
// are better spread over the chunk kernel (threshold 64); in dense water long
// lanes are common, balance each other within the wavefronts, and spreading
// them would only lengthen the chunk lists (threshold up to kLightMax).
// Which targets are heavy changes no result (the canonical order above),
// only the timing.
constexpr uint32_t kLightMax = 256;
constexpr int kLevels = 7;
__host__ __device__ constexpr uint32_t heavy_level(int k) {  // 64, 80, .., 128, 160, 192
  return k < 4 ? 64u + 16u * (uint32_t)k : 128u + 32u * (uint32_t)(k - 4);
}
constexpr uint32_t kLevelShare = 25;
// The per-level counts travel packed, three 21-bit fields to a 64-bit word
// (exact while a statistics slot counts fewer than 2^21 targets per level,
// i.e. below 2^27 targets; beyond that a count may spill into the next field
// and misjudge the threshold, which costs time only).
constexpr int kLevelBits = 21;
constexpr int kLevelWords = (kLevels + 2) / 3;

// Hit masks shared by the two passes of one solver iteration.  lambda_k and
// delta-p_k read the same positions x_k and walk the same candidate sequence,
// and every pair delta-p adds (KERNAL_EPS <= r < h, j != i) passes lambda's
// screening test (r^2 < h^2 padded): a pass policy with kMask = kMaskStore
// (lambda) records the 64-bit hit mask of each 64-candidate block of targets
// with at most kMaskBlocks blocks (2048 candidates: all but the deepest corner
// piles), and one with kMask = kMaskLoad (delta-p)
// uses those masks instead of screening.  Its add() still tests every pair it
// is handed, so the extra bits (coincident pairs, i itself) add nothing and
// the sums are those of its own screening.
constexpr int kMaskNone = 0, kMaskStore = 1, kMaskLoad = 2;
constexpr uint32_t kMaskBlocks = 32;
template <class P, class = void>
struct MaskMode {
  static constexpr int value = kMaskNone;
};
template <class P>
struct MaskMode<P, std::void_t<decltype(P::kMask)>> {
  static constexpr int value = P::kMask;
};

// A float4 array as a raw buffer resource, addressed by byte offsets: a
// 32-bit offset per load instead of a 64-bit address, and loads past the
// array's n elements return 0 instead of touching other memory.  The loads
// go through the LLVM buffer-load intrinsic with an explicit 3 x float
// result (12 bytes, buffer_load_dwordx3).
typedef int pbf_i32x4 __attribute__((ext_vector_type(4)));
typedef float pbf_f32x3 __attribute__((ext_vector_type(3)));
__device__ pbf_f32x3 pbf_raw_buffer_load_f32x3(pbf_i32x4 rsrc, int voffset, int soffset,
                                               int aux) __asm("llvm.amdgcn.raw.buffer.load.v3f32");
struct Float4Buf {
  pbf_i32x4 r;
};
__device__ __forceinline__ Float4Buf float4_buf(const float4* p, uint32_t n) {
  const uint64_t bytes = (uint64_t)n * 16u;
  const uint64_t a = (uint64_t)p;
  Float4Buf b;
  b.r.x = (int)(uint32_t)(a & 0xFFFFFFFFull);
  b.r.y = (int)(uint32_t)(a >> 32);  // stride 0
  b.r.z = (int)(bytes > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)bytes);
  b.r.w = 0x00020000;  // the raw (untyped) buffer format of gfx9
  return b;
}
// x, y, z of the element at byte offset `off` (a 12-byte load).
__device__ __forceinline__ float3 load_xyz_at(const Float4Buf& b, uint32_t off) {
  const pbf_f32x3 v = pbf_raw_buffer_load_f32x3(b.r, (int)off, 0, 0);
  return make_float3(v.x, v.y, v.z);
}

// Entries of a column table: nine columns and the closing entry.
constexpr int kSeqEntries = 10;

// A work-item's candidate sequence as a compact column table in LDS: entry
// k < nk is (first slot of the k-th non-empty column, sequence position of
// its first candidate); entry nk is (0, m).  `tab` points at entry 0 of this
// work-item and `stride` separates entries (LDS laid out [entry][lane] so
// lanes touch consecutive words).
struct SeqTable {
  uint2* tab;
  int stride;
  __device__ __forceinline__ uint2 at(int k) const { return tab[k * stride]; }
};

// Candidate neighbours of a particle whose query cell is q, in the
// reference's order (fluids/Simulator_kernel.cuh:78-88): dx, then dy, then
// dz, j ascending.  The three cells (x,y,z-1..z+1) of one column are
// consecutive cell ids, so their particles form one contiguous sorted range,
// read from the cell-start table cs (cells + 1 entries; cs[c] = first sorted
// slot whose cell id is >= c): [cs[first cell], cs[last cell + 1]).  Walking
// that range is the same j sequence as walking the three cells in turn.
// build_seq builds the table from the nine column ranges around query cell q
// (column k = 3*(dx+1) + (dy+1); outside the grid: empty) and returns the
// candidate count m.  The 18 cell-start loads are issued
// together (a column outside the grid reads a valid dummy address and is
// then treated as empty), so building the table costs one memory round trip,
// not one per column.
// The nine candidate columns of query cell q as (first slot, length), length
// 0 for a column outside the grid or empty; the 18 cell-start loads are
// issued together.
struct Cols {
  uint32_t lo[9], len[9];
};
__device__ __forceinline__ uint32_t load_cols(Cols& cl, const Cell& q,
                                              const uint32_t* __restrict__ cs,
                                              const SceneGeo& g) {
  uint32_t hi[9];
  bool in[9];
  const int z0 = q.z > 0 ? q.z - 1 : 0;
  const int z1 = q.z + 1 < g.dz ? q.z + 1 : g.dz - 1;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int X = q.x + k / 3 - 1;
    const int Y = q.y + k % 3 - 1;
    in[k] = (X >= 0) & (X < g.dx) & (Y >= 0) & (Y < g.dy);
    const uint32_t col = g.base + (in[k] ? (uint32_t)((X * g.dy + Y) * g.dz) : 0u);
    cl.lo[k] = cs[col + z0];
    hi[k] = cs[col + z1 + 1];
  }
  uint32_t m = 0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    cl.len[k] = (in[k] & (hi[k] > cl.lo[k])) ? hi[k] - cl.lo[k] : 0u;
    m += cl.len[k];
  }
  return m;
}

// Writes the column table of `cl`, column k's first slot mapped by
// slot_of(k, first slot), and returns the candidate count m.  Branch-free
// compaction: every column writes its entry at the next free index, which
// only advances past non-empty columns (an empty column's entry is
// overwritten by the next one or by the closing entry).
template <class SlotOf>
__device__ __forceinline__ uint32_t write_seq(const SeqTable& st, const Cols& cl,
                                              SlotOf&& slot_of) {
  uint32_t m = 0;
  int nk = 0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    st.tab[nk * st.stride] = make_uint2(slot_of(k, cl.lo[k]), m);
    nk += cl.len[k] ? 1 : 0;
    m += cl.len[k];
  }
  st.tab[nk * st.stride] = make_uint2(0u, m);
  return m;
}

__device__ __forceinline__ uint32_t build_seq(const SeqTable& st, const Cell& q,
                                              const uint32_t* __restrict__ cs,
                                              const SceneGeo& g) {
  Cols cl;
  load_cols(cl, q, cs, g);
  return write_seq(st, cl, [](int, uint32_t lo) { return lo; });
}

// Maps increasing sequence positions to sorted slots (positions must be < m
// and must not decrease between calls).  The cursor holds the current column
// (first slot b0, positions [c0, n0)) and the next one (first slot b1,
// positions [n0, n1)); a column exists after the current one exactly when
// n0 < m, so the lookahead never reads past the table's closing entry.
struct SeqCursor {
  const uint2* e;  // the current column's entry
  int stride;      // entries apart in LDS
  uint32_t m, b0, c0, n0, b1, n1;
  // With no column after the current one, b1 = n0 makes offsets8 map the
  // positions past the sequence's end to slots p (< m + 8): in range of a
  // staged workgroup's padded stage as well as of the particle arrays.
  __device__ __forceinline__ void load_next() {
    if (n0 < m) {
      b1 = e[stride].x;
      n1 = e[2 * stride].y;
    } else {
      b1 = n0;
      n1 = n0;
    }
  }
  __device__ __forceinline__ void init(const SeqTable& t, uint32_t u, uint32_t total) {
    e = t.tab;
    stride = t.stride;
    m = total;
    if (u != 0u) {
      // starting inside the sequence (a chunk item): the column holding u
      // from all entries' first positions, read together.  Entries 0..nk-1
      // are columns (first position < m), entry nk closes the table with m;
      // entries after it are stale, so only the run of entries below m
      // counts.
      uint32_t y[kSeqEntries];
#pragma unroll
      for (int k = 0; k < kSeqEntries; ++k) y[k] = e[k * stride].y;
      int col = 0;
      bool live = true;
#pragma unroll
      for (int k = 1; k < kSeqEntries; ++k) {
        live = live && y[k - 1] < total;
        col += (live && y[k] <= u) ? 1 : 0;
      }
      e += col * stride;
    }
    const uint2 e0 = e[0], e1 = e[stride];
    b0 = e0.x;
    c0 = e0.y;
    n0 = e1.y;
    load_next();
  }
  __device__ __forceinline__ void shift() {
    e += stride;
    b0 = b1;
    c0 = n0;
    n0 = n1;
    load_next();
  }
  __device__ __forceinline__ uint32_t slot(uint32_t u) {
    while (u >= n0) shift();
    return b0 + (u - c0);
  }
  // Byte offsets (16-byte elements) of positions p0 .. p0+7: off[t] - 16 t
  // is the offset of the current column's element for t < n0 - p0 and of
  // the next column's beyond, so each position costs one select.
  // Positions past the batch's last valid one get offsets the caller's
  // valid-bit mask discards: at most 7 elements past the end of the last
  // column or of the sequence's slot range (see load_next).  A batch
  // spanning more than two columns takes the per-position slot walk.
  __device__ __forceinline__ void offsets8(uint32_t p0, uint32_t last_valid, uint32_t off[8]) {
    const uint32_t last = p0 + 7u < last_valid ? p0 + 7u : last_valid;
    if (last < n1) {
      const uint32_t s = n0 - p0;  // positions of the current column in the batch
      const uint32_t oa = (b0 + (p0 - c0)) << 4;
      const uint32_t ob = (b1 + p0 - n0) << 4;  // wraps for t < s: not selected
#pragma unroll
      for (int t = 0; t < 8; ++t) off[t] = ((uint32_t)t < s ? oa : ob) + 16u * (uint32_t)t;
      if (last >= n0) shift();
    } else {
#pragma unroll
      for (int t = 0; t < 8; ++t) off[t] = slot(p0 + t < last_valid ? p0 + t : last_valid) << 4;
    }
  }
};

// The reload cursor: positions only (no lookahead), fewer registers; the
// slot of position u in the current column is u + delta.
struct SeqCursor1 {
  const uint2* e;
  int stride;
  uint32_t delta, next;
  __device__ __forceinline__ void from(const SeqCursor& c) {
    e = c.e;
    stride = c.stride;
    delta = c.b0 - c.c0;
    next = c.n0;
  }
  __device__ __forceinline__ uint32_t slot(uint32_t u) {
    while (u >= next) {
      e += stride;
      const uint2 e0 = e[0];
      delta = e0.x - e0.y;
      next = e[stride].y;
    }
    return u + delta;
  }
};

// The hit mask of the candidates [bb, be) (at most 64, one canonical block)
// of a target's sequence: eight screening loads in flight at a time, bit
// u - bb set where hit() holds.  c1 walks forward only.
template <class LoadS, class Hit>
__device__ __forceinline__ uint64_t screen_block(SeqCursor& c1, uint32_t bb, uint32_t be,
                                                 LoadS&& load_s, Hit&& hit) {
  uint64_t mask = 0;
#pragma unroll 1
  for (uint32_t u = bb; u < be; u += 8) {
    uint32_t off[8];
    c1.offsets8(u, be - 1u, off);
    float3 d[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) d[t] = load_s(off[t]);
    uint32_t bits = 0;
#pragma unroll
    for (int t = 0; t < 8; ++t) bits |= (hit(d[t]) ? 1u : 0u) << t;
    const uint32_t valid = be - u >= 8u ? 0xFFu : (1u << (be - u)) - 1u;
    mask |= (uint64_t)(bits & valid) << (u - bb);
  }
  return mask;
}

// Visits sequence positions [u0, u1) of a candidate sequence of m positions
// in order,
// calling add() only for candidates passing the cheap test hit() on their
// x, y, z (load_s(byte offset of the slot's 16-byte element)).  The
// positions are taken in groups
// of NB blocks of 64: each block of the group is screened eight loads at a
// time (across column boundaries, so a sparse neighbourhood costs a few
// memory round trips, not one or two per column) into a 64-bit hit mask,
// then the hits of the whole group are reloaded G at a time (a pass
// policy's kHitBatch) and accumulated in ascending order, with block_end()
// called once per block, in order, after its last hit.  Running the hits of
// several blocks through one loop keeps the lanes of a wavefront busy: they
// wait on the largest hit count of the group instead of that of every
// block.  Candidates failing hit() contribute exactly nothing in add(), so
// the sums equal the plain in-order sum.
// The screening is branch-free after the loads (all eight tests are
// evaluated and masked), so the compiler keeps the eight loads in flight
// together instead of sinking some of them into conditional blocks that
// wait for each load in turn; hit() must be branch-free as well.
// get_mask(b, mask) may supply block b's hit mask (b = position / 64; u0
// must be a multiple of 64) instead of screening it; put_mask(b, mask)
// receives each mask the sweep screened.
template <int G, int NB, class LoadS, class LoadF, class Hit, class Add, class BlockEnd,
          class GetMask, class PutMask>
__device__ __forceinline__ void sweep_seq(const SeqTable& st, uint32_t m, uint32_t u0,
                                          uint32_t u1, LoadS&& load_s, LoadF&& load_f,
                                          Hit&& hit, Add&& add, BlockEnd&& block_end,
                                          GetMask&& get_mask, PutMask&& put_mask) {
  if (u0 >= u1) return;
  SeqCursor c1;
  c1.init(st, u0, m);
  SeqCursor1 c2;
  c2.from(c1);
#pragma unroll 1
  for (uint32_t gb = u0; gb < u1; gb += 64u * NB) {
    uint64_t w[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const uint32_t bb = gb + 64u * (uint32_t)q;
      uint64_t mask = 0;
      if (bb < u1 && !get_mask(bb >> 6, mask)) {
        const uint32_t be = u1 - bb < 64u ? u1 : bb + 64u;
        mask = screen_block(c1, bb, be, load_s, hit);
        put_mask(bb >> 6, mask);
      }
      w[q] = mask;
    }
    const uint32_t left = (u1 - gb + 63u) >> 6;
    const uint32_t nblk = left < (uint32_t)NB ? left : (uint32_t)NB;
    // the hits of the group: w[0] holds block q's remaining hits (the words
    // move down as blocks run out)
    uint32_t q = 0, done = 0, last = gb;
#pragma unroll 1
    while (true) {
      uint64_t any = w[0];
#pragma unroll
      for (int k = 1; k < NB; ++k) any |= w[k];
      if (!any) break;
      uint32_t j[G], qt[G];
      bool h[G];
#pragma unroll
      for (int t = 0; t < G; ++t) {
        while (w[0] == 0u && q + 1u < nblk) {
#pragma unroll
          for (int k = 0; k + 1 < NB; ++k) w[k] = w[k + 1];
          w[NB - 1] = 0u;
          ++q;
        }
        h[t] = w[0] != 0u;
        // past the last hit, repeat it (the cursor only moves forward)
        const uint32_t pos = h[t] ? gb + 64u * q + (uint32_t)__builtin_ctzll(w[0]) : last;
        last = pos;
        w[0] &= w[0] - 1u;
        j[t] = c2.slot(pos);
        qt[t] = q;
      }
      decltype(load_f(0u)) d[G];
#pragma unroll
      for (int t = 0; t < G; ++t) d[t] = load_f(j[t]);
#pragma unroll
      for (int t = 0; t < G; ++t)
        if (h[t]) {
          while (done < qt[t]) {
            block_end();
            ++done;
          }
          add(j[t], d[t]);
        }
    }
    while (done < nblk) {
      block_end();
      ++done;
    }
  }
}

// Exclusive prefix sum of v over the active lanes of the wavefront; *total
// receives the sum.  All active lanes must call it.
__device__ __forceinline__ uint32_t wave_exclusive_sum(uint32_t v, uint32_t* total) {
  const int lane = threadIdx.x & 63;
  uint32_t incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(incl, off);
    if (lane >= off) incl += t;
  }
  *total = __shfl(incl, 63);
  return incl - v;
}

// Work lists of the heavy targets of one neighbour pass.
struct HeavyLists {
  // per pass parity: the record and chunk-item counters (rec_counter,
  // item_counter); counters[kThresholdWord]: the heavy threshold for the next
  // pass (0: kLightMax), set by each combine kernel from the kStatSlots
  // per-level target counts (stat_slot) the light kernel's wavefronts add
  // to.  Each word or slot has a 256-byte line of its own:
  // atomics on one line serialise, and every light wavefront reads the
  // threshold
  uint32_t* counters;
  uint4* recs;         // per heavy target: (i, first partial slot, chunks, 0)
  uint4* items;        // per chunk: (i, chunk index, partial slot, 0); i = ~0u: unused
  float* partials;     // per partial slot the pass's running sum, packed (at most
                       // kPartialWords floats)
  uint32_t item_cap;   // capacity of items / partials
  // hit masks of the current iteration: block b of target i at
  // masks[b * mask_stride + i], b < kMaskBlocks (null: not kept)
  uint64_t* masks;
  uint32_t mask_stride;
};
constexpr int kPartialWords = 8;
constexpr int kCounterLine = 64;  // words
constexpr int kThresholdWord = 4 * kCounterLine;
constexpr int kStatSlots = 64;
constexpr int kCounterWords = (5 + kStatSlots) * kCounterLine;
// Statistics slot k: kLevelWords words of packed per-level target counts.
__host__ __device__ __forceinline__ int stat_slot(int k) { return (5 + k) * kCounterLine; }
__host__ __device__ __forceinline__ int rec_counter(int parity) {
  return 2 * parity * kCounterLine;
}
__host__ __device__ __forceinline__ int item_counter(int parity) {
  return (2 * parity + 1) * kCounterLine;
}

// x^n for the s_corr exponent: repeated multiplication for small integer n
// (the reference default n=4), powf otherwise.
__device__ __forceinline__ float pow_corr(float w, const StepConsts& c) {
  if (c.n_corr_int == 4) {
    // the reference default; exactly the product the loop below forms for
    // n = 4 (1 * ((w*w) * (w*w)))
    const float w2 = w * w;
    return w2 * w2;
  }
  if (c.n_corr_int >= 0) {
    float r = 1.0f, b = w;
    int e = c.n_corr_int;
    while (e) {
      if (e & 1) r = r * b;
      b = b * b;
      e >>= 1;
    }
    return r;
  }
  return powf(w, c.n_corr);
}

// Arithmetic contract: the kernels are built with -ffp-contract=off, so every
// expression below is evaluated as written (IEEE FP32 +,-,*,/,sqrt and the
// explicit fmaf calls).  The host oracle's FP32 mirror executes the same
// sequence, which makes a correct GPU result equal to it bit for bit.

// Squared distance of two positions: dx*dx, then + dy*dy and + dz*dz as
// fused multiply-adds.
__device__ __forceinline__ float dist2(const float4& a, const float4& b, float& dx, float& dy,
                                       float& dz) {
  dx = a.x - b.x;
  dy = a.y - b.y;
  dz = a.z - b.z;
  float r2 = dx * dx;
  r2 = fmaf(dy, dy, r2);
  r2 = fmaf(dz, dz, r2);
  return r2;
}

// The same with the second position as a float3 (screening loads of x, y, z
// only).
__device__ __forceinline__ float dist2(const float4& a, const float3& b, float& dx, float& dy,
                                       float& dz) {
  dx = a.x - b.x;
  dy = a.y - b.y;
  dz = a.z - b.z;
  float r2 = dx * dx;
  r2 = fmaf(dy, dy, r2);
  r2 = fmaf(dz, dz, r2);
  return r2;
}

// x, y, z of element j of a float4 array (a 12-byte load).
__device__ __forceinline__ float3 load_xyz(const float4* __restrict__ a, uint32_t j) {
  const float* p = reinterpret_cast<const float*>(a + j);
  return make_float3(p[0], p[1], p[2]);
}

// Correctly rounded square root and quotient for the operand ranges of the
// pair terms, where no operand or result is zero, subnormal, infinite or NaN
// and no exponent is extreme: there the compiler's IEEE sequences
// (-fhip-fp32-correctly-rounded-divide-sqrt) apply no scaling and no
// special-case fix-up, and what remains is exactly the sequence below, so
// the results are the IEEE results bit for bit.  sqrt_pair: the hardware
// estimate, then the neighbour below or above it when its residual says so
// (inputs r2 in [KERNAL_EPS^2 * 0.99, h^2 * 1.0001); a smaller input, down
// to 0, still yields a value below KERNAL_EPS, which is all the callers
// test).  div_pair: reciprocal estimate, one Newton step, quotient with two
// residual corrections (numerators from 1e-18 to 1e6 over denominators from
// KERNAL_EPS to 1e4).
__device__ __forceinline__ float sqrt_pair(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float dn = __int_as_float(__float_as_int(s) - 1);
  const float up = __int_as_float(__float_as_int(s) + 1);
  float r = fmaf(-dn, s, x) <= 0.0f ? dn : s;
  r = fmaf(-up, s, x) > 0.0f ? up : r;
  return r;
}
__device__ __forceinline__ float div_pair(float a, float b) {
  float r = __builtin_amdgcn_rcpf(b);
  const float e = fmaf(-b, r, 1.0f);
  r = fmaf(e, r, r);
  float q = a * r;
  const float rem = fmaf(-b, q, a);
  q = fmaf(rem, r, q);
  const float rem2 = fmaf(-b, q, a);
  return fmaf(rem2, r, q);
}

// poly6 numerator times its coefficient, for r2 < h2: poly6 * (h2-r2)^3.
__device__ __forceinline__ float poly6_of(float r2, const StepConsts& c) {
  const float t = c.h2 - r2;
  return ((c.poly6 * t) * t) * t;
}

// DensityBoundary (fluids/Simulator.cu:139-161): per face, 0 beyond h,
// 2pi/3 at or behind the wall, (2pi/3)(h-d)^2(h+d) in between; faces summed
// in the reference's order.
__device__ __forceinline__ float wall_density_at(float d, const StepConsts& c) {
  if (d > c.h) return 0.0f;
  if (d <= 0.0f) return c.wall_c;
  return ((c.wall_c * (c.h - d)) * (c.h - d)) * (c.h + d);
}

__device__ __forceinline__ float boundary_density(float x, float y, float z, const SceneGeo& g,
                                                  const StepConsts& c) {
  float s = wall_density_at(g.ulx - x, c);
  s = s + wall_density_at(x - g.llx, c);
  s = s + wall_density_at(g.uly - y, c);
  s = s + wall_density_at(y - g.lly, c);
  s = s + wall_density_at(g.ulz - z, c);
  s = s + wall_density_at(z - g.llz, c);
  return s;
}

// Caller vector arrays: LAYOUT 4 = float4 (16 B), LAYOUT 3 = packed float3.
template <int LAYOUT>
__device__ __forceinline__ float4 load_vec(const float* __restrict__ a, uint32_t i) {
  if constexpr (LAYOUT == 4) {
    return reinterpret_cast<const float4*>(a)[i];
  } else {
    const float* p = a + (size_t)i * 3;
    return make_float4(p[0], p[1], p[2], 0.0f);
  }
}

template <int LAYOUT>
__device__ __forceinline__ void store_vec(float* __restrict__ a, uint32_t i, float x, float y,
                                          float z) {
  if constexpr (LAYOUT == 4) {
    reinterpret_cast<float4*>(a)[i] = make_float4(x, y, z, 0.0f);
  } else {
    float* p = a + (size_t)i * 3;
    p[0] = x;
    p[1] = y;
    p[2] = z;
  }
}

// Particles per cell, counted from the sorted keys (the input of the
// cell-start scan).  Equal keys are contiguous after the sort, so a run that
// lies inside one wavefront is that cell's only run and its first lane
// stores the count; a run touching a wavefront boundary (at most two per
// wavefront) may continue in the neighbouring wavefront, and each of its
// parts is added atomically.  The counts must be zero beforehand (the scan
// re-zeroes them).
__device__ __forceinline__ void count_sorted(const uint32_t* __restrict__ skey, int s, int n,
                                             uint32_t* __restrict__ cnt) {
  const int lane = threadIdx.x & 63;
  const bool valid = s < n;
  const uint32_t k = valid ? skey[s] : 0u;
  const uint32_t prev = __shfl_up(k, 1);
  const uint64_t vm = __ballot(valid);
  const bool head = valid && (lane == 0 || prev != k);
  const uint64_t hm = __ballot(head);
  if (head) {
    // run = [lane, next head or end of the valid lanes)
    const uint64_t later = lane == 63 ? 0ull : (hm >> (lane + 1)) << (lane + 1);
    const int end = later ? __ffsll((unsigned long long)later) - 1 : 64 - __clzll(vm);
    const int len = end - lane;
    const int after = s + len;  // sorted index just past the run
    const bool open_left = lane == 0 && s > 0 && skey[s - 1] == k;
    const bool open_right = !later && after < n && skey[after] == k;
    if (open_left || open_right)
      atomicAdd(&cnt[k], (uint32_t)len);
    else
      cnt[k] = (uint32_t)len;
  }
}

}  // namespace pbf
