// step.hip — the per-particle kernels of one PBF frame on gfx950.
//
// Data layout in HBM (per particle, sorted order unless noted):
//   xp   float4  predicted position, input order (dead after k_permute)
//   x0s  float4  start-of-frame position
//   xa   float4  current position iterate (w unused)
//   xl   float4  (position, lambda): written by the lambda pass, gathered by
//                the delta-p pass, so one 16-byte load brings a neighbour's
//                position and lambda (aliases xp)
//   vr   float4  (velocity before XSPH, rho): gathered by XSPH (aliases xp)
//   rho  float   density of the last lambda pass
//   cnt  uint32 per cell: particles per cell (kept all-zero between steps)
//   cs   uint32 per cell + 1: first sorted slot of each cell (exclusive scan
//        of cnt); a cell's particles are [cs[c], cs[c+1])
//
// The neighbour passes (lambda, delta-p, XSPH, density) are pass policies
// run by the generic kernels of pbf_pass.h.
#include "pbf_device.h"
#include "pbf_launch.h"
#include "pbf_pass.h"

namespace pbf {

namespace {

// ---------------------------------------------------------------------------
// Pass policies: what a target loads, what a candidate contributes, how the
// sums finish.  add() is the body of the reference's candidate loop.
// ---------------------------------------------------------------------------

// computeLambda (fluids/Simulator_kernel.cuh:52-129): rho sums poly6 over
// every candidate including i itself; the spiky gradient / pho0 is summed
// into grad_i and, for j != i, its squared norm into gradj_l2; then the wall
// term and lambda = -(rho/pho0 - 1) / (gradj_l2 + |grad_i|^2 + eps).
struct LambdaPass {
  static constexpr bool kXcdRanges = true;  // one contiguous range per XCD
  static constexpr int kMask = kMaskStore;  // hit masks for the delta-p pass
  // light workgroups stage their candidates' positions in LDS; add() uses j
  // only for the self test, which r >= KERNAL_EPS already excludes
#ifndef PBF_STAGE
#define PBF_STAGE 0
#endif
  static constexpr bool kStage = PBF_STAGE != 0;
  // hits reloaded per round of the sweep
  static constexpr int kHitBatch = 4;
  struct Args {
    const float4* __restrict__ x;
    float4* __restrict__ xl;
    float* __restrict__ rho;  // written only when non-null (last iteration, or a tape)
    float4* __restrict__ sd;  // tape only (else null): (sum of gradients, denominator)
  };
  struct Target {
    float4 x;
    uint32_t i;
  };
  struct Acc {
    float rho = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f, gl2 = 0.0f;
  };
  static __device__ __forceinline__ Target target(const Args& a, uint32_t i, const StepConsts&) {
    return Target{a.x[i], i};
  }
  static __device__ __forceinline__ const float4* screen(const Args& a) { return a.x; }
  static __device__ __forceinline__ float4 load_f(const Args& a, uint32_t j) { return a.x[j]; }
  static __device__ __forceinline__ bool hit(const Target& t, const float3& xj,
                                             const StepConsts& c) {
    float dx, dy, dz;
    return dist2(t.x, xj, dx, dy, dz) < c.h2_outer;
  }
  static __device__ __forceinline__ void add(Acc& s, const Target& t, uint32_t j, const float4& xj,
                                             const StepConsts& c, const Args&) {
    float dx, dy, dz;
    const float r2 = dist2(t.x, xj, dx, dy, dz);
    if (r2 < c.h2_outer) {
      if (r2 < c.h2) s.rho = s.rho + poly6_of(r2, c);
      // coincident pairs (piles clamped into a corner) have no gradient term:
      // skip their square root
      const float r = r2 >= c.keps2_lo ? sqrt_pair(r2) : 0.0f;
      if (r >= c.keps && r < c.h) {
        const float u = c.h - r;
        const float f = div_pair((c.spiky_rho0 * u) * u, r);
        const float ax = f * dx, ay = f * dy, az = f * dz;
        s.gx = s.gx + ax;
        s.gy = s.gy + ay;
        s.gz = s.gz + az;
        // j != i (Simulator_kernel.cuh:98): i itself has r = 0 < KERNAL_EPS
        // and never reaches this branch, so no index test is needed (j may
        // be a stage element index here)
        s.gl2 = fmaf(ax, ax, s.gl2);
        s.gl2 = fmaf(ay, ay, s.gl2);
        s.gl2 = fmaf(az, az, s.gl2);
      }
    }
  }
  static __device__ __forceinline__ void accumulate(Acc& tot, const Acc& p) {
    tot.rho = tot.rho + p.rho;
    tot.gx = tot.gx + p.gx;
    tot.gy = tot.gy + p.gy;
    tot.gz = tot.gz + p.gz;
    tot.gl2 = tot.gl2 + p.gl2;
  }
  static __device__ __forceinline__ void finish(const Args& a, const Target& t, const Acc& s,
                                                const StepConsts& c, const SceneGeo& g) {
    float rho = s.rho;
    if (c.k_boundary != 0.0f)
      rho = rho + c.k_boundary * boundary_density(t.x.x, t.x.y, t.x.z, g, c);
    float den = fmaf(s.gx, s.gx, s.gl2);
    den = fmaf(s.gy, s.gy, den);
    den = fmaf(s.gz, s.gz, den);
    den = den + c.lambda_eps;
    const float lam = -(rho / c.pho0 - 1.0f) / den;
    a.xl[t.i] = make_float4(t.x.x, t.x.y, t.x.z, lam);
    if (a.rho) a.rho[t.i] = rho;
    if (a.sd) a.sd[t.i] = make_float4(s.gx, s.gy, s.gz, den);
  }
};

// computetpos (fluids/Simulator_kernel.cuh:131-194): sum over j != i of
// (lambda_i + lambda_j + s_corr) * spikyGrad, / pho0, clamped to +-MAX_DP per
// axis, added, then clamped into the box shrunk by LIM_EPS.  Reads
// (x, lambda) from xl, writes the new iterate to xa (Jacobi; the ping-pong
// replaces thrust::copy_n at fluids/Simulator.cu:247-248).
struct DeltaPass {
  static constexpr bool kXcdRanges = true;  // one contiguous range per XCD
  static constexpr int kMask = kMaskLoad;   // the lambda pass's hit masks
  // hits reloaded per round of the sweep
  static constexpr int kHitBatch = 4;
  struct Args {
    const float4* __restrict__ xl;
    float4* __restrict__ xa;
    float4* __restrict__ dp;  // tape only (else null): the correction / pho0 before clamps
  };
  struct Target {
    float4 x;  // w = lambda_i
    uint32_t i;
  };
  struct Acc {
    float px = 0.0f, py = 0.0f, pz = 0.0f;
  };
  static __device__ __forceinline__ Target target(const Args& a, uint32_t i, const StepConsts&) {
    return Target{a.xl[i], i};
  }
  static __device__ __forceinline__ const float4* screen(const Args& a) { return a.xl; }
  static __device__ __forceinline__ float4 load_f(const Args& a, uint32_t j) { return a.xl[j]; }
  static __device__ __forceinline__ bool hit(const Target& t, const float3& xj,
                                             const StepConsts& c) {
    float dx, dy, dz;
    // pairs closer than KERNAL_EPS (coincident piles, and i itself at r2 = 0)
    // add nothing
    const float r2 = dist2(t.x, xj, dx, dy, dz);
    return (r2 < c.h2_outer) & (r2 >= c.keps2_lo);
  }
  static __device__ __forceinline__ void add(Acc& s, const Target& t, uint32_t j, const float4& xj,
                                             const StepConsts& c, const Args&) {
    if (j == t.i) return;
    float dx, dy, dz;
    const float r2 = dist2(t.x, xj, dx, dy, dz);
    if (r2 < c.h2_outer) {
      const float r = sqrt_pair(r2);
      if (r >= c.keps && r < c.h) {
        const float w = r2 < c.h2 ? poly6_of(r2, c) : 0.0f;
        const float corr = c.coef_corr * pow_corr(w, c);
        const float u = c.h - r;
        const float gs = div_pair((c.spiky * u) * u, r);
        const float f = ((t.x.w + xj.w) + corr) * gs;
        s.px = fmaf(f, dx, s.px);
        s.py = fmaf(f, dy, s.py);
        s.pz = fmaf(f, dz, s.pz);
      }
    }
  }
  static __device__ __forceinline__ void accumulate(Acc& tot, const Acc& p) {
    tot.px = tot.px + p.px;
    tot.py = tot.py + p.py;
    tot.pz = tot.pz + p.pz;
  }
  // The new iterate (stored into xa); also records dp on a tape.
  static __device__ __forceinline__ float3 update(const Args& a, const Target& t, const Acc& s,
                                                  const StepConsts& c, const SceneGeo& g) {
    const float ux = s.px / c.pho0, uy = s.py / c.pho0, uz = s.pz / c.pho0;
    if (a.dp) a.dp[t.i] = make_float4(ux, uy, uz, 0.0f);
    const float px = fminf(fmaxf(ux, -c.max_dp), c.max_dp);
    const float py = fminf(fmaxf(uy, -c.max_dp), c.max_dp);
    const float pz = fminf(fmaxf(uz, -c.max_dp), c.max_dp);
    const float x = fmaxf(fminf(t.x.x + px, g.hix), g.lox);
    const float y = fmaxf(fminf(t.x.y + py, g.hiy), g.loy);
    const float z = fmaxf(fminf(t.x.z + pz, g.hiz), g.loz);
    a.xa[t.i] = make_float4(x, y, z, 0.0f);
    return make_float3(x, y, z);
  }
  static __device__ __forceinline__ void finish(const Args& a, const Target& t, const Acc& s,
                                                const StepConsts& c, const SceneGeo& g) {
    update(a, t, s, c, g);
  }
};

// The last delta-p pass of a step, fused with updateVelocity
// (fluids/Simulator.cu:127-137, :267-274): v = (x_final - x_start) * (1/dt),
// packed with the last lambda pass's rho for XSPH, and the caller's pos / vel
// / iid receive their end-of-step contents (sorted start positions, pre-XSPH
// velocities, sorted identities).
template <int L>
struct DeltaLastPass : DeltaPass {
  struct Args : DeltaPass::Args {
    const float4* __restrict__ x0s;
    const float* __restrict__ rho;
    const uint32_t* __restrict__ iid_s;
    float4* __restrict__ vr;
    float* __restrict__ pos;
    float* __restrict__ vel;
    uint32_t* __restrict__ iid;
  };
  static __device__ __forceinline__ void finish(const Args& a, const Target& t, const Acc& s,
                                                const StepConsts& c, const SceneGeo& g) {
    const float3 x = update(a, t, s, c, g);
    const float4 x0 = a.x0s[t.i];
    const float vx = (x.x - x0.x) * c.inv_dt;
    const float vy = (x.y - x0.y) * c.inv_dt;
    const float vz = (x.z - x0.z) * c.inv_dt;
    a.vr[t.i] = make_float4(vx, vy, vz, a.rho[t.i]);
    store_vec<L>(a.vel, t.i, vx, vy, vz);
    store_vec<L>(a.pos, t.i, x0.x, x0.y, x0.z);
    a.iid[t.i] = a.iid_s[t.i];
  }
};

// computeXSPH (fluids/Simulator_kernel.cuh:196-239): nvel_i = v_i +
// c * sum_j 2 (v_j - v_i) W(|x_i - x_j|^2) / (rho_i + rho_j), j == i
// included (its term is 0), rho of the last lambda pass.  Pairs with W == 0
// are skipped: their reference term is exactly 0.  Also stores the final
// positions into the caller's npos.
template <int L>
struct XsphPass {
  static constexpr bool kXcdRanges = true;  // one contiguous range per XCD
  // hits reloaded per round of the sweep
  static constexpr int kHitBatch = 4;
  struct Args {
    const float4* __restrict__ xa;
    const float4* __restrict__ vr;  // (v, rho)
    float* __restrict__ npos;
    float* __restrict__ nvel;
  };
  struct Target {
    float4 x, v;
    uint32_t i;
  };
  struct Acc {
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
  };
  // Screening loads positions only; a neighbour's (v, rho) is loaded with
  // its position for the pairs inside h.
  struct Nb {
    float4 x, v;
  };
  static __device__ __forceinline__ Target target(const Args& a, uint32_t i, const StepConsts&) {
    return Target{a.xa[i], a.vr[i], i};
  }
  static __device__ __forceinline__ const float4* screen(const Args& a) { return a.xa; }
  static __device__ __forceinline__ Nb load_f(const Args& a, uint32_t j) {
    return Nb{a.xa[j], a.vr[j]};
  }
  static __device__ __forceinline__ bool hit(const Target& t, const float3& xj,
                                             const StepConsts& c) {
    float dx, dy, dz;
    return dist2(t.x, xj, dx, dy, dz) < c.h2;
  }
  static __device__ __forceinline__ void add(Acc& s, const Target& t, uint32_t, const Nb& nb,
                                             const StepConsts& c, const Args&) {
    float dx, dy, dz;
    const float r2 = dist2(t.x, nb.x, dx, dy, dz);
    if (r2 < c.h2) {
      const float4 vj = nb.v;
      const float w = poly6_of(r2, c);
      const float f = div_pair(2.0f * w, t.v.w + vj.w);
      s.ax = fmaf(vj.x - t.v.x, f, s.ax);
      s.ay = fmaf(vj.y - t.v.y, f, s.ay);
      s.az = fmaf(vj.z - t.v.z, f, s.az);
    }
  }
  static __device__ __forceinline__ void accumulate(Acc& tot, const Acc& p) {
    tot.ax = tot.ax + p.ax;
    tot.ay = tot.ay + p.ay;
    tot.az = tot.az + p.az;
  }
  static __device__ __forceinline__ void finish(const Args& a, const Target& t, const Acc& s,
                                                const StepConsts& c, const SceneGeo&) {
    store_vec<L>(a.nvel, t.i, fmaf(c.c_xsph, s.ax, t.v.x), fmaf(c.c_xsph, s.ay, t.v.y),
                 fmaf(c.c_xsph, s.az, t.v.z));
    store_vec<L>(a.npos, t.i, t.x.x, t.x.y, t.x.z);
  }
};

// rho on a fresh grid (the lambda pass's density term), for the density
// error statistic: |rho/pho0 - 1| per particle in FP64.
struct DensityPass {
  static constexpr bool kXcdRanges = true;  // one contiguous range per XCD
  // hits reloaded per round of the sweep
  static constexpr int kHitBatch = 4;
  struct Args {
    const float4* __restrict__ x;
    double* __restrict__ err;
  };
  struct Target {
    float4 x;
    uint32_t i;
  };
  struct Acc {
    float rho = 0.0f;
  };
  static __device__ __forceinline__ Target target(const Args& a, uint32_t i, const StepConsts&) {
    return Target{a.x[i], i};
  }
  static __device__ __forceinline__ const float4* screen(const Args& a) { return a.x; }
  static __device__ __forceinline__ float4 load_f(const Args& a, uint32_t j) { return a.x[j]; }
  static __device__ __forceinline__ bool hit(const Target& t, const float3& xj,
                                             const StepConsts& c) {
    float dx, dy, dz;
    return dist2(t.x, xj, dx, dy, dz) < c.h2;
  }
  static __device__ __forceinline__ void add(Acc& s, const Target& t, uint32_t, const float4& xj,
                                             const StepConsts& c, const Args&) {
    float dx, dy, dz;
    const float r2 = dist2(t.x, xj, dx, dy, dz);
    if (r2 < c.h2) s.rho = s.rho + poly6_of(r2, c);
  }
  static __device__ __forceinline__ void accumulate(Acc& tot, const Acc& p) {
    tot.rho = tot.rho + p.rho;
  }
  static __device__ __forceinline__ void finish(const Args& a, const Target& t, const Acc& s,
                                                const StepConsts& c, const SceneGeo& g) {
    float rho = s.rho;
    if (c.k_boundary != 0.0f)
      rho = rho + c.k_boundary * boundary_density(t.x.x, t.x.y, t.x.z, g, c);
    a.err[t.i] = fabs((double)rho / (double)c.pho0 - 1.0);
  }
};

// ---------------------------------------------------------------------------
// Streaming kernels.
// ---------------------------------------------------------------------------

// advect_kernel (fluids/Simulator_kernel.cuh:7-19) fused with getGridId
// (fluids/Simulator.cu:56-73).  vel += dt*g and npos = pos + dt*vel as fused
// multiply-adds; the updated velocity itself is not stored, since the
// reference overwrites it before any read (fluids/Simulator.cu:267-274).
// Also clears the heavy counters of the first neighbour pass.
template <int L, class S>
__global__ __launch_bounds__(kBlock) void k_predict_key(const float* __restrict__ pos,
                                                        const float* __restrict__ vel, int n,
                                                        StepConsts c, S sc,
                                                        float4* __restrict__ xp,
                                                        uint32_t* __restrict__ key,
                                                        uint32_t* __restrict__ val,
                                                        uint32_t* __restrict__ counters) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i == 0) {
    counters[rec_counter(0)] = 0u;
    counters[item_counter(0)] = 0u;
  }
  if (blockIdx.x * kBlock >= n) return;  // whole workgroup past the end
  const int ii = i < n ? i : n - 1;      // idle lanes repeat the last particle
  const float4 p = load_vec<L>(pos, ii);
  const float4 v = load_vec<L>(vel, ii);
  const float vx = fmaf(c.dt, 0.0f, v.x);
  const float vy = fmaf(c.dt, 0.0f, v.y);
  const float vz = fmaf(c.dt, c.gz, v.z);
  const float x = fmaf(c.dt, vx, p.x);
  const float y = fmaf(c.dt, vy, p.y);
  const float z = fmaf(c.dt, vz, p.z);
  const SceneGeo g = sc.at((uint32_t)ii, c);
  const uint32_t k = cell_id(cell_of(x, y, z, g, c.h), g);
  if (i >= n) return;
  xp[i] = make_float4(x, y, z, 0.0f);
  key[i] = k;
  val[i] = (uint32_t)i;
}

// Cell key of positions as they are (the fresh grid of the density error).
template <int L, class S>
__global__ __launch_bounds__(kBlock) void k_key_only(const float* __restrict__ pos, int n,
                                                     StepConsts c, S sc, float4* __restrict__ xp,
                                                     uint32_t* __restrict__ key,
                                                     uint32_t* __restrict__ val,
                                                     uint32_t* __restrict__ counters) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i == 0) {
    counters[rec_counter(0)] = 0u;
    counters[item_counter(0)] = 0u;
  }
  if (blockIdx.x * kBlock >= n) return;
  const int ii = i < n ? i : n - 1;
  const float4 p = load_vec<L>(pos, ii);
  const SceneGeo g = sc.at((uint32_t)ii, c);
  const uint32_t k = cell_id(cell_of(p.x, p.y, p.z, g, c.h), g);
  if (i >= n) return;
  xp[i] = make_float4(p.x, p.y, p.z, 0.0f);
  key[i] = k;
  val[i] = (uint32_t)i;
}

// Fill of the per-cell counts (solver creation, or after a step that failed
// between counting and the scan).  A fill kernel, not a memset node.
__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* __restrict__ a, long long n,
                                                     uint32_t v) {
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long k = (long long)blockIdx.x * kBlock + threadIdx.x; k < n; k += stride) a[k] = v;
}

// ---- cell-start table: exclusive scan of the per-cell counts -------------
// Replaces cudaMemset + computeGridRange (fluids/Simulator.cu:201-204,
// fluids/Simulator_kernel.cuh:21-50): cs[c] = number of particles in cells
// < c, which is the first sorted slot of cell c (empty cells: the next
// non-empty cell's start).  Kernels: tile sums, scan of the tile sums (large
// grids only), tile-local scan + offset (which also clears the counts for the
// next step).
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;
// Grids of up to this many tiles (4.2M cells) scan in two kernels.
constexpr int kFusedScanTiles = 2048;

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* sh) {
  // wave sums by shuffles, then the four wave sums
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  const uint32_t t = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return t;
}

// Exclusive scan over the 256 threads; *total gets the block sum.
__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* sh, uint32_t* total) {
  uint32_t tot;
  const uint32_t ex = wave_exclusive_sum(v, &tot);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sh[w] = tot;
  __syncthreads();
  uint32_t base = 0, all = 0;
#pragma unroll
  for (int k = 0; k < kBlock / 64; ++k) {
    base += k < w ? sh[k] : 0u;
    all += sh[k];
  }
  __syncthreads();
  *total = all;
  return base + ex;
}

__global__ __launch_bounds__(kBlock) void k_scan_tiles(const uint32_t* __restrict__ cnt,
                                                       long long cells,
                                                       uint32_t* __restrict__ tile_sum) {
  __shared__ uint32_t sh[kBlock / 64];
  const long long base = (long long)blockIdx.x * kScanTile;
  uint32_t v = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const long long c = base + (long long)k * kBlock + threadIdx.x;
    if (c < cells) v += cnt[c];
  }
  const uint32_t t = block_sum(v, sh);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = t;
}

// One workgroup: exclusive scan of the tile sums in place; cs[cells] = total.
__global__ __launch_bounds__(kBlock) void k_scan_tile_sums(uint32_t* __restrict__ tile_sum,
                                                           int ntiles, long long cells,
                                                           uint32_t* __restrict__ cs) {
  __shared__ uint32_t sh[kBlock / 64];
  uint32_t carry = 0;
  for (int t0 = 0; t0 < ntiles; t0 += kBlock) {
    const int t = t0 + (int)threadIdx.x;
    const uint32_t v = t < ntiles ? tile_sum[t] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive(v, sh, &total);
    if (t < ntiles) tile_sum[t] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) cs[cells] = carry;
}

// Tile-local exclusive scan plus the tile's offset; also clears the counts
// for the next step.  tile_off holds the tile offsets (kFused = false), or the
// tile sums themselves (kFused = true: each workgroup adds up the sums of the
// tiles before it, and the last one writes cs[cells]).  Thread t owns cells
// base + 8t .. base + 8t + 7, read and written as two 16-byte vectors.
template <bool kFused>
__global__ __launch_bounds__(kBlock) void k_scan_cells(uint32_t* __restrict__ cnt, long long cells,
                                                       const uint32_t* __restrict__ tile_off,
                                                       uint32_t* __restrict__ cs) {
  static_assert(kScanItems == 8, "two uint4 per thread");
  __shared__ uint32_t sh[kBlock / 64];
  uint32_t off;
  if (kFused) {
    uint32_t v = 0;
    for (int t = threadIdx.x; t < (int)blockIdx.x; t += kBlock) v += tile_off[t];
    off = block_sum(v, sh);
  } else {
    off = tile_off[blockIdx.x];
  }
  const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
  const bool whole = base + kScanItems <= cells;
  uint32_t v[kScanItems];
  if (whole) {
    const uint4 a0 = *reinterpret_cast<const uint4*>(cnt + base);
    const uint4 a1 = *reinterpret_cast<const uint4*>(cnt + base + 4);
    v[0] = a0.x, v[1] = a0.y, v[2] = a0.z, v[3] = a0.w;
    v[4] = a1.x, v[5] = a1.y, v[6] = a1.z, v[7] = a1.w;
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) v[k] = base + k < cells ? cnt[base + k] : 0u;
  }
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) sum += v[k];
  uint32_t total;
  const uint32_t first = off + block_exclusive(sum, sh, &total);
  uint32_t o[kScanItems];
  uint32_t run = first;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    o[k] = run;
    run += v[k];
  }
  if (whole) {
    *reinterpret_cast<uint4*>(cs + base) = make_uint4(o[0], o[1], o[2], o[3]);
    *reinterpret_cast<uint4*>(cs + base + 4) = make_uint4(o[4], o[5], o[6], o[7]);
    *reinterpret_cast<uint4*>(cnt + base) = make_uint4(0u, 0u, 0u, 0u);
    *reinterpret_cast<uint4*>(cnt + base + 4) = make_uint4(0u, 0u, 0u, 0u);
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
      if (base + k < cells) {
        cs[base + k] = o[k];
        cnt[base + k] = 0u;
      }
  }
  if (kFused && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) cs[cells] = off + total;
}

// Gather of the sorted payload (the zip permutation of
// fluids/Simulator.cu:196-198), with the per-cell counts of the sorted keys
// (computeGridRange's input, fluids/Simulator.cu:201-204).  Indices < n
// hold for any sorted index array; the guard keeps a corrupted one from
// reaching out of bounds.
template <int L>
__global__ __launch_bounds__(kBlock) void k_permute(
    const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sval,
    const float* __restrict__ pos, const float4* __restrict__ xp,
    const uint32_t* __restrict__ iid, int n, float4* __restrict__ x0s,
    float4* __restrict__ xa, uint32_t* __restrict__ iid_s, uint32_t* __restrict__ cnt) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (blockIdx.x * kBlock >= n) return;  // whole workgroup past the end
  count_sorted(skey, s, n, cnt);
  if (s >= n) return;
  const uint32_t src = sval[s] < (uint32_t)n ? sval[s] : 0u;
  x0s[s] = load_vec<L>(pos, src);
  xa[s] = xp[src];
  iid_s[s] = iid[src];
}

__global__ __launch_bounds__(kBlock) void k_permute_pos(const uint32_t* __restrict__ skey,
                                                        const uint32_t* __restrict__ sval,
                                                        const float4* __restrict__ xp, int n,
                                                        float4* __restrict__ xa,
                                                        uint32_t* __restrict__ cnt) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (blockIdx.x * kBlock >= n) return;
  count_sorted(skey, s, n, cnt);
  if (s >= n) return;
  xa[s] = xp[sval[s] < (uint32_t)n ? sval[s] : 0u];
}

constexpr int kRedBlock = 256;
constexpr int kRedItems = 16;  // particles per work-item in the first reduction

// Deterministic two-level reduction of the per-particle errors: blocks of
// kRedBlock * kRedItems particles to (sum, max) partials, then one block.
__global__ __launch_bounds__(kRedBlock) void k_reduce_err(const double* __restrict__ err, int n,
                                                          double* __restrict__ partial) {
  __shared__ double ssum[kRedBlock];
  __shared__ double smax[kRedBlock];
  const int base = blockIdx.x * kRedBlock * kRedItems;
  double s = 0.0, m = 0.0;
  for (int k = 0; k < kRedItems; ++k) {
    const int i = base + k * kRedBlock + (int)threadIdx.x;
    if (i < n) {
      s += err[i];
      m = fmax(m, err[i]);
    }
  }
  ssum[threadIdx.x] = s;
  smax[threadIdx.x] = m;
  __syncthreads();
  for (int off = kRedBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      ssum[threadIdx.x] += ssum[threadIdx.x + off];
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = ssum[0];
    partial[2 * blockIdx.x + 1] = smax[0];
  }
}

__global__ __launch_bounds__(kRedBlock) void k_reduce_partials(const double* __restrict__ partial,
                                                               int nb, int n,
                                                               double* __restrict__ out) {
  __shared__ double ssum[kRedBlock];
  __shared__ double smax[kRedBlock];
  double s = 0.0, m = 0.0;
  for (int b = threadIdx.x; b < nb; b += kRedBlock) {
    s += partial[2 * b];
    m = fmax(m, partial[2 * b + 1]);
  }
  ssum[threadIdx.x] = s;
  smax[threadIdx.x] = m;
  __syncthreads();
  for (int off = kRedBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      ssum[threadIdx.x] += ssum[threadIdx.x + off];
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = ssum[0] / (double)n;
    out[1] = smax[0];
  }
}

}  // namespace

template <int L, class S>
void predict_key_s(const float* pos, const float* vel, int n, const StepConsts& c, const S& sc,
                   float4* xp, uint32_t* key, uint32_t* val, uint32_t* counters,
                   hipStream_t st) {
  k_predict_key<L, S><<<blocks_for(n), kBlock, 0, st>>>(pos, vel, n, c, sc, xp, key, val,
                                                        counters);
}

template <int L, class S>
void key_only_s(const float* pos, int n, const StepConsts& c, const S& sc, float4* xp,
                uint32_t* key, uint32_t* val, uint32_t* counters, hipStream_t st) {
  k_key_only<L, S><<<blocks_for(n), kBlock, 0, st>>>(pos, n, c, sc, xp, key, val, counters);
}

void launch_predict_key(int layout, const float* pos, const float* vel, int n, const StepConsts& c,
                        const Scenes& scn, float4* xp, uint32_t* key, uint32_t* val,
                        uint32_t* counters, hipStream_t st) {
  const SceneTable tab{scn.geo, scn.sid};
  if (layout == 3) {
    if (scn.geo)
      predict_key_s<3>(pos, vel, n, c, tab, xp, key, val, counters, st);
    else
      predict_key_s<3>(pos, vel, n, c, OneScene{}, xp, key, val, counters, st);
  } else {
    if (scn.geo)
      predict_key_s<4>(pos, vel, n, c, tab, xp, key, val, counters, st);
    else
      predict_key_s<4>(pos, vel, n, c, OneScene{}, xp, key, val, counters, st);
  }
}

void launch_key_only(int layout, const float* pos, int n, const StepConsts& c, const Scenes& scn,
                     float4* xp, uint32_t* key, uint32_t* val, uint32_t* counters,
                     hipStream_t st) {
  const SceneTable tab{scn.geo, scn.sid};
  if (layout == 3) {
    if (scn.geo)
      key_only_s<3>(pos, n, c, tab, xp, key, val, counters, st);
    else
      key_only_s<3>(pos, n, c, OneScene{}, xp, key, val, counters, st);
  } else {
    if (scn.geo)
      key_only_s<4>(pos, n, c, tab, xp, key, val, counters, st);
    else
      key_only_s<4>(pos, n, c, OneScene{}, xp, key, val, counters, st);
  }
}

void launch_fill_u32(uint32_t* a, long long n, uint32_t v, hipStream_t st) {
  long long blocks = (n + kBlock - 1) / kBlock;
  if (blocks < 1) blocks = 1;
  if (blocks > 8192) blocks = 8192;
  k_fill_u32<<<(int)blocks, kBlock, 0, st>>>(a, n, v);
}

long long scan_tiles(long long cells) { return (cells + kScanTile - 1) / kScanTile; }

void launch_cell_starts(uint32_t* cnt, long long cells, uint32_t* tile_sum, uint32_t* cs,
                        hipStream_t st) {
  const int ntiles = (int)scan_tiles(cells);
  k_scan_tiles<<<ntiles, kBlock, 0, st>>>(cnt, cells, tile_sum);
  if (ntiles <= kFusedScanTiles) {
    // each workgroup adds up the tile sums before it (<= 2048 words)
    k_scan_cells<true><<<ntiles, kBlock, 0, st>>>(cnt, cells, tile_sum, cs);
    return;
  }
  k_scan_tile_sums<<<1, kBlock, 0, st>>>(tile_sum, ntiles, cells, cs);
  k_scan_cells<false><<<ntiles, kBlock, 0, st>>>(cnt, cells, tile_sum, cs);
}

void launch_permute(int layout, const uint32_t* skey, const uint32_t* sval, const float* pos,
                    const float4* xp, const uint32_t* iid, int n, float4* x0s, float4* xa,
                    uint32_t* iid_s, uint32_t* cnt, hipStream_t st) {
  if (layout == 3)
    k_permute<3><<<blocks_for(n), kBlock, 0, st>>>(skey, sval, pos, xp, iid, n, x0s, xa, iid_s,
                                                   cnt);
  else
    k_permute<4><<<blocks_for(n), kBlock, 0, st>>>(skey, sval, pos, xp, iid, n, x0s, xa, iid_s,
                                                   cnt);
}

void launch_permute_pos(const uint32_t* skey, const uint32_t* sval, const float4* xp, int n,
                        float4* xa, uint32_t* cnt, hipStream_t st) {
  k_permute_pos<<<blocks_for(n), kBlock, 0, st>>>(skey, sval, xp, n, xa, cnt);
}

void launch_lambda(const float4* xa, const uint32_t* cs, int n, const StepConsts& c,
                   const Scenes& scn, float4* xl, float* rho, float4* sd, const HeavyLists& hl,
                   int pass, const Shapes& sh, hipStream_t st) {
  launch_pass<LambdaPass>(LambdaPass::Args{xa, xl, rho, sd}, n, c, scn, cs, hl, pass, sh, st);
}

void launch_delta(const float4* xl, const uint32_t* cs, int n, const StepConsts& c,
                  const Scenes& scn, float4* xa, float4* dp, const VelocityOut* vo,
                  const HeavyLists& hl, int pass, const Shapes& sh, hipStream_t st) {
  const DeltaPass::Args base{xl, xa, dp};
  if (!vo) {
    launch_pass<DeltaPass>(base, n, c, scn, cs, hl, pass, sh, st);
  } else if (vo->layout == 3) {
    const DeltaLastPass<3>::Args a{base, vo->x0s, vo->rho, vo->iid_s, vo->vr, vo->pos, vo->vel,
                                   vo->iid};
    launch_pass<DeltaLastPass<3>>(a, n, c, scn, cs, hl, pass, sh, st);
  } else {
    const DeltaLastPass<4>::Args a{base, vo->x0s, vo->rho, vo->iid_s, vo->vr, vo->pos, vo->vel,
                                   vo->iid};
    launch_pass<DeltaLastPass<4>>(a, n, c, scn, cs, hl, pass, sh, st);
  }
}

void launch_xsph(int layout, const float4* xa, const float4* vr, const uint32_t* cs, int n,
                 const StepConsts& c, const Scenes& scn, float* npos, float* nvel,
                 const HeavyLists& hl, int pass, const Shapes& sh, hipStream_t st) {
  if (layout == 3)
    launch_pass<XsphPass<3>>(XsphPass<3>::Args{xa, vr, npos, nvel}, n, c, scn, cs, hl, pass, sh,
                             st);
  else
    launch_pass<XsphPass<4>>(XsphPass<4>::Args{xa, vr, npos, nvel}, n, c, scn, cs, hl, pass, sh,
                             st);
}

int density_partials(int n) { return (n + kRedBlock * kRedItems - 1) / (kRedBlock * kRedItems); }

void launch_density_error(const float4* xa, const uint32_t* cs, int n, const StepConsts& c,
                          const Scenes& scn, double* err, double* partial, double* out,
                          const HeavyLists& hl, const Shapes& sh, hipStream_t st) {
  launch_pass<DensityPass>(DensityPass::Args{xa, err}, n, c, scn, cs, hl, 0, sh, st);
  const int nb = density_partials(n);
  k_reduce_err<<<nb, kRedBlock, 0, st>>>(err, n, partial);
  k_reduce_partials<<<1, kRedBlock, 0, st>>>(partial, nb, n, out);
}

}  // namespace pbf
