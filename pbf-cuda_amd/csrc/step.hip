// step.hip — the per-particle kernels of one PBF frame on gfx950.
//
// Data layout in HBM (per particle, sorted order unless noted):
//   xp   float4  predicted position, input order (dead after k_permute)
//   x0s  float4  start-of-frame position
//   xa   float4  current position iterate (w unused)
//   xl   float4  (position, lambda): written by k_lambda, gathered by
//                k_delta, so one 16-byte load brings a neighbour's position
//                and lambda (aliases xp)
//   vr   float4  (velocity before XSPH, rho): gathered by k_xsph (aliases xp)
//   rho  float   density of the last lambda pass
//   ranges uint2 per cell: [start, end) of its particles, 0/0 when empty
// The neighbour kernels are one work-item per particle over the cell-sorted
// order; consecutive work-items share neighbour cells, so their gathers hit
// the same cache lines, and workgroups are remapped so each XCD's L2 serves
// one contiguous range of particles.
#include "pbf_device.h"
#include "pbf_launch.h"

namespace pbf {

namespace {

constexpr int kBlock = 256;

inline int blocks_for(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// advect_kernel (fluids/Simulator_kernel.cuh:7-19) fused with getGridId
// (fluids/Simulator.cu:56-73).  vel += dt*g and npos = pos + dt*vel as fused
// multiply-adds; the updated velocity itself is not stored, since the
// reference overwrites it before any read (fluids/Simulator.cu:267-274).
template <int L>
__global__ __launch_bounds__(kBlock) void k_predict_key(const float* __restrict__ pos,
                                                        const float* __restrict__ vel, int n,
                                                        StepConsts c, float4* __restrict__ xp,
                                                        uint32_t* __restrict__ key,
                                                        uint32_t* __restrict__ val) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 p = load_vec<L>(pos, i);
  const float4 v = load_vec<L>(vel, i);
  const float vx = fmaf(c.dt, 0.0f, v.x);
  const float vy = fmaf(c.dt, 0.0f, v.y);
  const float vz = fmaf(c.dt, c.gz, v.z);
  const float x = fmaf(c.dt, vx, p.x);
  const float y = fmaf(c.dt, vy, p.y);
  const float z = fmaf(c.dt, vz, p.z);
  xp[i] = make_float4(x, y, z, 0.0f);
  key[i] = cell_id(cell_of(x, y, z, c), c);
  val[i] = (uint32_t)i;
}

// Cell key of positions as they are (the fresh grid of the density error).
template <int L>
__global__ __launch_bounds__(kBlock) void k_key_only(const float* __restrict__ pos, int n,
                                                     StepConsts c, float4* __restrict__ xp,
                                                     uint32_t* __restrict__ key,
                                                     uint32_t* __restrict__ val) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 p = load_vec<L>(pos, i);
  xp[i] = make_float4(p.x, p.y, p.z, 0.0f);
  key[i] = cell_id(cell_of(p.x, p.y, p.z, c), c);
  val[i] = (uint32_t)i;
}

// cudaMemset of gridStart/gridEnd (fluids/Simulator.cu:201-203) as a fill
// kernel (a memset node must not sit in a captured graph), two cells per
// 16-byte store.
__global__ __launch_bounds__(kBlock) void k_clear_ranges(uint2* __restrict__ ranges,
                                                         long long cells) {
  const long long pairs = cells >> 1;
  uint4* r4 = reinterpret_cast<uint4*>(ranges);
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long k = (long long)blockIdx.x * kBlock + threadIdx.x; k < pairs; k += stride)
    r4[k] = make_uint4(0u, 0u, 0u, 0u);
  if ((cells & 1) && blockIdx.x == 0 && threadIdx.x == 0) ranges[cells - 1] = make_uint2(0u, 0u);
}

// Gather of the sorted payload (the zip permutation of
// fluids/Simulator.cu:196-198) fused with computeGridRange
// (fluids/Simulator_kernel.cuh:21-50): a slot whose key differs from the
// previous slot's opens its cell and closes the previous one.
template <int L>
__global__ __launch_bounds__(kBlock) void k_permute(
    const uint32_t* __restrict__ skey, const uint32_t* __restrict__ sval,
    const float* __restrict__ pos, const float4* __restrict__ xp, const uint32_t* __restrict__ iid,
    int n, int cells, float4* __restrict__ x0s, float4* __restrict__ xa,
    uint32_t* __restrict__ iid_s, uint2* __restrict__ ranges) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  const uint32_t k = skey[s];
  const uint32_t prev = s > 0 ? skey[s - 1] : 0xffffffffu;
  // keys < cells and indices < n hold for any sorted key/index array; the
  // guards keep a corrupted one from reaching out of bounds
  if (k < (uint32_t)cells) {
    if (s == 0 || prev != k) {
      ranges[k].x = (uint32_t)s;
      if (s > 0 && prev < (uint32_t)cells) ranges[prev].y = (uint32_t)s;
    }
    if (s == n - 1) ranges[k].y = (uint32_t)n;
  }
  const uint32_t src = sval[s] < (uint32_t)n ? sval[s] : 0u;
  x0s[s] = load_vec<L>(pos, src);
  xa[s] = xp[src];
  iid_s[s] = iid[src];
}

__global__ __launch_bounds__(kBlock) void k_permute_pos(const uint32_t* __restrict__ skey,
                                                        const uint32_t* __restrict__ sval,
                                                        const float4* __restrict__ xp, int n,
                                                        int cells, float4* __restrict__ xa,
                                                        uint2* __restrict__ ranges) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  const uint32_t k = skey[s];
  const uint32_t prev = s > 0 ? skey[s - 1] : 0xffffffffu;
  if (k < (uint32_t)cells) {
    if (s == 0 || prev != k) {
      ranges[k].x = (uint32_t)s;
      if (s > 0 && prev < (uint32_t)cells) ranges[prev].y = (uint32_t)s;
    }
    if (s == n - 1) ranges[k].y = (uint32_t)n;
  }
  xa[s] = xp[sval[s] < (uint32_t)n ? sval[s] : 0u];
}

// computeLambda (fluids/Simulator_kernel.cuh:52-129).  Query cell from the
// current position over the grid of this frame; rho sums poly6 over every
// candidate including i itself; the spiky gradient / pho0 is summed into
// grad_i and, for j != i, its squared norm into gradj_l2.
__global__ __launch_bounds__(kBlock) void k_lambda(const float4* __restrict__ xa,
                                                   const uint2* __restrict__ ranges, int n,
                                                   StepConsts c, float4* __restrict__ xl,
                                                   float* __restrict__ rho_out) {
  const int i = xcd_block_index() * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 xi = xa[i];
  const Cell q = cell_of(xi.x, xi.y, xi.z, c);
  float rho = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f, gl2 = 0.0f;
  for_each_candidate(q, ranges, c, [&](uint32_t j) {
    float dx, dy, dz;
    const float r2 = dist2(xi, xa[j], dx, dy, dz);
    if (r2 < c.h2_outer) {
      if (r2 < c.h2) rho = rho + poly6_of(r2, c);
      const float r = sqrtf(r2);
      if (r >= c.keps && r < c.h) {
        const float t = c.h - r;
        const float s = ((c.spiky_rho0 * t) * t) / r;
        const float ax = s * dx, ay = s * dy, az = s * dz;
        gx = gx + ax;
        gy = gy + ay;
        gz = gz + az;
        if (j != (uint32_t)i) {
          gl2 = fmaf(ax, ax, gl2);
          gl2 = fmaf(ay, ay, gl2);
          gl2 = fmaf(az, az, gl2);
        }
      }
    }
  });
  if (c.k_boundary != 0.0f) rho = rho + c.k_boundary * boundary_density(xi.x, xi.y, xi.z, c);
  float denom = fmaf(gx, gx, gl2);
  denom = fmaf(gy, gy, denom);
  denom = fmaf(gz, gz, denom);
  denom = denom + c.lambda_eps;
  const float lam = -(rho / c.pho0 - 1.0f) / denom;
  xl[i] = make_float4(xi.x, xi.y, xi.z, lam);
  if (rho_out) rho_out[i] = rho;
}

// computetpos (fluids/Simulator_kernel.cuh:131-194): sum over j != i of
// (lambda_i + lambda_j + s_corr) * spikyGrad, divided by pho0, clamped to
// +-MAX_DP per axis, added, then clamped into the box shrunk by LIM_EPS.
// Reads (x, lambda) from xl and writes the new iterate to xa (Jacobi; the
// ping-pong replaces thrust::copy_n at fluids/Simulator.cu:247-248).
__global__ __launch_bounds__(kBlock) void k_delta(const float4* __restrict__ xl,
                                                  const uint2* __restrict__ ranges, int n,
                                                  StepConsts c, float4* __restrict__ xa) {
  const int i = xcd_block_index() * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 xi = xl[i];
  const Cell q = cell_of(xi.x, xi.y, xi.z, c);
  float px = 0.0f, py = 0.0f, pz = 0.0f;
  for_each_candidate(q, ranges, c, [&](uint32_t j) {
    if (j == (uint32_t)i) return;
    const float4 xj = xl[j];
    float dx, dy, dz;
    const float r2 = dist2(xi, xj, dx, dy, dz);
    if (r2 < c.h2_outer) {
      const float r = sqrtf(r2);
      if (r >= c.keps && r < c.h) {
        const float w = r2 < c.h2 ? poly6_of(r2, c) : 0.0f;
        const float corr = c.coef_corr * pow_corr(w, c);
        const float t = c.h - r;
        const float gs = ((c.spiky * t) * t) / r;
        const float s = ((xi.w + xj.w) + corr) * gs;
        px = fmaf(s, dx, px);
        py = fmaf(s, dy, py);
        pz = fmaf(s, dz, pz);
      }
    }
  });
  px = fminf(fmaxf(px / c.pho0, -c.max_dp), c.max_dp);
  py = fminf(fmaxf(py / c.pho0, -c.max_dp), c.max_dp);
  pz = fminf(fmaxf(pz / c.pho0, -c.max_dp), c.max_dp);
  const float x = fmaxf(fminf(xi.x + px, c.hix), c.lox);
  const float y = fmaxf(fminf(xi.y + py, c.hiy), c.loy);
  const float z = fmaxf(fminf(xi.z + pz, c.hiz), c.loz);
  xa[i] = make_float4(x, y, z, 0.0f);
}

// h_updateVelocity (fluids/Simulator.cu:127-137, :267-274):
// v = (x_final - x_start) * (1/dt).  Also writes the caller's pos / vel /
// iid arrays with their end-of-step contents (sorted start positions,
// pre-XSPH velocities, sorted identities) and packs (v, rho) for XSPH.
template <int L>
__global__ __launch_bounds__(kBlock) void k_velocity(
    const float4* __restrict__ xa, const float4* __restrict__ x0s,
    const uint32_t* __restrict__ iid_s, const float* __restrict__ rho, int n, StepConsts c,
    float4* __restrict__ vr, float* __restrict__ pos, float* __restrict__ vel,
    uint32_t* __restrict__ iid) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  const float4 x = xa[s];
  const float4 x0 = x0s[s];
  const float vx = (x.x - x0.x) * c.inv_dt;
  const float vy = (x.y - x0.y) * c.inv_dt;
  const float vz = (x.z - x0.z) * c.inv_dt;
  vr[s] = make_float4(vx, vy, vz, rho[s]);
  store_vec<L>(vel, s, vx, vy, vz);
  store_vec<L>(pos, s, x0.x, x0.y, x0.z);
  iid[s] = iid_s[s];
}

// computeXSPH (fluids/Simulator_kernel.cuh:196-239): nvel_i = v_i +
// c * sum_j 2 (v_j - v_i) W(|x_i - x_j|^2) / (rho_i + rho_j), j == i
// included (its term is 0), rho of the last lambda pass.  Pairs with W == 0
// are skipped: their reference term is exactly 0.
template <int L>
__global__ __launch_bounds__(kBlock) void k_xsph(const float4* __restrict__ xa,
                                                 const float4* __restrict__ vr,
                                                 const uint2* __restrict__ ranges, int n,
                                                 StepConsts c, float* __restrict__ npos,
                                                 float* __restrict__ nvel) {
  const int i = xcd_block_index() * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 xi = xa[i];
  const float4 vi = vr[i];
  const Cell q = cell_of(xi.x, xi.y, xi.z, c);
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  for_each_candidate(q, ranges, c, [&](uint32_t j) {
    float dx, dy, dz;
    const float r2 = dist2(xi, xa[j], dx, dy, dz);
    if (r2 < c.h2) {
      const float w = poly6_of(r2, c);
      const float4 vj = vr[j];
      const float f = (2.0f * w) / (vi.w + vj.w);
      ax = fmaf(vj.x - vi.x, f, ax);
      ay = fmaf(vj.y - vi.y, f, ay);
      az = fmaf(vj.z - vi.z, f, az);
    }
  });
  store_vec<L>(nvel, i, fmaf(c.c_xsph, ax, vi.x), fmaf(c.c_xsph, ay, vi.y),
               fmaf(c.c_xsph, az, vi.z));
  store_vec<L>(npos, i, xi.x, xi.y, xi.z);
}

constexpr int kDensBlock = 256;

// |rho_i/pho0 - 1| on a fresh grid, block partial sums and maxima in FP64.
__global__ __launch_bounds__(kDensBlock) void k_density_error(const float4* __restrict__ xa,
                                                              const uint2* __restrict__ ranges,
                                                              int n, StepConsts c,
                                                              double* __restrict__ partial) {
  __shared__ double ssum[kDensBlock];
  __shared__ double smax[kDensBlock];
  const int i = blockIdx.x * kDensBlock + threadIdx.x;
  double e = 0.0;
  if (i < n) {
    const float4 xi = xa[i];
    const Cell q = cell_of(xi.x, xi.y, xi.z, c);
    float rho = 0.0f;
    for_each_candidate(q, ranges, c, [&](uint32_t j) {
      float dx, dy, dz;
      const float r2 = dist2(xi, xa[j], dx, dy, dz);
      if (r2 < c.h2) rho = rho + poly6_of(r2, c);
    });
    if (c.k_boundary != 0.0f) rho = rho + c.k_boundary * boundary_density(xi.x, xi.y, xi.z, c);
    e = fabs((double)rho / (double)c.pho0 - 1.0);
  }
  ssum[threadIdx.x] = e;
  smax[threadIdx.x] = e;
  __syncthreads();
  for (int off = kDensBlock / 2; off > 0; off >>= 1) {
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

// Fixed-order reduction of the block partials (deterministic).
__global__ __launch_bounds__(kDensBlock) void k_reduce_partials(const double* __restrict__ partial,
                                                                int nb, int n,
                                                                double* __restrict__ out) {
  __shared__ double ssum[kDensBlock];
  __shared__ double smax[kDensBlock];
  double s = 0.0, m = 0.0;
  for (int b = threadIdx.x; b < nb; b += kDensBlock) {
    s += partial[2 * b];
    m = fmax(m, partial[2 * b + 1]);
  }
  ssum[threadIdx.x] = s;
  smax[threadIdx.x] = m;
  __syncthreads();
  for (int off = kDensBlock / 2; off > 0; off >>= 1) {
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

void launch_predict_key(int layout, const float* pos, const float* vel, int n, const StepConsts& c,
                        float4* xp, uint32_t* key, uint32_t* val, hipStream_t st) {
  if (layout == 3)
    k_predict_key<3><<<blocks_for(n), kBlock, 0, st>>>(pos, vel, n, c, xp, key, val);
  else
    k_predict_key<4><<<blocks_for(n), kBlock, 0, st>>>(pos, vel, n, c, xp, key, val);
}

void launch_key_only(int layout, const float* pos, int n, const StepConsts& c, float4* xp,
                     uint32_t* key, uint32_t* val, hipStream_t st) {
  if (layout == 3)
    k_key_only<3><<<blocks_for(n), kBlock, 0, st>>>(pos, n, c, xp, key, val);
  else
    k_key_only<4><<<blocks_for(n), kBlock, 0, st>>>(pos, n, c, xp, key, val);
}

void launch_clear_ranges(uint2* ranges, long long cells, hipStream_t st) {
  long long blocks = ((cells >> 1) + kBlock - 1) / kBlock;
  if (blocks < 1) blocks = 1;
  if (blocks > 8192) blocks = 8192;
  k_clear_ranges<<<(int)blocks, kBlock, 0, st>>>(ranges, cells);
}

void launch_permute(int layout, const uint32_t* skey, const uint32_t* sval, const float* pos,
                    const float4* xp, const uint32_t* iid, int n, int cells, float4* x0s,
                    float4* xa, uint32_t* iid_s, uint2* ranges, hipStream_t st) {
  if (layout == 3)
    k_permute<3><<<blocks_for(n), kBlock, 0, st>>>(skey, sval, pos, xp, iid, n, cells, x0s, xa,
                                                   iid_s, ranges);
  else
    k_permute<4><<<blocks_for(n), kBlock, 0, st>>>(skey, sval, pos, xp, iid, n, cells, x0s, xa,
                                                   iid_s, ranges);
}

void launch_permute_pos(const uint32_t* skey, const uint32_t* sval, const float4* xp, int n,
                        int cells, float4* xa, uint2* ranges, hipStream_t st) {
  k_permute_pos<<<blocks_for(n), kBlock, 0, st>>>(skey, sval, xp, n, cells, xa, ranges);
}

void launch_lambda(const float4* xa, const uint2* ranges, int n, const StepConsts& c, float4* xl,
                   float* rho, hipStream_t st) {
  k_lambda<<<blocks_for(n), kBlock, 0, st>>>(xa, ranges, n, c, xl, rho);
}

void launch_delta(const float4* xl, const uint2* ranges, int n, const StepConsts& c, float4* xa,
                  hipStream_t st) {
  k_delta<<<blocks_for(n), kBlock, 0, st>>>(xl, ranges, n, c, xa);
}

void launch_velocity(int layout, const float4* xa, const float4* x0s, const uint32_t* iid_s,
                     const float* rho, int n, const StepConsts& c, float4* vr, float* pos,
                     float* vel, uint32_t* iid, hipStream_t st) {
  if (layout == 3)
    k_velocity<3><<<blocks_for(n), kBlock, 0, st>>>(xa, x0s, iid_s, rho, n, c, vr, pos, vel, iid);
  else
    k_velocity<4><<<blocks_for(n), kBlock, 0, st>>>(xa, x0s, iid_s, rho, n, c, vr, pos, vel, iid);
}

void launch_xsph(int layout, const float4* xa, const float4* vr, const uint2* ranges, int n,
                 const StepConsts& c, float* npos, float* nvel, hipStream_t st) {
  if (layout == 3)
    k_xsph<3><<<blocks_for(n), kBlock, 0, st>>>(xa, vr, ranges, n, c, npos, nvel);
  else
    k_xsph<4><<<blocks_for(n), kBlock, 0, st>>>(xa, vr, ranges, n, c, npos, nvel);
}

int density_blocks(int n) { return (n + kDensBlock - 1) / kDensBlock; }

void launch_density_error(const float4* xa, const uint2* ranges, int n, const StepConsts& c,
                          double* partial, double* out, hipStream_t st) {
  const int nb = density_blocks(n);
  k_density_error<<<nb, kDensBlock, 0, st>>>(xa, ranges, n, c, partial);
  k_reduce_partials<<<1, kDensBlock, 0, st>>>(partial, nb, n, out);
}

}  // namespace pbf
