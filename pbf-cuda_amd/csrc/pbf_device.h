// pbf_device.h — shared device-side definitions of the gfx950 PBF step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pbf {

// Per-step constants, built on the host once per step and passed by value
// (kernarg segment).  Field meanings follow the reference's functors
// (fluids/Simulator.cu:22-161) and constants (fluids/helper.h:4-9).
struct StepConsts {
  float llx, lly, llz;     // box lower corner (llim)
  float ulx, uly, ulz;     // box upper corner (ulim)
  int dx, dy, dz;          // grid dims, (int)ceilf((ulim-llim)/h)
  int cells;               // dx*dy*dz
  float h, h2;             // kernel radius, h*h
  float h2_outer;          // prefilter bound: r2 >= h2_outer => both kernels are 0
  float keps;              // KERNAL_EPS
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
  float lox, loy, loz;     // box clamp lower: llim + LIM_EPS, rounded to FP32
  float hix, hiy, hiz;     // box clamp upper: ulim - LIM_EPS, rounded to FP32
  float max_dp;            // MAX_DP
  float dt, inv_dt;        // dt, 1/dt (FP32, as h_updateVelocity)
  float gz;                // -g
  float c_xsph;
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

__device__ __forceinline__ Cell cell_of(float x, float y, float z, const StepConsts& c) {
  return Cell{cell_coord(x, c.llx, c.h, c.dx), cell_coord(y, c.lly, c.h, c.dy),
              cell_coord(z, c.llz, c.h, c.dz)};
}

// Cell id, x-major and z fastest (xyzToId, fluids/Simulator.cu:52).
__device__ __forceinline__ uint32_t cell_id(const Cell& q, const StepConsts& c) {
  return (uint32_t)((q.x * c.dy + q.y) * c.dz + q.z);
}

// Visits the candidate neighbours of a particle whose query cell is q, in
// the reference's order (fluids/Simulator_kernel.cuh:78-88): dx, then dy,
// then dz, j ascending.  The three cells (x,y,z-1..z+1) of one column are
// consecutive cell ids, so their particles form one contiguous sorted range:
// [first start, last end) over the column's non-empty cells.  Walking that
// range is the same j sequence as walking the three cells in turn.
template <class F>
__device__ __forceinline__ void for_each_candidate(const Cell& q, const uint2* __restrict__ ranges,
                                                   const StepConsts& c, F&& f) {
  const int z0 = q.z > 0 ? q.z - 1 : 0;
  const int z1 = q.z + 1 < c.dz ? q.z + 1 : c.dz - 1;
#pragma unroll 1
  for (int ox = -1; ox <= 1; ++ox) {
    const int X = q.x + ox;
    if (X < 0 || X >= c.dx) continue;
#pragma unroll 1
    for (int oy = -1; oy <= 1; ++oy) {
      const int Y = q.y + oy;
      if (Y < 0 || Y >= c.dy) continue;
      const int col = (X * c.dy + Y) * c.dz;
      uint32_t b = 0xffffffffu, e = 0u;
      for (int z = z0; z <= z1; ++z) {
        const uint2 r = ranges[col + z];
        if (r.x < r.y) {
          b = r.x < b ? r.x : b;
          e = r.y > e ? r.y : e;
        }
      }
#pragma unroll 1
      for (uint32_t j = b; j < e; ++j) f(j);
    }
  }
}

// x^n for the s_corr exponent: repeated multiplication for small integer n
// (the reference default n=4), powf otherwise.
__device__ __forceinline__ float pow_corr(float w, const StepConsts& c) {
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

__device__ __forceinline__ float boundary_density(float x, float y, float z, const StepConsts& c) {
  float s = wall_density_at(c.ulx - x, c);
  s = s + wall_density_at(x - c.llx, c);
  s = s + wall_density_at(c.uly - y, c);
  s = s + wall_density_at(y - c.lly, c);
  s = s + wall_density_at(c.ulz - z, c);
  s = s + wall_density_at(z - c.llz, c);
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

}  // namespace pbf
