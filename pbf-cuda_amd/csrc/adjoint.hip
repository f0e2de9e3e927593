// adjoint.hip — reverse-mode derivative of one PBF frame on gfx950.
//
// The frame is differentiated with its discrete choices held fixed (sort
// order, stale grid, candidate lists, kernel cutoffs, active clamps), along
// the trajectory the forward kernels computed, which a tape records: per
// solver iteration k the positions and lambdas xl[k], the lambda pass's
// (sum of gradients S, denominator D) and densities, and the correction
// before its clamps.  Derivatives follow the reference's formulas
// (fluids/Simulator_kernel.cuh:52-239): lambda_i = -(rho_i/pho0 - 1)/D_i with
// D_i = sum_{j!=i} |g_ij|^2 + |S_i|^2 + eps, g_ij = spikyGrad(x_i-x_j)/pho0;
// dp_i = sum_{j!=i} (lambda_i + lambda_j + s_corr) spikyGrad / pho0, clamped;
// XSPH v'_i = v_i + c sum_j 2 (v_j - v_i) W_ij / (rho_i + rho_j).
//
// Every backward neighbour pass is a pass policy of pbf_pass.h.  A pair term
// of target i (j in its candidate list cand(i)) depends on x_i and x_j; its
// derivatives belong to both.  The stale grid makes candidate lists
// one-sided (j in cand(i) does not imply i in cand(j)), and two modes
// collect the terms (see the comment before within_one below):
//  * scatter: i sweeps cand(i) on the sorted grid, gathers its own terms
//    and, where i is in cand(j), j's term about i; where it is not, i adds
//    d(its term)/dx_j to j with float atomics into "extra" arrays that a
//    later kernel folds in (their order of addition depends on timing);
//  * gather: i sweeps a grid built on the pass's own positions, which holds
//    every particle within h, and gathers every term that concerns it; no
//    atomics, every sum in a fixed order.
#include <utility>

#include "pbf_adjoint.h"
#include "pbf_pass.h"

namespace pbf {

namespace {

// Sorted (stale-grid) cell of a particle from its sorted key.
__device__ __forceinline__ Cell cell_of_key(uint32_t key, const SceneGeo& g) {
  const uint32_t local = key - g.base;
  const uint32_t yz = (uint32_t)(g.dy * g.dz);
  const uint32_t x = local / yz;
  const uint32_t r = local - x * yz;
  const uint32_t y = r / (uint32_t)g.dz;
  return Cell{(int)x, (int)y, (int)(r - y * (uint32_t)g.dz)};
}

// Query cells packed 10 bits per axis (x << 20 | y << 10 | z), valid when
// every grid dimension is at most 1024; a pointwise kernel computes them once
// per particle and position set, so the pair terms compare bit fields
// instead of dividing three coordinates per pair.
__device__ __forceinline__ bool packed_query_cells(const StepConsts& c) {
  return c.g.dx <= 1024 && c.g.dy <= 1024 && c.g.dz <= 1024;
}
__device__ __forceinline__ uint32_t pack_cell(const Cell& q) {
  return ((uint32_t)q.x << 20) | ((uint32_t)q.y << 10) | (uint32_t)q.z;
}

// Whether a particle whose sorted cell is ci is a candidate of the particle
// at position xj (ci in the 3x3x3 block around xj's query cell); qj is xj's
// packed query cell (read only when the packing is valid).
__device__ __forceinline__ bool is_candidate_of(const Cell& ci, const float4& xj, uint32_t qj,
                                                const StepConsts& c) {
  if (packed_query_cells(c)) {
    const int qx = (int)(qj >> 20), qy = (int)((qj >> 10) & 1023u), qz = (int)(qj & 1023u);
    return abs(ci.x - qx) <= 1 && abs(ci.y - qy) <= 1 && abs(ci.z - qz) <= 1;
  }
  const Cell q = cell_of(xj.x, xj.y, xj.z, c.g, c.h);
  return abs(ci.x - q.x) <= 1 && abs(ci.y - q.y) <= 1 && abs(ci.z - q.z) <= 1;
}

// dW/dx_i of poly6 at d = x_i - x_j (r2 < h2): -6 poly6 (h2 - r2)^2 d.
__device__ __forceinline__ float3 poly6_grad(const float3& d, float r2, const StepConsts& c) {
  const float t = c.h2 - r2;
  const float f = -6.0f * c.poly6 * t * t;
  return make_float3(f * d.x, f * d.y, f * d.z);
}

// 1/x to about one ulp (v_rcp_f32 and one Newton step).  The adjoint only
// has to agree with the FP64 adjoint to a relative 1e-4, so each pair
// divides once and multiplies by the reciprocal where the forward formulas
// divide.
__device__ __forceinline__ float recip(float x) {
  const float y = __builtin_amdgcn_rcpf(x);
  return fmaf(fmaf(-x, y, 1.0f), y, y);
}

// Hessian of k (h - r)^2 d / r (the spiky gradient with coefficient k):
// phi I + psi d d^T, phi = k u^2 / r, psi = -k u (h + r) / r^3, u = h - r.
struct SpikyHess {
  float phi, psi;
};
__device__ __forceinline__ SpikyHess spiky_hess(float k, float r, float inv_r,
                                                const StepConsts& c) {
  const float u = c.h - r;
  return SpikyHess{k * u * u * inv_r, -k * u * (c.h + r) * (inv_r * inv_r * inv_r)};
}
__device__ __forceinline__ float3 hess_apply(const SpikyHess& H, const float3& d,
                                             const float3& a) {
  const float pd = H.psi * (d.x * a.x + d.y * a.y + d.z * a.z);
  return make_float3(H.phi * a.x + pd * d.x, H.phi * a.y + pd * d.y, H.phi * a.z + pd * d.z);
}

// d s_corr / d w for s_corr = coef_corr * w^n_corr.
__device__ __forceinline__ float dscorr_dw(float w, const StepConsts& c) {
  if (c.n_corr_int == 0 || c.n_corr == 0.0f) return 0.0f;
  if (c.n_corr_int > 0) {
    float r = 1.0f;
    for (int e = 1; e < c.n_corr_int; ++e) r = r * w;
    return c.coef_corr * c.n_corr * r;
  }
  return w > 0.0f ? c.coef_corr * c.n_corr * powf(w, c.n_corr - 1.0f) : 0.0f;
}

__device__ __forceinline__ float dot3(const float3& a, const float3& b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ float3 xyz(const float4& v) { return make_float3(v.x, v.y, v.z); }

// The scattered pair sums: kScatterCopies copies of n float4 entries, stored
// component-major ([component][copy][n] floats) so the four atomic adds of
// one scatter land on four cache lines.  A workgroup adds to copy
// blockIdx.x % kScatterCopies.
__device__ __forceinline__ size_t scatter_copy(uint32_t n) {
  return (size_t)(blockIdx.x % kScatterCopies) * n;
}
// Adds (x, y, z[, w]) at copy offset e (scatter_copy(n) + j).
__device__ __forceinline__ void scatter_add(float4* ex, uint32_t n, size_t e, float x, float y,
                                            float z, float w, bool with_w) {
  float* f = reinterpret_cast<float*>(ex);
  const size_t cs = (size_t)kScatterCopies * n;
  atomicAdd(&f[e], x);
  atomicAdd(&f[cs + e], y);
  atomicAdd(&f[2 * cs + e], z);
  if (with_w) atomicAdd(&f[3 * cs + e], w);
}
// Sum of the copies of entry i (in copy order), which are cleared.
__device__ __forceinline__ float4 take_scattered(float4* __restrict__ ex, int i, int n) {
  float* f = reinterpret_cast<float*>(ex);
  const size_t cs = (size_t)kScatterCopies * n;
  float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int k = 0; k < kScatterCopies; ++k) {
      const size_t idx = q * cs + (size_t)k * n + i;
      v[q] = v[q] + f[idx];
      f[idx] = 0.0f;
    }
  return make_float4(v[0], v[1], v[2], v[3]);
}

// Two modes share each backward pass policy:
//  * scatter (kGather = false): targets in the frame's sorted order sweep the
//    sorted (stale) grid around their query cell, so every candidate's own
//    term counts; j's term about i is gathered when i is a candidate of j
//    and otherwise i's term about j is scattered to j (float atomics);
//  * gather (kGather = true): targets and candidates in the order of a grid
//    built on the pass's own positions (sorted by query cell).  A target
//    sweeps that grid around its query cell, which holds every particle
//    within h of it, and tests both memberships per pair from the packed
//    (sorted cell, query cell) of each slot: its own term counts when j is
//    its candidate, j's term about it when it is j's candidate.  Every term
//    is gathered by the particle it belongs to, in a fixed order: no atomics.
//    A pair within h whose query cells are two apart (possible only through
//    rounding when |x_i - x_j| is within ~1e-6 h of h, where every term and
//    derivative is ~0) is the only pair this sweep can miss.

// Whether packed cells a and b are at most one cell apart on every axis.
__device__ __forceinline__ bool within_one(uint32_t a, uint32_t b) {
  const int dx = (int)(a >> 20) - (int)(b >> 20);
  const int dy = (int)((a >> 10) & 1023u) - (int)((b >> 10) & 1023u);
  const int dz = (int)(a & 1023u) - (int)(b & 1023u);
  return (uint32_t)(dx + 1) <= 2u && (uint32_t)(dy + 1) <= 2u && (uint32_t)(dz + 1) <= 2u;
}

// Memberships of a pair (target t, candidate j): own = j is a candidate of
// t (t's own term about j exists), other = t is a candidate of j.
struct Members {
  bool own, other;
};
template <bool kGather>
__device__ __forceinline__ Members members(const Cell& ct, uint2 tc, const float4& xj, uint2 jc,
                                           const StepConsts& c) {
  if constexpr (kGather) {
    // packed (sorted cell, query cell) of t and j
    return Members{within_one(jc.x, tc.y), within_one(tc.x, jc.y)};
  } else {
    // the sweep ran over t's candidates; jc.y: j's packed query cell
    return Members{true, is_candidate_of(ct, xj, jc.y, c)};
  }
}

// ---- XSPH -----------------------------------------------------------------
template <bool kGather>
struct BwdXsphPass {
  // scatter mode: workgroups in launch order, round-robin over the XCDs, so a
  // crowded cell's targets scatter from all eight XCDs, each into its own
  // copy; gather mode: one contiguous range per XCD
  static constexpr bool kXcdRanges = kGather;
  // hits reloaded per round of the sweep
  static constexpr int kHitBatch = 4;
  struct Args {
    const float4* __restrict__ x;    // final positions
    const float4* __restrict__ vr;   // (v before XSPH, rho of the last lambda pass)
    const float4* __restrict__ gv;   // adjoint of v after XSPH
    const uint32_t* __restrict__ key;  // scatter: sorted keys
    float4* __restrict__ out_x;      // gathered (dL/dx, dL/drho)
    float4* __restrict__ out_v;      // gathered dL/dv (pair terms)
    float4* __restrict__ ex_x;       // scatter: (dL/dx, dL/drho), kScatterCopies x ex_n
    float4* __restrict__ ex_v;       // scatter: dL/dv, kScatterCopies x ex_n
    uint32_t ex_n;
    const uint32_t* __restrict__ qc;  // scatter: packed query cells of x
    const uint2* __restrict__ cq;     // gather: packed (sorted cell, query cell)
  };
  struct Target {
    float4 x, vr, gv;
    Cell cell;
    uint2 c;
    uint32_t i;
  };
  struct Acc {
    float vx = 0.0f, vy = 0.0f, vz = 0.0f, xx = 0.0f, xy = 0.0f, xz = 0.0f, r = 0.0f;
  };
  struct Nb {
    float4 x, vr, gv;
    uint2 c;
  };
  static __device__ __forceinline__ Target target(const Args& a, uint32_t i, const StepConsts& c) {
    if constexpr (kGather)
      return Target{a.x[i], a.vr[i], a.gv[i], Cell{}, a.cq[i], i};
    else
      return Target{a.x[i], a.vr[i], a.gv[i], cell_of_key(a.key[i], c.g), make_uint2(0u, 0u), i};
  }
  static __device__ __forceinline__ const float4* screen(const Args& a) { return a.x; }
  static __device__ __forceinline__ Nb load_f(const Args& a, uint32_t j) {
    if constexpr (kGather)
      return Nb{a.x[j], a.vr[j], a.gv[j], a.cq[j]};
    else
      return Nb{a.x[j], a.vr[j], a.gv[j], make_uint2(0u, a.qc[j])};
  }
  static __device__ __forceinline__ bool hit(const Target& t, const float3& xj,
                                             const StepConsts& c) {
    float dx, dy, dz;
    return dist2(t.x, xj, dx, dy, dz) < c.h2;
  }
  static __device__ __forceinline__ void add(Acc& s, const Target& t, uint32_t j, const Nb& nb,
                                             const StepConsts& c, const Args& a) {
    float dx, dy, dz;
    const float r2 = dist2(t.x, nb.x, dx, dy, dz);
    if (!(r2 < c.h2)) return;
    const Members mb = members<kGather>(t.cell, t.c, nb.x, nb.c, c);
    const float3 d = make_float3(dx, dy, dz);
    const float W = poly6_of(r2, c);
    const float3 P = poly6_grad(d, r2, c);
    const float inv_den = recip(t.vr.w + nb.vr.w);  // 1 / (rho_i + rho_j)
    const float k2 = 2.0f * c.c_xsph * inv_den;
    const float f = k2 * W;
    const float3 dv = make_float3(nb.vr.x - t.vr.x, nb.vr.y - t.vr.y, nb.vr.z - t.vr.z);
    const float kr = f * inv_den;  // 2 c W / (rho_i + rho_j)^2
    // own term of v'_i
    const float3 B = xyz(t.gv);
    const float u = dot3(B, dv);
    const float fx = k2 * u;
    const float fr = kr * u;
    if (mb.own) {
      s.vx -= f * B.x;
      s.vy -= f * B.y;
      s.vz -= f * B.z;
      s.xx += fx * P.x;
      s.xy += fx * P.y;
      s.xz += fx * P.z;
      s.r -= fr;
    }
    if (mb.other) {
      // j's term about i
      const float3 Bj = xyz(nb.gv);
      const float uj = -dot3(Bj, dv);
      const float gx = k2 * uj;
      s.vx += f * Bj.x;
      s.vy += f * Bj.y;
      s.vz += f * Bj.z;
      s.xx += gx * P.x;
      s.xy += gx * P.y;
      s.xz += gx * P.z;
      s.r -= kr * uj;
    } else if constexpr (!kGather) {
      if (j != t.i) {
        // j never visits i: deliver the own term's derivatives about j
        const size_t e = scatter_copy(a.ex_n) + j;
        scatter_add(a.ex_v, a.ex_n, e, f * B.x, f * B.y, f * B.z, 0.0f, false);
        scatter_add(a.ex_x, a.ex_n, e, -fx * P.x, -fx * P.y, -fx * P.z, -fr, true);
      }
    }
  }
  static __device__ __forceinline__ void accumulate(Acc& tot, const Acc& p) {
    tot.vx += p.vx;
    tot.vy += p.vy;
    tot.vz += p.vz;
    tot.xx += p.xx;
    tot.xy += p.xy;
    tot.xz += p.xz;
    tot.r += p.r;
  }
  static __device__ __forceinline__ void finish(const Args& a, const Target& t, const Acc& s,
                                                const StepConsts&, const SceneGeo&) {
    a.out_x[t.i] = make_float4(s.xx, s.xy, s.xz, s.r);
    a.out_v[t.i] = make_float4(s.vx, s.vy, s.vz, 0.0f);
  }
};

// ---- delta-p ----------------------------------------------------------------
template <bool kGather>
struct BwdDeltaPass {
  // scatter mode: round-robin workgroups (see BwdXsphPass); gather mode: one
  // contiguous range per XCD
  static constexpr bool kXcdRanges = kGather;
  // hit masks for the lambda pass of the same iteration, which sweeps the
  // same candidates of the same positions (its hit test below)
  static constexpr int kMask = kMaskStore;
  // hits reloaded per round of the sweep
  static constexpr int kHitBatch = 4;
  struct Args {
    const float4* __restrict__ xl;   // (x, lambda) of the iteration
    const float4* __restrict__ A;    // (adjoint of the correction sum (before / pho0),
                                     //  scatter: packed query cell of xl as float bits)
    const uint32_t* __restrict__ key;  // scatter: sorted keys
    float4* __restrict__ out;        // gathered (dL/dx, dL/dlambda)
    float4* __restrict__ ex;         // scatter: (dL/dx, dL/dlambda), kScatterCopies x ex_n
    uint32_t ex_n;
    const uint2* __restrict__ cq;    // gather: packed (sorted cell, query cell)
  };
  struct Target {
    float4 x, A;
    Cell cell;
    uint2 c;
    uint32_t i;
  };
  struct Acc {
    float x = 0.0f, y = 0.0f, z = 0.0f, l = 0.0f;
  };
  struct Nb {
    float4 x, A;
    uint2 c;
  };
  static __device__ __forceinline__ Target target(const Args& a, uint32_t i, const StepConsts& c) {
    if constexpr (kGather)
      return Target{a.xl[i], a.A[i], Cell{}, a.cq[i], i};
    else
      return Target{a.xl[i], a.A[i], cell_of_key(a.key[i], c.g), make_uint2(0u, 0u), i};
  }
  static __device__ __forceinline__ const float4* screen(const Args& a) { return a.xl; }
  static __device__ __forceinline__ Nb load_f(const Args& a, uint32_t j) {
    if constexpr (kGather) {
      return Nb{a.xl[j], a.A[j], a.cq[j]};
    } else {
      const float4 A = a.A[j];
      return Nb{a.xl[j], A, make_uint2(0u, __float_as_uint(A.w))};
    }
  }
  static __device__ __forceinline__ bool hit(const Target& t, const float3& xj,
                                             const StepConsts& c) {
    float dx, dy, dz;
    // i itself and coincident pairs (r2 = 0) add nothing; pairs closer than
    // KERNAL_EPS add nothing here either (add() drops them) but are hits of
    // the lambda pass, which uses these masks
    const float r2 = dist2(t.x, xj, dx, dy, dz);
    return (r2 < c.h2_outer) & (r2 > 0.0f);
  }
  static __device__ __forceinline__ void add(Acc& s, const Target& t, uint32_t j, const Nb& nb,
                                             const StepConsts& c, const Args& a) {
    if (j == t.i) return;
    float dx, dy, dz;
    const float r2 = dist2(t.x, nb.x, dx, dy, dz);
    if (!(r2 < c.h2_outer)) return;
    const float r = sqrt_pair(r2);
    if (!(r >= c.keps && r < c.h)) return;
    const Members mb = members<kGather>(t.cell, t.c, nb.x, nb.c, c);
    const float3 d = make_float3(dx, dy, dz);
    const float w = r2 < c.h2 ? poly6_of(r2, c) : 0.0f;
    const float q = (t.x.w + nb.x.w) + c.coef_corr * pow_corr(w, c);
    const SpikyHess H = spiky_hess(c.spiky, r, recip(r), c);
    const float3 G = make_float3(H.phi * d.x, H.phi * d.y, H.phi * d.z);
    float3 ds = make_float3(0.0f, 0.0f, 0.0f);
    if (r2 < c.h2) {
      const float k1 = dscorr_dw(w, c);
      const float3 P = poly6_grad(d, r2, c);
      ds = make_float3(k1 * P.x, k1 * P.y, k1 * P.z);
    }
    // own term: A_i . (q G(d))
    const float3 Ai = xyz(t.A);
    const float AG = dot3(Ai, G);
    const float3 HA = hess_apply(H, d, Ai);
    const float3 gi = make_float3(q * HA.x + AG * ds.x, q * HA.y + AG * ds.y, q * HA.z + AG * ds.z);
    if (mb.own) {
      s.x += gi.x;
      s.y += gi.y;
      s.z += gi.z;
      s.l += AG;
    }
    if (mb.other) {
      // j's term about i: A_j . (q G(x_j - x_i)) = -q A_j . G(d)
      const float3 Aj = xyz(nb.A);
      const float AGj = dot3(Aj, G);
      const float3 HAj = hess_apply(H, d, Aj);
      s.x -= q * HAj.x + AGj * ds.x;
      s.y -= q * HAj.y + AGj * ds.y;
      s.z -= q * HAj.z + AGj * ds.z;
      s.l -= AGj;
    } else if constexpr (!kGather) {
      const size_t e = scatter_copy(a.ex_n) + j;
      scatter_add(a.ex, a.ex_n, e, -gi.x, -gi.y, -gi.z, AG, true);
    }
  }
  static __device__ __forceinline__ void accumulate(Acc& tot, const Acc& p) {
    tot.x += p.x;
    tot.y += p.y;
    tot.z += p.z;
    tot.l += p.l;
  }
  static __device__ __forceinline__ void finish(const Args& a, const Target& t, const Acc& s,
                                                const StepConsts&, const SceneGeo&) {
    a.out[t.i] = make_float4(s.x, s.y, s.z, s.l);
  }
};

// ---- lambda (density and gradient norms) ------------------------------------
template <bool kGather>
struct BwdLambdaPass {
  // scatter mode: round-robin workgroups (see BwdXsphPass); gather mode: one
  // contiguous range per XCD
  static constexpr bool kXcdRanges = kGather;
  // the delta-p pass's hit masks (its hit test is this pass's)
  static constexpr int kMask = kMaskLoad;
  // hits reloaded per round of the sweep
  static constexpr int kHitBatch = 4;
  struct Args {
    const float4* __restrict__ xr;   // (x, dL/drho)
    const float4* __restrict__ sd;   // (S, dL/dD)
    const uint32_t* __restrict__ key;  // scatter: sorted keys
    const float4* gin;  // adjoint of x before the pass (sorted order)
    float4* gout;       // gin + gathered (sorted order; may alias gin)
    float4* __restrict__ ex;         // scatter: dL/dx, kScatterCopies x ex_n
    uint32_t ex_n;
    const uint32_t* __restrict__ qc;   // scatter: packed query cells of x
    const uint2* __restrict__ cq;      // gather: packed (sorted cell, query cell)
    const uint32_t* __restrict__ slot;  // gather: sorted slot of each query slot
  };
  struct Target {
    float4 x, sd;
    Cell cell;
    uint2 c;
    uint32_t i;
  };
  struct Acc {
    float x = 0.0f, y = 0.0f, z = 0.0f;
  };
  struct Nb {
    float4 x, sd;
    uint2 c;
  };
  static __device__ __forceinline__ Target target(const Args& a, uint32_t i, const StepConsts& c) {
    if constexpr (kGather)
      return Target{a.xr[i], a.sd[i], Cell{}, a.cq[i], i};
    else
      return Target{a.xr[i], a.sd[i], cell_of_key(a.key[i], c.g), make_uint2(0u, 0u), i};
  }
  static __device__ __forceinline__ const float4* screen(const Args& a) { return a.xr; }
  static __device__ __forceinline__ Nb load_f(const Args& a, uint32_t j) {
    if constexpr (kGather)
      return Nb{a.xr[j], a.sd[j], a.cq[j]};
    else
      return Nb{a.xr[j], a.sd[j], make_uint2(0u, a.qc[j])};
  }
  static __device__ __forceinline__ bool hit(const Target& t, const float3& xj,
                                             const StepConsts& c) {
    float dx, dy, dz;
    // coincident pairs (and i itself) have d = 0: every term below vanishes
    // (the delta-p pass stores masks of this test)
    const float r2 = dist2(t.x, xj, dx, dy, dz);
    return (r2 < c.h2_outer) & (r2 > 0.0f);
  }
  static __device__ __forceinline__ void add(Acc& s, const Target& t, uint32_t j, const Nb& nb,
                                             const StepConsts& c, const Args& a) {
    float dx, dy, dz;
    const float r2 = dist2(t.x, nb.x, dx, dy, dz);
    if (!(r2 < c.h2_outer)) return;
    const float3 d = make_float3(dx, dy, dz);
    const Members mb = members<kGather>(t.cell, t.c, nb.x, nb.c, c);
    float3 gi = make_float3(0.0f, 0.0f, 0.0f);  // own terms' d/dx_i
    float3 gr = make_float3(0.0f, 0.0f, 0.0f);  // j's terms' d/dx_i
    if (r2 < c.h2) {
      const float3 P = poly6_grad(d, r2, c);
      gi = make_float3(t.x.w * P.x, t.x.w * P.y, t.x.w * P.z);
      gr = make_float3(nb.x.w * P.x, nb.x.w * P.y, nb.x.w * P.z);
    }
    const float r = sqrt_pair(r2);
    if (r >= c.keps && r < c.h && j != t.i) {
      const SpikyHess H = spiky_hess(c.spiky_rho0, r, recip(r), c);
      const float3 g = make_float3(H.phi * d.x, H.phi * d.y, H.phi * d.z);
      const float3 Vi = make_float3(g.x + t.sd.x, g.y + t.sd.y, g.z + t.sd.z);
      const float3 Hi = hess_apply(H, d, Vi);
      const float ki = 2.0f * t.sd.w;
      gi = make_float3(gi.x + ki * Hi.x, gi.y + ki * Hi.y, gi.z + ki * Hi.z);
      const float3 Vj = make_float3(g.x - nb.sd.x, g.y - nb.sd.y, g.z - nb.sd.z);
      const float3 Hj = hess_apply(H, d, Vj);
      const float kj = 2.0f * nb.sd.w;
      gr = make_float3(gr.x + kj * Hj.x, gr.y + kj * Hj.y, gr.z + kj * Hj.z);
    }
    if (mb.own) {
      s.x += gi.x;
      s.y += gi.y;
      s.z += gi.z;
    }
    if (mb.other) {
      s.x += gr.x;
      s.y += gr.y;
      s.z += gr.z;
    } else if constexpr (!kGather) {
      if (j != t.i)
        scatter_add(a.ex, a.ex_n, scatter_copy(a.ex_n) + j, -gi.x, -gi.y, -gi.z, 0.0f, false);
    }
  }
  static __device__ __forceinline__ void accumulate(Acc& tot, const Acc& p) {
    tot.x += p.x;
    tot.y += p.y;
    tot.z += p.z;
  }
  static __device__ __forceinline__ void finish(const Args& a, const Target& t, const Acc& s,
                                                const StepConsts&, const SceneGeo&) {
    const uint32_t k = kGather ? a.slot[t.i] : t.i;
    const float4 g = a.gin[k];
    a.gout[k] = make_float4(g.x + s.x, g.y + s.y, g.z + s.z, 0.0f);
  }
};

// ---- pointwise kernels -------------------------------------------------------
// In gather mode the pass outputs are in query order: work-item u handles
// query slot u, whose sorted slot is perm[u]; sorted-order arrays are read
// and written at perm[u].  In scatter mode both orders are the sorted one.

// After the XSPH pass: dL/dx_final, dL/dx_start (from v = (x - x0)/dt) and
// dL/drho of the last lambda pass; scatter mode: adds and clears the extras.
template <bool kGather>
__global__ __launch_bounds__(kBlock) void k_bwd_velocity(
    int n, StepConsts c, const uint32_t* __restrict__ perm, const float4* __restrict__ gx_out,
    const float4* __restrict__ gv_out, const float4* __restrict__ out_x,
    const float4* __restrict__ out_v, float4* __restrict__ ex_x, float4* __restrict__ ex_v,
    float4* __restrict__ gx, float4* __restrict__ gx0, float* __restrict__ grho) {
  const int u = blockIdx.x * kBlock + threadIdx.x;
  if (u >= n) return;
  const int i = kGather ? (int)perm[u] : u;
  float4 ex = make_float4(0.0f, 0.0f, 0.0f, 0.0f), ev = ex;
  if constexpr (!kGather) {
    ex = take_scattered(ex_x, u, n);
    ev = take_scattered(ex_v, u, n);
  }
  const float4 ox = out_x[u], ov = out_v[u], go = gx_out[i], gvo = gv_out[i];
  const float vx = gvo.x + ov.x + ev.x, vy = gvo.y + ov.y + ev.y, vz = gvo.z + ov.z + ev.z;
  gx[i] = make_float4(go.x + ox.x + ex.x + vx * c.inv_dt, go.y + ox.y + ex.y + vy * c.inv_dt,
                      go.z + ox.z + ex.z + vz * c.inv_dt, 0.0f);
  gx0[i] = make_float4(-vx * c.inv_dt, -vy * c.inv_dt, -vz * c.inv_dt, 0.0f);
  grho[i] = ox.w + ex.w;
}

// Query cell of position x: written packed (qc, scatter mode) or as the sort
// key of the pass's grid with its sorted slot (qkey, qval, gather mode).
__device__ __forceinline__ uint32_t note_query_cell(int i, const float4& x, const StepConsts& c,
                                                    uint32_t* __restrict__ qc,
                                                    uint32_t* __restrict__ qkey,
                                                    uint32_t* __restrict__ qval) {
  const Cell q = cell_of(x.x, x.y, x.z, c.g, c.h);
  const uint32_t packed = pack_cell(q);
  if (qc) qc[i] = packed;
  if (qkey) {
    qkey[i] = cell_id(q, c.g);
    qval[i] = (uint32_t)i;
  }
  return packed;
}

// Position update of iteration k (fluids/Simulator_kernel.cuh:184-192):
// x' = clamp_box(x + clamp(dp, +-MAX_DP)).  gx = dL/dx' -> gxn = identity
// part of dL/dx, A = dL/d(correction sum), r0 += dL/dpho0 through the / pho0;
// A.w holds the packed query cell of x, also noted for the passes (see
// note_query_cell).
__global__ __launch_bounds__(kBlock) void k_bwd_clamp(
    int n, StepConsts c, const float4* __restrict__ gx, const float4* __restrict__ xl,
    const float4* __restrict__ dp, float4* __restrict__ gxn, float4* __restrict__ A,
    uint32_t* __restrict__ qc, uint32_t* __restrict__ qkey, uint32_t* __restrict__ qval,
    double* __restrict__ r0, const int* __restrict__ escale) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 g = gx[i], x = xl[i], u = dp[i];
  const float gv[3] = {g.x, g.y, g.z}, xv[3] = {x.x, x.y, x.z}, uv[3] = {u.x, u.y, u.z};
  const float lo[3] = {c.g.lox, c.g.loy, c.g.loz}, hi[3] = {c.g.hix, c.g.hiy, c.g.hiz};
  float e[3], a[3];
  double rr = 0.0;
  for (int q = 0; q < 3; ++q) {
    const float y = xv[q] + fminf(fmaxf(uv[q], -c.max_dp), c.max_dp);
    e[q] = (y > lo[q] && y < hi[q]) ? gv[q] : 0.0f;
    const float db = (uv[q] > -c.max_dp && uv[q] < c.max_dp) ? e[q] : 0.0f;
    rr += (double)db * (-(double)uv[q] / (double)c.pho0);
    a[q] = db / c.pho0;
  }
  const uint32_t q = note_query_cell(i, x, c, qc, qkey, qval);
  gxn[i] = make_float4(e[0], e[1], e[2], 0.0f);
  A[i] = make_float4(a[0], a[1], a[2], __uint_as_float(q));
  r0[i] += escale ? ldexp(rr, *escale) : rr;
}

// Query cells of the final positions (for the XSPH pass).
__global__ __launch_bounds__(kBlock) void k_bwd_qcell(int n, StepConsts c,
                                                      const float4* __restrict__ x,
                                                      uint32_t* __restrict__ qc,
                                                      uint32_t* __restrict__ qkey,
                                                      uint32_t* __restrict__ qval) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  note_query_cell(i, x[i], c, qc, qkey, qval);
}

// Gather mode, once per frame: packed sorted (stale-grid) cell of each
// sorted slot.
__global__ __launch_bounds__(kBlock) void k_bwd_sorted_cells(int n, StepConsts c,
                                                             const uint32_t* __restrict__ key,
                                                             uint32_t* __restrict__ pc) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  pc[i] = pack_cell(cell_of_key(key[i], c.g));
}

// Gather mode: the inputs of an iteration's delta-p pass in query order
// (skey: sorted query keys, perm: their sorted slots), with the per-cell
// counts of the query grid (cnt all zero beforehand).
__global__ __launch_bounds__(kBlock) void k_bwd_gather_iter(
    int n, const uint32_t* __restrict__ skey, const uint32_t* __restrict__ perm,
    const float4* __restrict__ xl, const float4* __restrict__ A, const uint32_t* __restrict__ pc,
    float4* __restrict__ xq, float4* __restrict__ aq, uint2* __restrict__ cq,
    uint32_t* __restrict__ cnt) {
  const int u = blockIdx.x * kBlock + threadIdx.x;
  if (blockIdx.x * kBlock >= n) return;  // whole workgroup past the end
  count_sorted(skey, u, n, cnt);
  if (u >= n) return;
  const uint32_t s = perm[u] < (uint32_t)n ? perm[u] : 0u;
  const float4 a = A[s];
  xq[u] = xl[s];
  aq[u] = a;
  cq[u] = make_uint2(pc[s], __float_as_uint(a.w));
}

// Gather mode: the inputs of the XSPH pass in query order, with the per-cell
// counts of the query grid (cnt all zero beforehand).
__global__ __launch_bounds__(kBlock) void k_bwd_gather_xsph(
    int n, StepConsts c, const uint32_t* __restrict__ skey, const uint32_t* __restrict__ perm,
    const float4* __restrict__ x, const float4* __restrict__ vr, const float4* __restrict__ gv,
    const uint32_t* __restrict__ pc, float4* __restrict__ xq, float4* __restrict__ vrq,
    float4* __restrict__ gvq, uint2* __restrict__ cq, uint32_t* __restrict__ cnt) {
  const int u = blockIdx.x * kBlock + threadIdx.x;
  if (blockIdx.x * kBlock >= n) return;  // whole workgroup past the end
  count_sorted(skey, u, n, cnt);
  if (u >= n) return;
  const uint32_t s = perm[u] < (uint32_t)n ? perm[u] : 0u;
  const float4 p = x[s];
  const Cell q = cell_of(p.x, p.y, p.z, c.g, c.h);
  xq[u] = p;
  vrq[u] = vr[s];
  gvq[u] = gv[s];
  cq[u] = make_uint2(pc[s], pack_cell(q));
}

// DensityBoundary slope (fluids/Simulator.cu:139-161) at wall distance d.
__device__ __forceinline__ float wall_slope(float d, const StepConsts& c) {
  if (d > c.h || d <= 0.0f) return 0.0f;
  return -c.wall_c * (c.h - d) * (c.h + 3.0f * d);
}

// lambda = -(rho/pho0 - 1)/D: from dL/dlambda (gathered, plus in scatter
// mode the scattered extras, which it clears) to dL/drho, dL/dD; folds the
// delta-p pair terms into gxn; writes the lambda pass's per-particle inputs
// (xr, sdb: in the pass's order).
template <bool kGather>
__global__ __launch_bounds__(kBlock) void k_bwd_lambda_pre(
    int n, StepConsts c, int last, const uint32_t* __restrict__ perm,
    const float4* __restrict__ out, float4* __restrict__ ex, const float4* __restrict__ xl,
    const float4* __restrict__ sd, const float* __restrict__ rho,
    const float* __restrict__ grho_x, float4* __restrict__ gxn, float4* __restrict__ xr,
    float4* __restrict__ sdb, double* __restrict__ r0, const int* __restrict__ escale) {
  const int u = blockIdx.x * kBlock + threadIdx.x;
  if (u >= n) return;
  const int i = kGather ? (int)perm[u] : u;
  float4 e = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if constexpr (!kGather) e = take_scattered(ex, u, n);
  const float4 o = out[u], x = xl[i], s = sd[i];
  const float gl = o.w + e.w;
  const float D = s.w, rh = rho[i];
  const float C = rh / c.pho0 - 1.0f;
  const float gC = -gl / D;
  const float gD = gl * C / (D * D);
  float gr = gC / c.pho0;
  if (last) gr += grho_x[i];
  const double dr0 = (double)gC * (-(double)rh / ((double)c.pho0 * (double)c.pho0)) +
                     (double)gD * (-2.0 * ((double)D - (double)c.lambda_eps) / (double)c.pho0);
  r0[i] += escale ? ldexp(dr0, *escale) : dr0;
  float4 g = gxn[i];
  g.x += o.x + e.x;
  g.y += o.y + e.y;
  g.z += o.z + e.z;
  if (c.k_boundary != 0.0f) {
    const float k = gr * c.k_boundary;
    g.x += k * (wall_slope(x.x - c.g.llx, c) - wall_slope(c.g.ulx - x.x, c));
    g.y += k * (wall_slope(x.y - c.g.lly, c) - wall_slope(c.g.uly - x.y, c));
    g.z += k * (wall_slope(x.z - c.g.llz, c) - wall_slope(c.g.ulz - x.z, c));
  }
  gxn[i] = g;
  xr[u] = make_float4(x.x, x.y, x.z, gr);
  sdb[u] = make_float4(s.x, s.y, s.z, gD);
}

// Scatter mode: gx = gxn + scattered extras; clears the extras.
__global__ __launch_bounds__(kBlock) void k_bwd_fold(int n, const float4* __restrict__ gxn,
                                                     float4* __restrict__ ex,
                                                     float4* __restrict__ gx) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 g = gxn[i], e = take_scattered(ex, i, n);
  gx[i] = make_float4(g.x + e.x, g.y + e.y, g.z + e.z, 0.0f);
}

// Per-frame rescaling of the adjoint state (tapes with PBF_TAPE_RESCALE):
// the largest finite |value| of two arrays, as float bits (all values are
// non-negative, so the unsigned order of the bits is the float order) ...
__global__ __launch_bounds__(kBlock) void k_absmax(const float* __restrict__ a, int na,
                                                   const float* __restrict__ b, int nb,
                                                   uint32_t* __restrict__ out) {
  uint32_t m = 0u;
  const int stride = gridDim.x * kBlock;
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < na + nb; k += stride) {
    const float v = fabsf(k < na ? a[k] : b[k - na]);
    const uint32_t bits = __float_as_uint(v);
    m = (v <= 3.4028235e38f && bits > m) ? bits : m;  // NaN and inf are left out
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = __shfl_xor(m, off);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0 && m) atomicMax(out, m);
}

// ... and both arrays multiplied by 2^e (exact) so that the largest becomes
// [1, 2); escale -= e (true values = stored x 2^escale).  Every workgroup
// derives e from the same maximum; workgroup 0 updates escale.
__global__ __launch_bounds__(kBlock) void k_rescale(float* __restrict__ a, int na,
                                                    float* __restrict__ b, int nb,
                                                    const uint32_t* __restrict__ maxbits,
                                                    int* __restrict__ escale) {
  const float mx = __uint_as_float(*maxbits);
  const int e = mx > 0.0f ? -ilogbf(mx) : 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) *escale -= e;
  if (e == 0) return;
  const int stride = gridDim.x * kBlock;
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < na + nb; k += stride) {
    float* p = k < na ? a + k : b + (k - na);
    *p = ldexpf(*p, e);
  }
}

// Prediction x[0] = x + dt (v + dt g) (fluids/Simulator_kernel.cuh:14-15),
// back to the frame's input order through the sort permutation.
template <int L>
__global__ __launch_bounds__(kBlock) void k_bwd_predict(int n, StepConsts c,
                                                        const uint32_t* __restrict__ sval,
                                                        const float4* __restrict__ gx,
                                                        const float4* __restrict__ gx0,
                                                        float* __restrict__ gpos,
                                                        float* __restrict__ gvel) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  const uint32_t src = sval[s] < (uint32_t)n ? sval[s] : 0u;
  const float4 g = gx[s], g0 = gx0[s];
  store_vec<L>(gpos, src, g0.x + g.x, g0.y + g.y, g0.z + g.z);
  store_vec<L>(gvel, src, c.dt * g.x, c.dt * g.y, c.dt * g.z);
}

__global__ void k_fill_u32_one(uint32_t* __restrict__ a) {
  if (threadIdx.x == 0) *a = 0u;
}

__global__ __launch_bounds__(kBlock) void k_fill_f4(float4* __restrict__ a, int n) {
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
    a[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ __launch_bounds__(kBlock) void k_fill_f64(double* __restrict__ a, int n) {
  const int stride = gridDim.x * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) a[i] = 0.0;
}

// Layout conversion of a caller vector array to float4.
template <int L>
__global__ __launch_bounds__(kBlock) void k_to_f4(const float* __restrict__ a, int n,
                                                  float4* __restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  out[i] = load_vec<L>(a, (uint32_t)i);
}

// Sum of n doubles in a fixed order (independent of the hardware's
// scheduling): kSumParts contiguous segments, each summed by one workgroup
// (strided per work-item, then a fixed tree), then the segment sums by one
// workgroup.  part holds kSumParts doubles.
constexpr int kSumParts = 1024;

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[kBlock];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  return sh[0];
}

__global__ __launch_bounds__(kBlock) void k_sum_parts(const double* __restrict__ a, int n,
                                                      double* __restrict__ part) {
  const long long seg = ((long long)n + kSumParts - 1) / kSumParts;
  const long long lo = (long long)blockIdx.x * seg;
  const long long hi = lo + seg < n ? lo + seg : n;
  double v = 0.0;
  for (long long i = lo + threadIdx.x; i < hi; i += kBlock) v += a[i];
  const double t = block_sum(v);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(kBlock) void k_sum_total(const double* __restrict__ part,
                                                      double* __restrict__ out) {
  double v = 0.0;
  for (int i = threadIdx.x; i < kSumParts; i += kBlock) v += part[i];
  const double t = block_sum(v);
  if (threadIdx.x == 0) out[0] = t;
}

int fill_blocks(int n) {
  const int b = (n + kBlock - 1) / kBlock;
  return b < 1 ? 1 : (b > 4096 ? 4096 : b);
}

}  // namespace

void launch_bwd_frame(const BwdFrame& f, int n, const StepConsts& c, hipStream_t st) {
  if (f.q) k_bwd_sorted_cells<<<blocks_for(n), kBlock, 0, st>>>(n, c, f.key, f.q->pc);
}

namespace {
// Gather mode: sorts the sorted slots by the query keys in q.k0 (slots in
// q.v0) and returns (sorted keys, their sorted slots).
std::pair<const uint32_t*, const uint32_t*> sort_query(const BwdQuery& q, int n,
                                                       hipStream_t st) {
  const int which =
      radix_sort_pairs(q.k0, q.v0, q.k1, q.v1, n, q.key_bits, q.hist, q.totals, st);
  return which ? std::make_pair(q.k1, q.v1) : std::make_pair(q.k0, q.v0);
}
}  // namespace

void launch_bwd_xsph(const BwdFrame& f, int n, const StepConsts& c, const uint32_t* cs,
                     const HeavyLists& hl, int pass, const Shapes& sh, hipStream_t st) {
  if (f.q) {
    const BwdQuery& q = *f.q;
    k_bwd_qcell<<<blocks_for(n), kBlock, 0, st>>>(n, c, f.x_final, nullptr, q.k0, q.v0);
    const auto kv = sort_query(q, n, st);
    k_bwd_gather_xsph<<<blocks_for(n), kBlock, 0, st>>>(n, c, kv.first, kv.second, f.x_final,
                                                        f.vr, f.gv_out, q.pc, q.xq, q.vrq,
                                                        q.gvq, q.cq, q.cnt);
    launch_cell_starts(q.cnt, c.cells, q.tile_sum, q.cs, st);
    using P = BwdXsphPass<true>;
    launch_pass<P>(P::Args{q.xq, q.vrq, q.gvq, nullptr, f.out_x, f.out_v, nullptr, nullptr,
                           (uint32_t)n, nullptr, q.cq},
                   n, c, Scenes{}, q.cs, hl, pass, sh, st);
    k_bwd_velocity<true><<<blocks_for(n), kBlock, 0, st>>>(n, c, kv.second, f.gx_out, f.gv_out,
                                                           f.out_x, f.out_v, nullptr, nullptr,
                                                           f.gx, f.gx0, f.grho_x);
    return;
  }
  k_bwd_qcell<<<blocks_for(n), kBlock, 0, st>>>(n, c, f.x_final, f.qc, nullptr, nullptr);
  using P = BwdXsphPass<false>;
  launch_pass<P>(P::Args{f.x_final, f.vr, f.gv_out, f.key, f.out_x, f.out_v, f.ex_x, f.ex_v,
                         (uint32_t)n, f.qc, nullptr},
                 n, c, Scenes{}, cs, hl, pass, sh, st);
  k_bwd_velocity<false><<<blocks_for(n), kBlock, 0, st>>>(n, c, nullptr, f.gx_out, f.gv_out,
                                                          f.out_x, f.out_v, f.ex_x, f.ex_v,
                                                          f.gx, f.gx0, f.grho_x);
}

void launch_bwd_iteration(const BwdFrame& f, const BwdIter& it, int last, int n,
                          const StepConsts& c, const uint32_t* cs, const HeavyLists& hl, int pass,
                          const Shapes& sh, hipStream_t st) {
  if (f.q) {
    const BwdQuery& q = *f.q;
    k_bwd_clamp<<<blocks_for(n), kBlock, 0, st>>>(n, c, f.gx, it.xl, it.dp, f.gxn, f.A, nullptr,
                                                  q.k0, q.v0, f.r0, f.escale);
    const auto kv = sort_query(q, n, st);
    k_bwd_gather_iter<<<blocks_for(n), kBlock, 0, st>>>(n, kv.first, kv.second, it.xl, f.A, q.pc,
                                                        q.xq, q.aq, q.cq, q.cnt);
    launch_cell_starts(q.cnt, c.cells, q.tile_sum, q.cs, st);
    using PD = BwdDeltaPass<true>;
    launch_pass<PD>(PD::Args{q.xq, q.aq, nullptr, f.out_x, nullptr, (uint32_t)n, q.cq}, n, c,
                    Scenes{}, q.cs, hl, pass, sh, st);
    k_bwd_lambda_pre<true><<<blocks_for(n), kBlock, 0, st>>>(n, c, last, kv.second, f.out_x,
                                                             nullptr, it.xl, it.sd, it.rho,
                                                             f.grho_x, f.gxn, q.xr, q.sdb, f.r0,
                                                             f.escale);
    using PL = BwdLambdaPass<true>;
    launch_pass<PL>(PL::Args{q.xr, q.sdb, nullptr, f.gxn, f.gx, nullptr, (uint32_t)n, nullptr,
                             q.cq, kv.second},
                    n, c, Scenes{}, q.cs, hl, pass + 1, sh, st);
    return;
  }
  k_bwd_clamp<<<blocks_for(n), kBlock, 0, st>>>(n, c, f.gx, it.xl, it.dp, f.gxn, f.A, f.qc,
                                                nullptr, nullptr, f.r0, f.escale);
  using PD = BwdDeltaPass<false>;
  launch_pass<PD>(PD::Args{it.xl, f.A, f.key, f.out_x, f.ex_x, (uint32_t)n, nullptr}, n, c,
                  Scenes{}, cs, hl, pass, sh, st);
  k_bwd_lambda_pre<false><<<blocks_for(n), kBlock, 0, st>>>(n, c, last, nullptr, f.out_x, f.ex_x,
                                                            it.xl, it.sd, it.rho, f.grho_x, f.gxn,
                                                            f.xr, f.sdb, f.r0, f.escale);
  using PL = BwdLambdaPass<false>;
  launch_pass<PL>(PL::Args{f.xr, f.sdb, f.key, f.gxn, f.gxn, f.ex_x, (uint32_t)n, f.qc, nullptr,
                           nullptr},
                  n, c, Scenes{}, cs, hl, pass + 1, sh, st);
  k_bwd_fold<<<blocks_for(n), kBlock, 0, st>>>(n, f.gxn, f.ex_x, f.gx);
}

void launch_bwd_predict(int layout, const BwdFrame& f, const uint32_t* sval, int n,
                        const StepConsts& c, float* gpos, float* gvel, hipStream_t st) {
  if (layout == 3)
    k_bwd_predict<3><<<blocks_for(n), kBlock, 0, st>>>(n, c, sval, f.gx, f.gx0, gpos, gvel);
  else
    k_bwd_predict<4><<<blocks_for(n), kBlock, 0, st>>>(n, c, sval, f.gx, f.gx0, gpos, gvel);
}

void launch_rescale(float* a, int na, float* b, int nb, uint32_t* maxbits, int* escale,
                    hipStream_t st) {
  k_fill_u32_one<<<1, 64, 0, st>>>(maxbits);
  const int blocks = fill_blocks(na + nb);
  k_absmax<<<blocks, kBlock, 0, st>>>(a, na, b, nb, maxbits);
  k_rescale<<<blocks, kBlock, 0, st>>>(a, na, b, nb, maxbits, escale);
}

void launch_fill_f4(float4* a, int n, hipStream_t st) {
  k_fill_f4<<<fill_blocks(n), kBlock, 0, st>>>(a, n);
}

void launch_fill_f64(double* a, int n, hipStream_t st) {
  k_fill_f64<<<fill_blocks(n), kBlock, 0, st>>>(a, n);
}

void launch_to_f4(int layout, const float* a, int n, float4* out, hipStream_t st) {
  if (layout == 3)
    k_to_f4<3><<<blocks_for(n), kBlock, 0, st>>>(a, n, out);
  else
    k_to_f4<4><<<blocks_for(n), kBlock, 0, st>>>(a, n, out);
}

void launch_sum_f64(const double* a, int n, double* part, double* out, hipStream_t st) {
  k_sum_parts<<<kSumParts, kBlock, 0, st>>>(a, n, part);
  k_sum_total<<<1, kBlock, 0, st>>>(part, out);
}

}  // namespace pbf
