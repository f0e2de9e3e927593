// pbf_adjoint.h — host-side launchers of the backward (adjoint) kernels of
// one PBF frame (adjoint.hip).  All arrays are in the frame's sorted order
// unless noted; float4 arrays use w as stated.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbf_device.h"
#include "pbf_launch.h"

namespace pbf {

// Gather mode (see adjoint.hip): each pass runs over a grid built on its own
// positions.  Device arrays of n entries unless noted.
struct BwdQuery {
  uint32_t *k0, *v0, *k1, *v1;    // radix sort buffers: query keys, sorted slots
  uint32_t *hist, *totals;        // radix sort scratch (radix_sort_pairs)
  int key_bits;                   // bits of the query keys
  uint32_t *cnt, *tile_sum, *cs;  // per-cell counts (all zero between uses), scan tiles,
                                  // cell starts (max cells + 1)
  uint32_t* pc;                   // packed sorted cell of each sorted slot (per frame)
  float4 *xq, *aq, *xr, *sdb;     // an iteration's pass inputs in query order
  float4 *vrq, *gvq;              // the XSPH pass's inputs in query order (x in xq)
  uint2* cq;                      // packed (sorted cell, query cell) in query order
};

// Scatter mode's copies of the scattered pair sums: workgroup b adds to copy b % 8 (the XCD
// it runs on, as workgroups are dispatched round-robin), which spreads the
// atomic adds a crowded cell sends to one particle over eight addresses
// (same-address atomics serialise); the next pointwise kernel adds the
// copies.
constexpr int kScatterCopies = 8;

// One frame's backward state.
struct BwdFrame {
  // recorded forward (read)
  const float4* x_final;  // positions after the last iteration
  const float4* vr;       // (velocity before XSPH, rho of the last lambda pass)
  const uint32_t* key;    // sorted cell keys
  // adjoints of the frame's outputs (read)
  const float4* gx_out;   // dL/d(final positions)
  const float4* gv_out;   // dL/d(velocities after XSPH)
  // working arrays
  float4 *out_x, *out_v;  // gathered pair sums
  float4 *ex_x, *ex_v;    // scatter mode: scattered pair sums, kScatterCopies copies of n
                          // entries, component-major (all zero between kernels)
  float4* gx;             // dL/d(positions) of the current iteration
  float4* gxn;            // the same for the iteration before
  float4* gx0;            // dL/d(start positions) from v = (x - x0)/dt
  float4* A;              // (dL/d(correction sum), packed query cell as float bits)
  uint32_t* qc;           // packed query cells of the positions of the current pass
  float4 *xr, *sdb;       // lambda-pass inputs: (x, dL/drho), (S, dL/dD)
  float* grho_x;          // dL/drho of the last lambda pass, from XSPH
  double* r0;             // per-slot dL/dpho0 accumulators
  const BwdQuery* q;      // gather mode's arrays; null: scatter mode
  const int* escale;      // rescaled adjoints: true = stored x 2^*escale (r0 gets true
                          // values); null: no rescaling
};

// One solver iteration's record.
struct BwdIter {
  const float4* xl;   // (positions entering the iteration, lambda)
  const float4* sd;   // (sum of lambda gradients, denominator)
  const float* rho;   // densities
  const float4* dp;   // correction / pho0 before the clamps
};

// Once per frame before its passes (gather mode: the packed sorted cells).
void launch_bwd_frame(const BwdFrame& f, int n, const StepConsts& c, hipStream_t st);
// XSPH and velocity update: fills gx (dL/d final positions), gx0, grho_x.
// Uses the neighbour passes `pass` (one pass number).
void launch_bwd_xsph(const BwdFrame& f, int n, const StepConsts& c, const uint32_t* cs,
                     const HeavyLists& hl, int pass, const Shapes& sh, hipStream_t st);
// Iteration k, last = (k == niter - 1): gx (dL/dx after the iteration) ->
// gx (dL/dx before it).  Uses the neighbour passes `pass` and `pass + 1`.
void launch_bwd_iteration(const BwdFrame& f, const BwdIter& it, int last, int n,
                          const StepConsts& c, const uint32_t* cs, const HeavyLists& hl, int pass,
                          const Shapes& sh, hipStream_t st);
// Prediction: dL/d(start positions, velocities) in the frame's input order
// (sval: sorted slot -> input index), caller layout.
void launch_bwd_predict(int layout, const BwdFrame& f, const uint32_t* sval, int n,
                        const StepConsts& c, float* gpos, float* gvel, hipStream_t st);

// Rescales the adjoint arrays a (na floats) and b (nb floats) by a power of
// two so that their largest finite |value| lies in [1, 2), and subtracts the
// exponent applied from *escale (maxbits: one device word of scratch).
void launch_rescale(float* a, int na, float* b, int nb, uint32_t* maxbits, int* escale,
                    hipStream_t st);
void launch_fill_f4(float4* a, int n, hipStream_t st);
void launch_fill_f64(double* a, int n, hipStream_t st);
// Caller-layout vector array -> float4.
void launch_to_f4(int layout, const float* a, int n, float4* out, hipStream_t st);
// out[0] = sum of a[0..n) in a fixed order (one workgroup).
// Fixed-order sum of n doubles into out[0]; part holds 1024 doubles.
void launch_sum_f64(const double* a, int n, double* part, double* out, hipStream_t st);

}  // namespace pbf
