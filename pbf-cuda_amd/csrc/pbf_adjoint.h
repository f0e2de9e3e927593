// pbf_adjoint.h — host-side launchers of the backward (adjoint) kernels of
// one PBF frame (adjoint.hip).  All arrays are in the frame's sorted order
// unless noted; float4 arrays use w as stated.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbf_device.h"
#include "pbf_launch.h"

namespace pbf {

// One frame's backward state.
// Copies of the scattered pair sums: workgroup b adds to copy b % 8 (the XCD
// it runs on, as workgroups are dispatched round-robin), which spreads the
// atomic adds a crowded cell sends to one particle over eight addresses
// (same-address atomics serialise); the next pointwise kernel adds the
// copies.
constexpr int kScatterCopies = 8;

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
  float4 *ex_x, *ex_v;    // scattered pair sums: kScatterCopies copies of n entries,
                          // component-major (all zero between kernels)
  float4* gx;             // dL/d(positions) of the current iteration
  float4* gxn;            // the same for the iteration before
  float4* gx0;            // dL/d(start positions) from v = (x - x0)/dt
  float4* A;              // (dL/d(correction sum), packed query cell as float bits)
  uint32_t* qc;           // packed query cells of the positions of the current pass
  float4 *xr, *sdb;       // lambda-pass inputs: (x, dL/drho), (S, dL/dD)
  float* grho_x;          // dL/drho of the last lambda pass, from XSPH
  double* r0;             // per-slot dL/dpho0 accumulators
};

// One solver iteration's record.
struct BwdIter {
  const float4* xl;   // (positions entering the iteration, lambda)
  const float4* sd;   // (sum of lambda gradients, denominator)
  const float* rho;   // densities
  const float4* dp;   // correction / pho0 before the clamps
};

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

void launch_fill_f4(float4* a, int n, hipStream_t st);
void launch_fill_f64(double* a, int n, hipStream_t st);
// Caller-layout vector array -> float4.
void launch_to_f4(int layout, const float* a, int n, float4* out, hipStream_t st);
// out[0] = sum of a[0..n) in a fixed order (one workgroup).
// Fixed-order sum of n doubles into out[0]; part holds 1024 doubles.
void launch_sum_f64(const double* a, int n, double* part, double* out, hipStream_t st);

}  // namespace pbf
