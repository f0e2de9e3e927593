// pbf_launch.h — host-side launchers of the gfx950 kernels (stream-ordered,
// no synchronisation).  Every launcher assumes its arguments were validated
// by the caller (api.cpp): n > 0, buffers sized for n and for the cells of c.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbf_device.h"

namespace pbf {

// Scene geometry of a launch: geo == nullptr is the single scene of
// StepConsts::g; otherwise a batch, with the device table geo and the
// per-index scene ids sid.
struct Scenes {
  const SceneGeo* geo = nullptr;
  const uint16_t* sid = nullptr;
};

// Launch shapes of the neighbour passes.  Results do not depend on them
// (every shape sums in the canonical order); only the speed does.
struct Shapes {
  int light_block = 256;   // workgroup size of the light kernels: 64, 128 or 256
  int chunk_blocks = 2048;  // grid cap of the chunk kernels (grid-stride)
  int heavy = 0;            // heavy-target threshold; 0: adaptive
};

// sort.hip
int radix_tiles(int n);
// Sorts (k0,v0) by the low key_bits bits of k0, stably; (k1,v1) is scratch.
// Returns 0 if the result is in (k0,v0), 1 if in (k1,v1).  hist holds
// 256*radix_tiles(n) words, totals 256.
int radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, int n, int key_bits,
                     uint32_t* hist, uint32_t* totals, hipStream_t st);

// step.hip.  Neighbour passes take the heavy-target work lists; `pass`
// numbers the neighbour passes of one step from 0 (predict / key_only clear
// the parity-0 counters, each light kernel clears the other parity's).
void launch_predict_key(int layout, const float* pos, const float* vel, int n, const StepConsts& c,
                        const Scenes& scn, float4* xp, uint32_t* key, uint32_t* val,
                        uint32_t* counters, hipStream_t st);
void launch_key_only(int layout, const float* pos, int n, const StepConsts& c, const Scenes& scn,
                     float4* xp, uint32_t* key, uint32_t* val, uint32_t* counters,
                     hipStream_t st);
void launch_fill_u32(uint32_t* a, long long n, uint32_t v, hipStream_t st);
// Cell-start table from the per-cell counts (which it clears): cs holds
// cells + 1 entries, tile_sum scan_tiles(cells) entries.
long long scan_tiles(long long cells);
void launch_cell_starts(uint32_t* cnt, long long cells, uint32_t* tile_sum, uint32_t* cs,
                        hipStream_t st);
// The payload gathers after the sort also count the particles per cell from
// the sorted keys skey into cnt (all zero beforehand), the input of
// launch_cell_starts.
void launch_permute(int layout, const uint32_t* skey, const uint32_t* sval, const float* pos,
                    const float4* xp, const uint32_t* iid, int n, float4* x0s, float4* xa,
                    uint32_t* iid_s, uint32_t* cnt, hipStream_t st);
void launch_permute_pos(const uint32_t* skey, const uint32_t* sval, const float4* xp, int n,
                        float4* xa, uint32_t* cnt, hipStream_t st);
// rho (may be null) receives the densities; sd and dp (null outside a tape)
// receive (sum of lambda gradients, denominator) and the correction / pho0
// before the clamps, per particle.
void launch_lambda(const float4* xa, const uint32_t* cs, int n, const StepConsts& c,
                   const Scenes& scn, float4* xl, float* rho, float4* sd, const HeavyLists& hl,
                   int pass, const Shapes& sh, hipStream_t st);
// The end-of-step outputs of updateVelocity, fused into the last delta-p
// launch: vr = (v, rho) for XSPH, the caller's pos (sorted start positions),
// vel (velocities before XSPH) and iid (sorted identities).
struct VelocityOut {
  int layout;
  const float4* x0s;
  const float* rho;
  const uint32_t* iid_s;
  float4* vr;
  float* pos;
  float* vel;
  uint32_t* iid;
};
// vo: null except on the step's last iteration.
void launch_delta(const float4* xl, const uint32_t* cs, int n, const StepConsts& c,
                  const Scenes& scn, float4* xa, float4* dp, const VelocityOut* vo,
                  const HeavyLists& hl, int pass, const Shapes& sh, hipStream_t st);
void launch_xsph(int layout, const float4* xa, const float4* vr, const uint32_t* cs, int n,
                 const StepConsts& c, const Scenes& scn, float* npos, float* nvel,
                 const HeavyLists& hl, int pass, const Shapes& sh, hipStream_t st);
// err holds n doubles, partial 2*density_partials(n), out (device) 2.
int density_partials(int n);
void launch_density_error(const float4* xa, const uint32_t* cs, int n, const StepConsts& c,
                          const Scenes& scn, double* err, double* partial, double* out,
                          const HeavyLists& hl, const Shapes& sh, hipStream_t st);

}  // namespace pbf
