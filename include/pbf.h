/*
 * pbf.h — C ABI of the MI355X (gfx950) Position Based Fluids step.
 *
 * Drop-in for the `Simulator` class of naeioi/PBF-CUDA
 * (fluids/Simulator.h:7-61, fluids/Simulator.cpp:10-136).  The reference
 * takes OpenGL buffer names and reads its parameters from a global
 * singleton; this ABI takes device pointers, an explicit HIP stream and an
 * explicit parameter block, and reports errors as status codes instead of
 * exiting the process.
 *
 * Conventions
 *  - Every vector array is either PBF_LAYOUT_FLOAT4 (16 B per particle,
 *    x,y,z,w; w is not read and is written as 0) or PBF_LAYOUT_FLOAT3
 *    (12 B packed, the layout of the reference's GL vertex buffers,
 *    fluids/FluidSystem.cpp:64-84).  The layout is fixed at pbf_create.
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *    Work is stream-ordered; no call below synchronises the device except
 *    where its comment says so.
 *  - All functions return PBF_OK (0) or a PBF_ERR_* code; the message of the
 *    last failure on a solver is available from pbf_last_error(solver), and
 *    the last failure of a call without a solver (pbf_create) from
 *    pbf_last_error(NULL).  Nothing calls exit().
 *  - A solver is used from one host thread at a time; one solver per stream.
 */
#ifndef PBF_H_
#define PBF_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBF_ABI_VERSION 4

/* status codes */
#define PBF_OK 0
#define PBF_ERR_INVALID_ARGUMENT 1 /* null pointer, n out of range, bad parameter */
#define PBF_ERR_CAPACITY 2         /* n > capacity, or the box needs more cells than allowed */
#define PBF_ERR_HIP 3              /* a HIP runtime call failed (message holds hipGetErrorString) */
#define PBF_ERR_NO_DEVICE 4        /* no HIP device visible */
#define PBF_ERR_STATE 5            /* call not valid in the solver's current state */

/* caller array layouts */
#define PBF_LAYOUT_FLOAT3 3
#define PBF_LAYOUT_FLOAT4 4

typedef struct pbf_vec3 {
  float x, y, z;
} pbf_vec3;

/*
 * Fluid parameters: the solver fields of GUIParams (fluids/GUIParams.h:6-17),
 * same names and meaning.  Reference defaults are set in
 * fluids/FluidSystem.cpp:15-25.  Gravity acts along -z
 * (fluids/Simulator.cu:174).
 */
typedef struct pbf_params {
  int niter;         /* solver iterations per step */
  float pho0;        /* rest density (mass is implicitly 1) */
  float g;           /* gravity magnitude, acting along -z */
  float h;           /* kernel radius == grid cell size */
  float dt;          /* time step */
  float lambda_eps;  /* relaxation epsilon in the lambda denominator */
  float delta_q;     /* s_corr reference distance */
  float k_corr;      /* s_corr strength */
  float n_corr;      /* s_corr exponent */
  float k_boundaryDensity; /* wall density weight (0 = off) */
  float c_XSPH;      /* XSPH viscosity coefficient */
} pbf_params;

/* Opaque solver handle.  Replaces `Simulator` (fluids/Simulator.h:7). */
typedef struct pbf_solver pbf_solver;

/*
 * Create a solver on the current HIP device.
 * Replaces Simulator::Simulator(params, ulim, llim) (fluids/Simulator.h:10-26),
 * which sized its scratch by MAX_PARTICLE_NUM (fluids/helper.h:10) and four
 * times the initial box cells; here the caller states `capacity` (maximum
 * particles) and `max_cells` (maximum grid cells over the solver's life;
 * <= 0 means 4x the cells of the initial box, as the reference).
 * Allocates device scratch (hipMalloc), so it synchronises: ~(750 B per
 * particle of capacity + 8 B per cell), most of it the work list of crowded
 * cells (8 chunk items of 48 B per particle, enough for piles of ~1/3 of a
 * million particles in one cell; beyond the list a crowded target is summed
 * by its own work-item, correct but slow) and the hit masks the two passes of
 * an iteration share (256 B per particle).
 */
int pbf_create(const pbf_params* params, pbf_vec3 ulim, pbf_vec3 llim,
               int capacity, long long max_cells, int layout,
               pbf_solver** out);

/* Free the solver and its scratch.  Replaces Simulator::~Simulator
 * (fluids/Simulator.h:27-34).  NULL is accepted. */
void pbf_destroy(pbf_solver* s);

/* Replace the parameters.  Replaces Simulator::loadParams
 * (fluids/Simulator.cpp:101-115), which read the GUIParams singleton. */
int pbf_set_params(pbf_solver* s, const pbf_params* params);

/* Read the parameters back.  Replaces Simulator::saveParams
 * (fluids/Simulator.cpp:117-130). */
int pbf_get_params(const pbf_solver* s, pbf_params* out);

/* Move the box.  Replaces Simulator::setLim (fluids/Simulator.cpp:132-136).
 * The grid dims are recomputed on the next step, as the reference does every
 * step (fluids/Simulator.cu:187-188).  Fails with PBF_ERR_CAPACITY if the
 * new box needs more than max_cells cells. */
int pbf_set_lim(pbf_solver* s, pbf_vec3 ulim, pbf_vec3 llim);

/* Read the box back. */
int pbf_get_lim(const pbf_solver* s, pbf_vec3* ulim, pbf_vec3* llim);

/*
 * One PBF frame.  Replaces Simulator::step(d_pos, d_npos, d_vel, d_nvel,
 * d_iid, nparticle) (fluids/Simulator.h:37, fluids/Simulator.cpp:10-99),
 * with device pointers instead of GL buffer names.
 *
 * In:  pos, vel (start-of-frame state), iid (particle identities).
 * Out: npos, nvel (end-of-frame state).  As in the reference, every array is
 *      left in the new (cell-sorted, stable) order: pos holds the sorted
 *      start positions, vel the velocities before XSPH, iid the sorted
 *      identities.  The caller ping-pongs (pos,vel) <-> (npos,nvel) between
 *      frames (fluids/FluidSystem.cpp:112-117).
 * Stream-ordered, does not synchronise.  0 <= n <= capacity.
 */
int pbf_step(pbf_solver* s, float* pos, float* npos, float* vel, float* nvel,
             uint32_t* iid, int n, void* stream);

/*
 * Sweeping wall of the reference's FluidSystem (fluids/FluidSystem.cpp:36-38,
 * :104-110): for frame f >= start_frame,
 *   phi = sinf(w * (f - start_frame)),
 *   ulim = ulim0 + a_ulim * phi,  llim = llim0 + a_llim * phi  (FP32);
 * before start_frame the box is (ulim0, llim0).
 */
typedef struct pbf_sweep {
  pbf_vec3 ulim0, llim0; /* resting box */
  pbf_vec3 a_ulim, a_llim; /* amplitudes */
  float w;                 /* radians per frame */
  int start_frame;
} pbf_sweep;

/* The box of frame `frame` under `sweep`.  Host only (no device call). */
int pbf_sweep_box(const pbf_sweep* sweep, int frame, pbf_vec3* ulim, pbf_vec3* llim);

/*
 * FluidSystem::stepSimulate repeated for `nframes` frames
 * (fluids/FluidSystem.cpp:99-120): per frame, when sweep != NULL the box is
 * set to pbf_sweep_box(sweep, frame) (as pbf_set_lim), then one pbf_step
 * with tic-toc buffers: frames first_frame + 2k step (a -> b), the others
 * (b -> a).  *which receives 0 if the newest state is in (pos_a, vel_a),
 * 1 if in (pos_b, vel_b).  Stream-ordered, does not synchronise.  Stops at
 * the first failing frame (its status is returned; *which then names the
 * newest completed state).
 */
int pbf_run_frames(pbf_solver* s, float* pos_a, float* vel_a, float* pos_b, float* vel_b,
                   uint32_t* iid, int n, int first_frame, int nframes, const pbf_sweep* sweep,
                   void* stream, int* which);

/*
 * Density error of a particle set: mean and max over particles of
 * |rho_i/pho0 - 1|, with rho_i recomputed on `pos` by a fresh grid (self
 * included, same kernel and wall term as the lambda pass,
 * fluids/Simulator_kernel.cuh:77-120).  No counterpart in the reference; it
 * is the multi-frame density statistic.  Writes out[0]=mean, out[1]=max into
 * host memory, so it synchronises `stream`.  Does not modify `pos`.
 */
int pbf_density_error(pbf_solver* s, const float* pos, int n, void* stream,
                      double out[2]);

/* Per-phase GPU timing with HIP events (replaces the Logger phase timers,
 * fluids/Logger.cpp:17-82, fluids/Simulator.cpp:44-79, without device-wide
 * syncs).  enable != 0 starts recording; records are resolved by
 * pbf_get_timing, which synchronises the recorded events. */
#define PBF_PHASE_PREDICT 0 /* predict + cell key */
#define PBF_PHASE_SORT 1    /* radix sort of (key, index) */
#define PBF_PHASE_GRID 2    /* cell clear + permute + cell ranges */
#define PBF_PHASE_LAMBDA 3  /* one lambda launch */
#define PBF_PHASE_DELTA 4   /* one delta-p launch */
#define PBF_PHASE_VEL 5     /* velocity update: fused into the last delta-p launch, so
                               it records nothing of its own */
#define PBF_PHASE_XSPH 6    /* XSPH viscosity */
#define PBF_PHASE_STEP 7    /* whole step */
#define PBF_NUM_PHASES 8
int pbf_set_timing(pbf_solver* s, int enable);
/* ms_total[p] = summed GPU ms of phase p, count[p] = number of intervals
 * since the last call; resets the accumulators. */
int pbf_get_timing(pbf_solver* s, double ms_total[PBF_NUM_PHASES],
                   long long count[PBF_NUM_PHASES]);

/*
 * Batched solver: `nscenes` independent scenes in one set of particle
 * arrays, stepped together by one launch sequence.  No counterpart in the
 * reference (one scene per Simulator).  Scene k owns the index range
 * [off_k, off_k + counts[k]), off_k = counts[0] + ... + counts[k-1], of every
 * array, and its own box (ulim[k], llim[k]); all scenes share `params`.
 * Each scene's result equals, bit for bit, stepping that scene alone with
 * pbf_step in a single-scene solver, and each scene's arrays are left in
 * that scene's own sorted order inside its range.  Batched boxes are fixed
 * (pbf_set_lim / pbf_get_lim / pbf_run_frames with a sweep return
 * PBF_ERR_STATE); pbf_set_params applies to every scene.  `max_cells` bounds
 * the cells of all scenes together (<= 0: 4x the initial total).
 * 1 <= nscenes <= 65535; counts[k] >= 0; total <= 2^30.  Synchronises.
 */
int pbf_create_batched(const pbf_params* params, int nscenes, const pbf_vec3* ulim,
                       const pbf_vec3* llim, const int* counts, long long max_cells, int layout,
                       pbf_solver** out);

/* Scenes of a solver: *nscenes = 1 for pbf_create; counts (optional, host,
 * nscenes entries) receives each scene's particle count (for a single-scene
 * solver: its capacity). */
int pbf_get_scenes(const pbf_solver* s, int* nscenes, int* counts);

/* One frame of every scene of a batched solver (n = the total count).
 * pbf_step on a batched solver is the same call with n == the total. */
int pbf_step_batched(pbf_solver* s, float* pos, float* npos, float* vel, float* nvel,
                     uint32_t* iid, void* stream);

/*
 * Graph mode (no counterpart in the reference).  With enable != 0, pbf_step
 * records its launch sequence into a HIP graph the first time it sees an
 * (arrays, n, stream) combination, and replays that graph on later calls
 * with the same combination: one graph launch per step instead of ~40
 * kernel launches.  Up to 4 graphs are kept (a tic-toc needs 2).  A graph
 * stays valid across parameter, box and shape changes: when pbf_set_params
 * (with new values), pbf_set_lim or pbf_set_shapes changed what the
 * sequence depends on, the next step with that combination records it
 * again and updates the instantiated graph in place, so the reference's
 * calling pattern (loadParams every frame, a sweeping setLim,
 * fluids/FluidSystem.cpp:99-120) keeps replaying two graphs.
 * pbf_set_graph drops them.  Steps on the null stream, or on a stream the
 * caller is capturing, run eagerly.  With timing enabled, graph steps record
 * PBF_PHASE_STEP only.  Results are identical to eager steps.
 */
int pbf_set_graph(pbf_solver* s, int enable);

/* Graph bookkeeping of a solver since pbf_create: graphs recorded from
 * scratch and instantiated, in-place updates after a parameter / box /
 * shape change, replays, and the graphs currently kept. */
typedef struct pbf_graph_info {
  long long recorded;
  long long instantiated;
  long long updated;
  long long replayed;
  int live;
} pbf_graph_info;
int pbf_get_graph_info(const pbf_solver* s, pbf_graph_info* out);

/*
 * Launch shapes of the neighbour passes (no counterpart in the reference):
 * light_block = workgroup size of the one-target-per-lane kernels (64, 128
 * or 256), chunk_blocks = grid cap of the crowded-cell chunk kernels
 * (64..65536).  Every shape sums in the same order, so results never depend
 * on them; only the speed does.  Defaults: 256, 2048.
 */
typedef struct pbf_shapes {
  int light_block;
  int chunk_blocks;
} pbf_shapes;
int pbf_set_shapes(pbf_solver* s, const pbf_shapes* shapes);
int pbf_get_shapes(const pbf_solver* s, pbf_shapes* out);

/*
 * Crowded targets (no counterpart in the reference): a target with more
 * candidate neighbours than the threshold has its 64-candidate chunks spread
 * over the chunk kernel instead of one lane.  0 (the default) = adaptive:
 * the smallest of 64, 80, 96, 112, 128, 160, 192 such that the previous
 * neighbour pass found at most 1 in 25 targets with more candidates than it
 * and at most 256, else 256; otherwise a fixed count >= 64.  Like the launch
 * shapes, it changes no result, only the speed.
 */
int pbf_set_heavy_threshold(pbf_solver* s, int threshold);
int pbf_get_heavy_threshold(const pbf_solver* s, int* out);

/*
 * Autotuner: times one step from the given state for each candidate shape
 * (light_block in {64, 128, 256} x chunk_blocks in {512, 1024, 2048, 4096}),
 * `reps` times each (median), on private copies of pos / vel / iid (the
 * caller's arrays are not modified), checks that every shape produced the
 * same bits, and installs the fastest.  *best and *best_ms (optional)
 * receive it and its median step time.  Allocates and synchronises.
 * n = the particle count (the total for a batched solver).
 */
int pbf_autotune(pbf_solver* s, const float* pos, const float* vel, const uint32_t* iid, int n,
                 int reps, void* stream, pbf_shapes* best, double* best_ms);

/*
 * Differentiation (no counterpart in the reference).  A tape runs `nframes`
 * frames of a single-scene solver, keeping each frame's input state (x, v,
 * iid) as a checkpoint; the backward pass walks the frames last to first,
 * recomputing each frame's solver iterations from its checkpoint and
 * running the adjoint kernels of that frame.  Gradients are of the frame
 * map with its discrete choices held fixed (sort order, grid, candidate
 * lists, kernel cutoffs, active clamps).
 *
 * pbf_tape_create: tape for n particles and nframes frames under the
 *   solver's current niter (device memory ~ (36 nframes + 52 niter + 620) B
 *   per particle).  Allocates and synchronises.
 * pbf_tape_create_ex: the same with flags.  PBF_TAPE_STORE: the forward run
 *   keeps every frame's solver iterations, sort and cell-start table, and
 *   the backward pass reads them instead of recomputing each frame (device
 *   memory ~ nframes ((52 niter + 76) B per particle + 4 B per cell of the
 *   solver's largest box) more; the gradients are the same bits).
 *   PBF_TAPE_GATHER, PBF_TAPE_SCATTER: how the backward pass collects the
 *   pair terms the stale grid makes one-sided (j a candidate of i but not i
 *   of j).  GATHER runs each backward neighbour pass over a grid built on
 *   that pass's own positions, so every particle gathers all of its terms:
 *   no atomics, and the gradients do not depend on thread timing; it needs
 *   grids of at most 1024 cells per axis (else pbf_tape_backward fails with
 *   PBF_ERR_INVALID_ARGUMENT).  SCATTER lets i add such a term to j with a
 *   float atomic (the summation order, and so the last bits, can vary from
 *   run to run).  Without either flag, tapes of 2^18 particles or more
 *   gather when their grids allow it and smaller tapes scatter (there the
 *   per-pass sort of the gather costs more than the atomics).
 *   PBF_TAPE_RESCALE: before the backward pass of each frame the adjoint
 *   state is multiplied by a power of two that brings its largest finite
 *   magnitude into [1, 2), so gradients that grow by many orders of
 *   magnitude over the frames (chaotic flows, piles against a wall) stay in
 *   FP32 range; gpos0 and gvel0 then hold the true gradients x 2^-e with e
 *   from pbf_tape_gradient_scale (exact powers of two: the same bits as an
 *   unscaled pass wherever that one stays finite), and *grho0 is the true
 *   dL/dpho0 (accumulated in FP64).
 * pbf_tape_forward: runs the frames from (pos, vel, iid) (caller layout,
 *   not modified); the final state is pbf_tape_state(t, nframes).  With a
 *   sweep, frame k runs in the box pbf_sweep_box(sweep, first_frame + k)
 *   (as pbf_run_frames; the solver keeps the last box); the backward pass
 *   differentiates each frame in its own box.
 * pbf_tape_state: device pointers of checkpoint `frame` (0 = the inputs,
 *   nframes = the final state; caller layout, sorted as the frames left it).
 * pbf_tape_read: copies checkpoint `frame` into caller arrays (any may be
 *   NULL), stream-ordered.
 * pbf_tape_backward: gpos, gvel = dL/d(final positions, final velocities)
 *   in the final state's order (either may be NULL = zero); writes gpos0,
 *   gvel0 = dL/d(initial positions, velocities) in the order of pos/vel as
 *   given to pbf_tape_forward, and *grho0 = dL/dpho0 (host; may be NULL).
 *   Synchronises.  The parameters must be those of the forward run.
 */
typedef struct pbf_tape pbf_tape;
#define PBF_TAPE_STORE 1
#define PBF_TAPE_GATHER 2
#define PBF_TAPE_SCATTER 4
#define PBF_TAPE_RESCALE 8
int pbf_tape_create(pbf_solver* s, int n, int nframes, pbf_tape** out);
int pbf_tape_create_ex(pbf_solver* s, int n, int nframes, int flags, pbf_tape** out);
void pbf_tape_destroy(pbf_tape* t);
/* After pbf_tape_backward on a PBF_TAPE_RESCALE tape: the gradients it wrote
 * are the true ones x 2^-log2_scale (0 for other tapes). */
int pbf_tape_gradient_scale(const pbf_tape* t, int* log2_scale);
int pbf_tape_forward(pbf_solver* s, pbf_tape* t, const float* pos, const float* vel,
                     const uint32_t* iid, int first_frame, const pbf_sweep* sweep,
                     void* stream);
int pbf_tape_state(const pbf_tape* t, int frame, float** pos, float** vel, uint32_t** iid);
int pbf_tape_read(const pbf_tape* t, int frame, float* pos, float* vel, uint32_t* iid,
                  void* stream);
int pbf_tape_backward(pbf_solver* s, pbf_tape* t, const float* gpos, const float* gvel,
                      float* gpos0, float* gvel0, double* grho0, void* stream);

/* Message of the last failure on `s` (or of the last pbf_create failure on
 * this thread when s is NULL).  Never NULL. */
const char* pbf_last_error(const pbf_solver* s);

/* Static name of a status code. */
const char* pbf_status_string(int code);

/*
 * Allocation-free step.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same frame as
 * pbf_step (same results, bit for bit) on float4 arrays (PBF_LAYOUT_FLOAT4)
 * without a solver: parameters and box are passed per call, and the step's
 * scratch is a fixed device array inside the library (25 MiB per GPU, held
 * until the process ends), so the call allocates nothing.  Limits:
 * n <= 32768 particles and a box of at most 65536 cells (else
 * PBF_ERR_CAPACITY).  All calls on one GPU share that GPU's scratch: calls
 * whose work overlaps on the device (different streams, or host threads
 * with work in flight on different streams) race; calls on one stream are
 * ordered.  The scratch keeps the previous call's contents; each call clears
 * what must start at zero with kernels (no memset), so it can be captured.
 * Errors are reported by pbf_last_error(NULL).
 */
int pbf_step_noalloc(const pbf_params* params, pbf_vec3 ulim, pbf_vec3 llim, float* pos,
                     float* npos, float* vel, float* nvel, uint32_t* iid, int n, void* stream);

/* PBF_ABI_VERSION of the built library. */
int pbf_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PBF_H_ */
