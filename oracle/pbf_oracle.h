/*
 * pbf_oracle.h — TEST INFRASTRUCTURE, not product code.
 *
 * Host C++ FP64 restatement of the reference PBF step (naeioi/PBF-CUDA,
 * fluids/Simulator.cpp:10-99, fluids/Simulator.cu:22-274,
 * fluids/Simulator_kernel.cuh:7-239).  It is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library never links or calls it.
 *
 * Arithmetic: the parts that produce integers (predicted positions that feed
 * the cell keys, grid dims, cell keys, query cells) are computed in FP32
 * exactly as the GPU computes them, so orderings compare bit for bit.
 * orc_step computes everything else (densities, lambdas, corrections,
 * velocities, XSPH) in FP64; orc_step_f32 is the FP32 mirror of the GPU's
 * exact operation sequence.
 */
#ifndef PBF_ORACLE_H_
#define PBF_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order and meaning as pbf_params (GUIParams.h:6-17). */
typedef struct orc_params {
  int niter;
  float pho0, g, h, dt, lambda_eps, delta_q, k_corr, n_corr, k_boundaryDensity, c_XSPH;
} orc_params;

typedef struct orc_box {
  float ulim[3];
  float llim[3];
} orc_box;

/* Options of orc_step. */
typedef struct orc_options {
  int stride;         /* floats per particle in vector arrays: 3 or 4 */
  int brute_force;    /* 1: every pair j != i is a candidate (O(N^2)); only
                         equal to the grid path while the grid is fresh */
  int threads;        /* OpenMP threads, 0 = runtime default */
  int query_cell_f64; /* 1: query cells from the FP64 position; 0 (default):
                         from the position rounded to FP32, as the GPU */
} orc_options;

/* Diagnostics of the last orc_step (sorted order, FP64).  Any pointer may be
 * NULL.  Arrays are n long (x, v: 3n). */
typedef struct orc_diag {
  double* x_final;   /* final positions (3n) */
  double* v_final;   /* final velocities after XSPH (3n) */
  double* v_pre;     /* velocities before XSPH (3n) */
  double* lambda;    /* lambda of the last iteration */
  double* rho;       /* density of the last iteration */
  double* dp_raw_first; /* first iteration's sum over j of (lam_i+lam_j+s_corr)gradW, before /pho0 and clamps (3n) */
  uint32_t* perm;    /* sorted slot -> input index */
  uint32_t* keys;    /* sorted cell keys */
} orc_diag;

/* Grid dims (int)ceilf((ulim-llim)/h) in FP32, fluids/Simulator.cu:187-188.
 * Returns 0, or -1 for an invalid box. */
int orc_grid_dims(const orc_box* box, float h, int dims[3]);

/* Cell key of FP32 positions, fluids/Simulator.cu:56-73. */
void orc_cell_keys(const float* x, int stride, int n, const orc_box* box, float h,
                   const int dims[3], uint32_t* keys);

/* Stable ascending order of keys (std::stable_sort): perm[slot] = index. */
void orc_stable_order(const uint32_t* keys, int n, uint32_t* perm);

/* [start,end) per cell over sorted keys; empty cells 0/0
 * (fluids/Simulator.cu:201-204, fluids/Simulator_kernel.cuh:21-50). */
void orc_cell_ranges(const uint32_t* sorted_keys, int n, long long cells, uint32_t* start,
                     uint32_t* end);

/* SPH kernels of fluids/Simulator.cu:75-107 and wall density :139-161, FP64.
 * grad receives 3 doubles. */
double orc_poly6(double h, double r2);
void orc_spiky_grad(double h, const double r[3], double grad[3]);
double orc_boundary_density(const orc_box* box, double h, const double p[3]);

/* One frame of Simulator::step on caller arrays, same contract as pbf_step:
 * reads pos, vel, iid; writes npos, nvel (FP32-rounded oracle results) and
 * leaves pos, vel, iid in sorted order (start positions, pre-XSPH
 * velocities, identities).  Returns 0 or -1 (invalid input). */
int orc_step(const orc_params* p, const orc_box* box, int n, float* pos, float* npos,
             float* vel, float* nvel, uint32_t* iid, const orc_options* opt, orc_diag* diag);

/* The same frame with the FP32 mirror: the step's fixed FP32 operation
 * sequence (the one the GPU kernels execute), so a correct GPU result equals
 * this one bit for bit.  diag arrays receive the FP32 values widened. */
int orc_step_f32(const orc_params* p, const orc_box* box, int n, float* pos, float* npos,
                 float* vel, float* nvel, uint32_t* iid, const orc_options* opt, orc_diag* diag);

/* mean/max |rho_i/pho0 - 1| with rho from a fresh grid over `x` (stride
 * floats per particle), self included, plus the wall term.  FP64 rho;
 * err_out (may be NULL) receives |rho_i/pho0 - 1| per particle in input
 * order. */
int orc_density_error(const orc_params* p, const orc_box* box, int n, const float* x,
                      int stride, int threads, double out[2], double* err_out);

/* Same with FP32 rho (the GPU's operation sequence); err_out (may be NULL)
 * receives |rho_i/pho0 - 1| per particle in input order. */
int orc_density_error_f32(const orc_params* p, const orc_box* box, int n, const float* x,
                          int stride, int threads, double out[2], double* err_out);

/*
 * Rollout of `nframes` frames from (pos0, vel0, iid0) (input order, 3n
 * doubles each) with the state kept in T between frames (fp32 = 0: FP64, the
 * reference of the rollout; fp32 = 1: the FP32 mirror, the GPU's trajectory
 * operation for operation).  boxes (nframes entries, may be NULL = `box`
 * for every frame) gives each frame's box.  pho0 > 0 overrides p->pho0 (for finite
 * differences in the rest density).  Loss L = (1/n) sum_i |x_final(iid i) -
 * target(iid i)|^2 with target (3n, indexed by identity).  x_final (3n, by
 * identity), loss and the gradients may be NULL; the gradients dL/dpos0,
 * dL/dvel0 (input order) and dL/dpho0 come from the adjoint of every frame,
 * evaluated in FP64 along the rollout's own trajectory, with its discrete
 * choices (order, candidate lists, cutoffs, clamps) held fixed.
 * Returns 0 or -1.
 */
int orc_rollout(const orc_params* p, const orc_box* box, const orc_box* boxes, int n,
                const double* pos0, const double* vel0, const uint32_t* iid0, int nframes,
                double pho0, int fp32, int threads, const double* target, double* loss,
                double* x_final, double* g_pos0, double* g_vel0, double* g_rho0);

/* OpenMP threads the oracle uses by default. */
int orc_max_threads(void);

#ifdef __cplusplus
}
#endif

#endif /* PBF_ORACLE_H_ */
