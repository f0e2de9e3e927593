// api.cpp — the C ABI (include/pbf.h): solver handle, scratch ownership,
// validation, and the stream-ordered launch sequence of one frame.
//
// Replaces the Simulator facade of naeioi/PBF-CUDA
// (fluids/Simulator.h:7-61, fluids/Simulator.cpp:10-136).  Differences that
// are deliberate: device pointers instead of GL buffer names (no per-step
// register/map/unregister, Simulator.cpp:16-36 and :84-98), an explicit
// stream with no device-wide syncs (the reference has five per step,
// Simulator.cpp:44-79), explicit parameters instead of the GUIParams
// singleton (Simulator.cpp:101-115), and status codes instead of exit()
// (checkCudaErrors).
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "pbf.h"
#include "pbf_adjoint.h"
#include "pbf_launch.h"

namespace {

// fluids/helper.h:4-9 (M_PI is the reference's literal).
constexpr double kPi = 3.14159265359;
constexpr double kLimEps = 1e-3;
constexpr float kKernelEps = 1e-4f;
constexpr float kMaxDp = 0.1f;
constexpr int kMaxIter = 1 << 16;
constexpr long long kMaxCellsLimit = 1ll << 31;
// Tapes from this many particles run the gather-mode adjoint by default.
constexpr int kGatherMinParticles = 1 << 18;
constexpr int kMaxParticles = 1 << 30;
constexpr int kMaxScenes = 65535;

thread_local std::string g_create_error;

// roctx ranges around the step and its phases, for rocprofv3 --marker-trace.
// The rocprofiler-sdk roctx library is looked up at run time; without it the
// markers are no-ops (no link-time dependency).
struct Roctx {
  using Push = int (*)(const char*);
  using Pop = int (*)();
  Push push = nullptr;
  Pop pop = nullptr;
  Roctx() {
    void* h = dlopen("librocprofiler-sdk-roctx.so.1", RTLD_NOW | RTLD_LOCAL);
    if (!h) return;
    push = reinterpret_cast<Push>(dlsym(h, "roctxRangePushA"));
    pop = reinterpret_cast<Pop>(dlsym(h, "roctxRangePop"));
    if (!push || !pop) push = nullptr;
  }
};
const Roctx& roctx() {
  static const Roctx r;
  return r;
}
void mark_push(const char* name) {
  if (roctx().push) roctx().push(name);
}
void mark_pop() {
  if (roctx().push) roctx().pop();
}
const char* const kPhaseName[PBF_NUM_PHASES] = {"pbf.predict", "pbf.sort",     "pbf.grid",
                                                "pbf.lambda",  "pbf.delta",    "pbf.velocity",
                                                "pbf.xsph",    "pbf.step"};

std::string vformat(const char* fmt, va_list ap) {
  char buf[512];
  vsnprintf(buf, sizeof(buf), fmt, ap);
  return std::string(buf);
}

struct Timer {
  bool on = false;
  struct Rec {
    int phase;
    hipEvent_t a, b;
  };
  std::vector<hipEvent_t> pool;
  std::vector<Rec> recs;
  hipEvent_t open[PBF_NUM_PHASES] = {};
  double ms[PBF_NUM_PHASES] = {};
  long long cnt[PBF_NUM_PHASES] = {};
};

}  // namespace

// Everything besides the arrays that a recorded step's launch sequence
// depends on: the step constants (parameters and box), the parameter block
// (iteration count) and the launch shapes, as the bytes of those objects
// (StepConsts has tail padding, which make_consts zero-fills; a byte array
// keeps it through copies, so memcmp compares values only).
struct GraphSig {
  unsigned char b[sizeof(pbf::StepConsts) + sizeof(pbf_params) + sizeof(pbf::Shapes)];
};

// A recorded step: the launch sequence of one (arrays, n, stream)
// combination as an instantiated HIP graph, recorded under `sig`.
struct GraphEntry {
  const void* key[6] = {};
  int n = 0;
  GraphSig sig;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  unsigned long long used = 0;
};
constexpr int kMaxGraphs = 4;

struct pbf_solver {
  int device = 0;
  pbf_params prm{};
  pbf_vec3 ulim{}, llim{};
  int capacity = 0;
  int layout = PBF_LAYOUT_FLOAT4;
  long long max_cells = 0;

  // batch of scenes (nscenes == 0: single-scene solver, box in ulim/llim)
  int nscenes = 0;
  std::vector<int> counts;
  std::vector<pbf_vec3> sc_ulim, sc_llim;
  long long batch_cells = 0;
  pbf::SceneGeo* geo_d = nullptr;  // per-scene geometry (device)
  uint16_t* sid_d = nullptr;       // scene of each index (device)

  pbf::Shapes shapes;
  bool graph_on = false;
  std::vector<GraphEntry> graphs;
  unsigned long long graph_clock = 0;
  pbf_graph_info ginfo{};  // counts of graph recordings, updates, replays

  void* arena = nullptr;
  uint32_t *key0 = nullptr, *val0 = nullptr, *key1 = nullptr, *val1 = nullptr;
  uint32_t *iid_s = nullptr, *hist = nullptr, *totals = nullptr;
  float4 *xp = nullptr, *x0s = nullptr, *xa = nullptr;
  float4* vr = nullptr;  // (velocity before XSPH, rho of the last lambda pass)
  float* rho = nullptr;
  uint32_t* counters = nullptr;  // heavy-target work list counters
  pbf::HeavyLists hl{};
  double *dens_err = nullptr, *dens_partial = nullptr, *dens_out = nullptr;
  uint32_t* cnt = nullptr;       // particles per cell, all zero between steps
  uint32_t* cs = nullptr;        // cell-start table (max_cells + 1)
  uint32_t* tile_sum = nullptr;  // scan partials
  bool cnt_dirty = false;        // a step failed between counting and the scan

  Timer timer;
  std::string err;
};

// Checkpoints and backward scratch of a differentiated run (one allocation).
struct pbf_tape {
  int device = 0;
  int n = 0, nframes = 0, niter = 0, layout = PBF_LAYOUT_FLOAT4;
  bool forward_done = false;
  void* arena = nullptr;
  std::vector<float*> cpos, cvel;    // per checkpoint 0..nframes (caller layout)
  std::vector<uint32_t*> ciid;
  float *wpos = nullptr, *wvel = nullptr, *wnpos = nullptr, *wnvel = nullptr;
  uint32_t* wiid = nullptr;
  std::vector<float4*> xl, sd, dp;   // per iteration
  std::vector<float*> rho;
  float4 *out_x = nullptr, *out_v = nullptr, *ex_x = nullptr, *ex_v = nullptr, *gx = nullptr,
         *gxn = nullptr, *gx0 = nullptr, *A = nullptr, *xr = nullptr, *sdb = nullptr,
         *gxo = nullptr, *gvo = nullptr;
  float* grho_x = nullptr;
  uint32_t* qc = nullptr;           // packed query cells
  // gather mode (PBF_TAPE_GATHER, or automatic): the query grid's sort
  // buffers and scratch, the packed sorted cells; its query-order arrays
  // live in the memory of ex_x and ex_v, which scatter mode needs all zero
  int adjoint_flag = 0;             // PBF_TAPE_GATHER, PBF_TAPE_SCATTER or 0 (automatic)
  // PBF_TAPE_RESCALE: the adjoint state is renormalised by a power of two
  // before each frame's backward; the outputs are true values x 2^log2_scale
  bool rescale = false;
  uint32_t* maxbits = nullptr;      // scratch word of the renormalisation
  int* escale = nullptr;            // device: the exponent of the current scale
  int log2_scale = 0;               // host: the exponent of the last backward's outputs
  bool ex_dirty = false;            // ex_x / ex_v hold gather-mode data
  uint32_t *qk0 = nullptr, *qv0 = nullptr, *qk1 = nullptr, *qv1 = nullptr, *qhist = nullptr,
           *qtotals = nullptr, *pc = nullptr;
  float *gp[2] = {nullptr, nullptr}, *gv[2] = {nullptr, nullptr};  // inter-frame adjoints
  double *r0 = nullptr, *r0_sum = nullptr;  // r0_sum: total, then 1024 partial sums
  pbf_params prm{};                 // parameters of the forward run
  std::vector<pbf_vec3> ul, ll;     // box of each frame of the forward run
  // PBF_TAPE_STORE: what the backward pass would otherwise recompute, per
  // frame (iteration arrays indexed frame * niter + iteration)
  bool store = false;
  std::vector<float4*> f_xl, f_sd, f_dp, f_xa, f_vr;
  std::vector<float*> f_rho;
  std::vector<uint32_t*> f_key, f_val, f_cs;
};

namespace {

int fail(pbf_solver* s, int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::string m = vformat(fmt, ap);
  va_end(ap);
  if (s)
    s->err = m;
  else
    g_create_error = m;
  return code;
}

int fail_hip(pbf_solver* s, hipError_t e, const char* what) {
  return fail(s, PBF_ERR_HIP, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

// Restores the caller's current device on scope exit.
struct DeviceGuard {
  int prev = -1;
  bool ok = true;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) ok = hipSetDevice(dev) == hipSuccess;
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

bool finitef(float v) { return std::isfinite(v); }

int validate_params(pbf_solver* s, const pbf_params* p) {
  if (!p) return fail(s, PBF_ERR_INVALID_ARGUMENT, "params is NULL");
  const float v[] = {p->pho0, p->g, p->h, p->dt, p->lambda_eps, p->delta_q,
                     p->k_corr, p->n_corr, p->k_boundaryDensity, p->c_XSPH};
  for (float f : v)
    if (!finitef(f)) return fail(s, PBF_ERR_INVALID_ARGUMENT, "non-finite parameter");
  if (p->niter < 1 || p->niter > kMaxIter)
    return fail(s, PBF_ERR_INVALID_ARGUMENT, "niter must be in [1, %d], got %d", kMaxIter,
                p->niter);
  if (!(p->h > 0.0f)) return fail(s, PBF_ERR_INVALID_ARGUMENT, "h must be > 0");
  if (!(p->dt > 0.0f)) return fail(s, PBF_ERR_INVALID_ARGUMENT, "dt must be > 0");
  if (!(p->pho0 > 0.0f)) return fail(s, PBF_ERR_INVALID_ARGUMENT, "pho0 must be > 0");
  if (!(p->lambda_eps > 0.0f)) return fail(s, PBF_ERR_INVALID_ARGUMENT, "lambda_eps must be > 0");
  if (!(p->delta_q >= 0.0f && p->delta_q < p->h))
    return fail(s, PBF_ERR_INVALID_ARGUMENT, "delta_q must be in [0, h)");
  return PBF_OK;
}

// Grid dims and clamp bounds of one box (the reference recomputes the dims
// in every buildGridHash, fluids/Simulator.cu:187-188, because the box may
// move); `base` is the box's first global cell id.
int make_geo(pbf_solver* s, const pbf_params& p, pbf_vec3 ul, pbf_vec3 ll, long long base,
             pbf::SceneGeo& g, long long& cells_out) {
  const float ulv[3] = {ul.x, ul.y, ul.z};
  const float llv[3] = {ll.x, ll.y, ll.z};
  int d[3];
  for (int a = 0; a < 3; ++a) {
    if (!finitef(ulv[a]) || !finitef(llv[a]))
      return fail(s, PBF_ERR_INVALID_ARGUMENT, "non-finite box");
    const float diff = ulv[a] - llv[a];
    const float q = std::ceil(diff / p.h);
    if (!(diff > 0.0f)) return fail(s, PBF_ERR_INVALID_ARGUMENT, "ulim must exceed llim on every axis");
    if (!(q >= 1.0f) || q > 1073741824.0f)
      return fail(s, PBF_ERR_CAPACITY, "box/h gives an unrepresentable grid");
    d[a] = (int)q;
  }
  const long long cells = (long long)d[0] * d[1] * d[2];
  if (cells > kMaxCellsLimit || base + cells > kMaxCellsLimit)
    return fail(s, PBF_ERR_CAPACITY, "grid has %lld cells (> 2^31)", base + cells);
  std::memset(&g, 0, sizeof(g));
  g.llx = ll.x;
  g.lly = ll.y;
  g.llz = ll.z;
  g.ulx = ul.x;
  g.uly = ul.y;
  g.ulz = ul.z;
  // Box clamp of computetpos (fluids/Simulator_kernel.cuh:190-192): the bound
  // ulim - LIM_EPS is formed in double there; rounding it to FP32 once gives
  // the same clamped FP32 value.
  g.lox = (float)((double)ll.x + kLimEps);
  g.loy = (float)((double)ll.y + kLimEps);
  g.loz = (float)((double)ll.z + kLimEps);
  g.hix = (float)((double)ul.x - kLimEps);
  g.hiy = (float)((double)ul.y - kLimEps);
  g.hiz = (float)((double)ul.z - kLimEps);
  g.dx = d[0];
  g.dy = d[1];
  g.dz = d[2];
  g.base = (uint32_t)base;
  cells_out = cells;
  return PBF_OK;
}

// Every per-step constant, computed on the host once per step, for the box
// (ul, ll) (single scene; a batch overrides cells and ignores g).
int make_consts(pbf_solver* s, const pbf_params& p, pbf_vec3 ul, pbf_vec3 ll,
                pbf::StepConsts& c) {
  std::memset(&c, 0, sizeof(c));
  long long cells = 0;
  const int rc = make_geo(s, p, ul, ll, 0, c.g, cells);
  if (rc) return rc;
  c.cells = (int)cells;
  c.h = p.h;
  c.h2 = p.h * p.h;
  c.h2_outer = c.h2 * 1.0001f;
  c.keps = kKernelEps;
  // a strict lower bound of the pairs whose gradient term can be non-zero:
  // sqrt is monotone and correctly rounded, so r2 < 0.99 keps^2 gives
  // sqrtf(r2) <= sqrtf(0.99 keps^2) < keps
  c.keps2_lo = (float)(0.99 * (double)kKernelEps * (double)kKernelEps);
  const double h = (double)p.h;
  // getPoly6 / getSpikyGrad coefficients (fluids/Simulator.cu:77-83, :94-98),
  // evaluated in FP64 and stored as FP32.
  const double poly6 = 315.0 / (64.0 * kPi * std::pow(h, 9));
  const double spiky = -45.0 / (kPi * std::pow(h, 6));
  c.poly6 = (float)poly6;
  c.spiky = (float)spiky;
  c.spiky_rho0 = (float)(spiky / (double)p.pho0);
  c.pho0 = p.pho0;
  c.lambda_eps = p.lambda_eps;
  // coef_corr = -k_corr / W(delta_q)^n_corr (fluids/Simulator.cu:235).
  const double dq = (double)p.delta_q;
  const double wd = h * h - dq * dq;
  const double wdq = poly6 * wd * wd * wd;
  const double coef_corr = -(double)p.k_corr / std::pow(wdq, (double)p.n_corr);
  if (!std::isfinite(coef_corr) || !std::isfinite((float)coef_corr))
    return fail(s, PBF_ERR_INVALID_ARGUMENT, "k_corr / W(delta_q)^n_corr is not finite");
  c.coef_corr = (float)coef_corr;
  c.n_corr = p.n_corr;
  c.n_corr_int = (p.n_corr >= 0.0f && p.n_corr <= 16.0f && std::floor(p.n_corr) == p.n_corr)
                     ? (int)p.n_corr
                     : -1;
  c.k_boundary = p.k_boundaryDensity;
  c.wall_c = (float)(2.0 * kPi / 3.0);
  c.max_dp = kMaxDp;
  c.dt = p.dt;
  c.inv_dt = 1.0f / p.dt;  // h_updateVelocity, fluids/Simulator.cu:129
  c.gz = -p.g;
  c.c_xsph = p.c_XSPH;
  return PBF_OK;
}

int key_bits_for(long long cells) {
  int b = 1;
  while (b < 32 && (1ll << b) < cells) ++b;
  return b;
}

size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

// ---- timing ----
hipEvent_t take_event(Timer& t) {
  if (!t.pool.empty()) {
    hipEvent_t e = t.pool.back();
    t.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void tbegin(pbf_solver* s, int phase, hipStream_t st) {
  Timer& t = s->timer;
  if (!t.on) return;
  hipEvent_t e = take_event(t);
  if (e && hipEventRecord(e, st) == hipSuccess) t.open[phase] = e;
}

void tend(pbf_solver* s, int phase, hipStream_t st) {
  Timer& t = s->timer;
  if (!t.on || !t.open[phase]) return;
  hipEvent_t e = take_event(t);
  if (e && hipEventRecord(e, st) == hipSuccess) {
    t.recs.push_back({phase, t.open[phase], e});
  } else {
    t.pool.push_back(t.open[phase]);
    if (e) t.pool.push_back(e);
  }
  t.open[phase] = nullptr;
}

int check_vec_ptr(pbf_solver* s, const void* p, const char* name) {
  if (!p) return fail(s, PBF_ERR_INVALID_ARGUMENT, "%s is NULL", name);
  const uintptr_t a = (uintptr_t)p;
  if (s->layout == PBF_LAYOUT_FLOAT4 && (a & 15u))
    return fail(s, PBF_ERR_INVALID_ARGUMENT, "%s is not 16-byte aligned (float4 layout)", name);
  if (a & 3u) return fail(s, PBF_ERR_INVALID_ARGUMENT, "%s is not 4-byte aligned", name);
  return PBF_OK;
}

// Geometry of every scene of a batch under parameters p: scene k's cell ids
// follow scene k-1's.
int make_batch_geo(pbf_solver* s, const pbf_params& p, std::vector<pbf::SceneGeo>& geo,
                   long long& total) {
  geo.assign((size_t)s->nscenes, pbf::SceneGeo{});
  total = 0;
  for (int k = 0; k < s->nscenes; ++k) {
    long long cells = 0;
    const int rc = make_geo(s, p, s->sc_ulim[k], s->sc_llim[k], total, geo[k], cells);
    if (rc) return rc;
    total += cells;
  }
  return PBF_OK;
}

void drop_graphs(pbf_solver* s) {
  for (auto& g : s->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  s->graphs.clear();
}

// What a tape keeps of one frame's forward pass (all null outside a tape):
// per iteration k the (x, lambda) array, (S, D), densities and the
// correction before its clamps; and, set by enqueue_step, where the sorted
// keys and the sort permutation ended up.
struct Record {
  float4** xl = nullptr;
  float4** sd = nullptr;
  float** rho = nullptr;
  float4** dp = nullptr;
  const uint32_t* skey = nullptr;
  const uint32_t* sval = nullptr;
  // a backward pass's recompute of the frame: the XSPH pass is left out (its
  // outputs, the next frame's state, come from the checkpoints), so the last
  // neighbour pass run is the last delta-p pass
  bool skip_xsph = false;
};

// Caller arrays of one step.
struct StepArrays {
  float *pos, *npos, *vel, *nvel;
  uint32_t* iid;
  int n;
};

pbf::Scenes scenes_of(const pbf_solver* s) {
  pbf::Scenes sc;
  if (s->nscenes) {
    sc.geo = s->geo_d;
    sc.sid = s->sid_d;
  }
  return sc;
}

// Constants of the next step, with its cell count checked against max_cells.
int step_consts(pbf_solver* s, pbf::StepConsts& c) {
  const int rc = make_consts(s, s->prm, s->ulim, s->llim, c);
  if (rc) return rc;
  if (s->nscenes) c.cells = (int)s->batch_cells;
  if (c.cells > s->max_cells)
    return fail(s, PBF_ERR_CAPACITY, "box needs %d cells > max_cells %lld", c.cells, s->max_cells);
  return PBF_OK;
}

// The launch sequence of one frame, on validated arguments.  timed: record
// the phase events (never inside a graph capture).  rec: keep the frame's
// iterations for a backward pass.
void enqueue_step(pbf_solver* s, const StepArrays& a, const pbf::StepConsts& c, hipStream_t st,
                  bool timed, Record* rec = nullptr) {
  auto tb = [&](int ph) {
    mark_push(kPhaseName[ph]);
    if (timed) tbegin(s, ph, st);
  };
  auto te = [&](int ph) {
    if (timed) tend(s, ph, st);
    mark_pop();
  };
  const int L = s->layout;
  const int n = a.n;
  const pbf::Scenes scn = scenes_of(s);
  const pbf::Shapes& sh = s->shapes;
  tb(PBF_PHASE_STEP);
  // advect() + getGridId (Simulator.cpp:50, Simulator.cu:165-193)
  tb(PBF_PHASE_PREDICT);
  pbf::launch_predict_key(L, a.pos, a.vel, n, c, scn, s->xp, s->key0, s->val0, s->counters, st);
  te(PBF_PHASE_PREDICT);
  // thrust::sort_by_key (Simulator.cu:196-198)
  tb(PBF_PHASE_SORT);
  const int which = pbf::radix_sort_pairs(s->key0, s->val0, s->key1, s->val1, n,
                                          key_bits_for(c.cells), s->hist, s->totals, st);
  te(PBF_PHASE_SORT);
  const uint32_t* skey = which ? s->key1 : s->key0;
  const uint32_t* sval = which ? s->val1 : s->val0;
  if (rec) {
    rec->skey = skey;
    rec->sval = sval;
  }
  // memset + computeGridRange (Simulator.cu:200-204): the payload gather
  // counts the particles per cell, the scan turns the counts into cell starts
  tb(PBF_PHASE_GRID);
  pbf::launch_permute(L, skey, sval, a.pos, s->xp, a.iid, n, s->x0s, s->xa, s->iid_s, s->cnt,
                      st);
  pbf::launch_cell_starts(s->cnt, c.cells, s->tile_sum, s->cs, st);
  te(PBF_PHASE_GRID);
  // niter x correctDensity (Simulator.cpp:62-66, Simulator.cu:213-249);
  // xl (x, lambda) aliases xp, which is dead after the permute (a tape keeps
  // one xl per iteration instead).
  int pass = 0;
  const int NI = s->prm.niter;
  for (int it = 0; it < NI; ++it) {
    float4* xl = rec ? rec->xl[it] : s->xp;
    float* rho = rec ? rec->rho[it] : (it + 1 == NI ? s->rho : nullptr);
    tb(PBF_PHASE_LAMBDA);
    pbf::launch_lambda(s->xa, s->cs, n, c, scn, xl, rho, rec ? rec->sd[it] : nullptr, s->hl,
                       pass++, sh, st);
    te(PBF_PHASE_LAMBDA);
    // the last iteration also does updateVelocity (Simulator.cpp:72,
    // Simulator.cu:267-274) as each particle's final position is known
    const pbf::VelocityOut vo{L,         s->x0s, rec ? rec->rho[NI - 1] : s->rho, s->iid_s,
                              s->vr,     a.pos,  a.vel,                            a.iid};
    tb(PBF_PHASE_DELTA);
    pbf::launch_delta(xl, s->cs, n, c, scn, s->xa, rec ? rec->dp[it] : nullptr,
                      it + 1 == NI ? &vo : nullptr, s->hl, pass++, sh, st);
    te(PBF_PHASE_DELTA);
  }
  // correctVelocity (Simulator.cpp:77, Simulator.cu:251-265)
  if (!(rec && rec->skip_xsph)) {
    tb(PBF_PHASE_XSPH);
    pbf::launch_xsph(L, s->xa, s->vr, s->cs, n, c, scn, a.npos, a.nvel, s->hl, pass, sh, st);
    te(PBF_PHASE_XSPH);
  }
  te(PBF_PHASE_STEP);
}

// Records the step's launch sequence on `st` (stream capture of
// enqueue_step) into *graph.
int capture_step(pbf_solver* s, const StepArrays& a, const pbf::StepConsts& c, hipStream_t st,
                 hipGraph_t* graph) {
  (void)hipGetLastError();
  hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return fail_hip(s, e, "hipStreamBeginCapture");
  enqueue_step(s, a, c, st, false);
  const hipError_t le = hipGetLastError();
  *graph = nullptr;
  e = hipStreamEndCapture(st, graph);
  if (le != hipSuccess || e != hipSuccess) {
    if (*graph) (void)hipGraphDestroy(*graph);
    *graph = nullptr;
    return fail_hip(s, le != hipSuccess ? le : e, "recording the step graph");
  }
  return PBF_OK;
}

GraphSig make_sig(const pbf_solver* s, const pbf::StepConsts& c) {
  GraphSig g;
  std::memcpy(g.b, &c, sizeof(c));
  std::memcpy(g.b + sizeof(c), &s->prm, sizeof(pbf_params));
  std::memcpy(g.b + sizeof(c) + sizeof(pbf_params), &s->shapes, sizeof(pbf::Shapes));
  return g;
}

// The recorded graph of this (arrays, n, stream) combination.  None: the
// step is recorded and instantiated (the least recently used of kMaxGraphs
// entries makes room).  One recorded under other constants, parameters or
// shapes (a moving box, a changed rest density, ...): the step is recorded
// again and the instantiated graph updated in place
// (hipGraphExecUpdate; the launch sequence has the same topology), or
// re-instantiated if the update is refused (a different sort pass count).
int find_or_record(pbf_solver* s, const StepArrays& a, const pbf::StepConsts& c, hipStream_t st,
                   GraphEntry** out) {
  const void* key[6] = {a.pos, a.npos, a.vel, a.nvel, a.iid, (const void*)st};
  const GraphSig sig = make_sig(s, c);
  for (auto& g : s->graphs) {
    if (g.n == a.n && std::memcmp(g.key, key, sizeof(key)) == 0) {
      g.used = ++s->graph_clock;
      *out = &g;
      if (std::memcmp(&g.sig, &sig, sizeof(sig)) == 0) return PBF_OK;
      hipGraph_t graph = nullptr;
      int rc = capture_step(s, a, c, st, &graph);
      if (rc) return rc;
      hipGraphNode_t bad = nullptr;
      hipGraphExecUpdateResult res = hipGraphExecUpdateError;
      hipError_t e = hipGraphExecUpdate(g.exec, graph, &bad, &res);
      if (e == hipSuccess && res == hipGraphExecUpdateSuccess) {
        ++s->ginfo.updated;
      } else {
        (void)hipGetLastError();
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
          (void)hipGraphDestroy(graph);
          return fail_hip(s, e, "hipGraphInstantiate");
        }
        (void)hipGraphExecDestroy(g.exec);
        g.exec = exec;
        ++s->ginfo.instantiated;
      }
      (void)hipGraphDestroy(g.graph);
      g.graph = graph;
      g.sig = sig;
      return PBF_OK;
    }
  }
  if ((int)s->graphs.size() >= kMaxGraphs) {
    size_t lru = 0;
    for (size_t k = 1; k < s->graphs.size(); ++k)
      if (s->graphs[k].used < s->graphs[lru].used) lru = k;
    (void)hipGraphExecDestroy(s->graphs[lru].exec);
    (void)hipGraphDestroy(s->graphs[lru].graph);
    s->graphs.erase(s->graphs.begin() + (long)lru);
  }
  hipGraph_t graph = nullptr;
  int rc = capture_step(s, a, c, st, &graph);
  if (rc) return rc;
  hipGraphExec_t exec = nullptr;
  hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return fail_hip(s, e, "hipGraphInstantiate");
  }
  GraphEntry g;
  std::memcpy(g.key, key, sizeof(key));
  g.n = a.n;
  g.sig = sig;
  g.graph = graph;
  g.exec = exec;
  g.used = ++s->graph_clock;
  s->graphs.push_back(g);
  ++s->ginfo.recorded;
  ++s->ginfo.instantiated;
  *out = &s->graphs.back();
  return PBF_OK;
}

// Offsets of the solver's scratch arrays in one arena for `capacity`
// particles, `max_cells` cells and `nscenes` scenes (0: a single scene).
struct ArenaLayout {
  size_t key0, val0, key1, val1, iid, hist, tot, xp, x0s, xa, vr, rho, recs, items, part, cnt,
      de, dp, dout, cnt2, cs, ts, geo, sid, mask;
  size_t icap, total;
};

ArenaLayout arena_layout(int capacity, long long max_cells, int nscenes) {
  const size_t N = (size_t)capacity;
  const size_t tiles = (size_t)pbf::radix_tiles(capacity);
  const size_t dblk = (size_t)pbf::density_partials(capacity);
  ArenaLayout L{};
  // chunk items of crowded targets (48 B each with their partial sums): 8
  // per particle.  A pile of P particles in one cell needs ~P^2/64 items (the
  // 4M-particle sweeping-wall run piles 17,000 into a corner cell and needs
  // ~18M); targets beyond the list are summed by their own work-item, which
  // is correct but serial
  L.icap = std::min<size_t>(8 * N + 4096, 0xF0000000u);
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    const size_t o = off;
    off = align_up(off + bytes);
    return o;
  };
  L.key0 = carve(4 * N);
  L.val0 = carve(4 * N);
  L.key1 = carve(4 * N);
  L.val1 = carve(4 * N);
  L.iid = carve(4 * N);
  L.hist = carve(4 * 256 * tiles);
  L.tot = carve(4 * 256);
  L.xp = carve(16 * N);
  L.x0s = carve(16 * N);
  L.xa = carve(16 * N);
  L.vr = carve(16 * N);
  L.rho = carve(4 * N);
  L.recs = carve(16 * N);
  L.items = carve(16 * L.icap);
  L.part = carve(4 * pbf::kPartialWords * L.icap);
  L.cnt = carve(4 * pbf::kCounterWords);
  L.de = carve(8 * N);
  L.dp = carve(16 * dblk);
  L.dout = carve(16);
  L.cnt2 = carve(4 * (size_t)max_cells);
  L.cs = carve(4 * ((size_t)max_cells + 1));
  L.ts = carve(4 * ((size_t)pbf::scan_tiles(max_cells) + 1));
  L.geo = carve(sizeof(pbf::SceneGeo) * (size_t)(nscenes > 0 ? nscenes : 1));
  L.sid = carve(nscenes > 0 ? 2 * N : 2);
  L.mask = carve(8 * pbf::kMaskBlocks * N);
  L.total = off;
  return L;
}

// Points the solver's scratch arrays into the arena at `b` (laid out by
// arena_layout for `capacity` particles).
void bind_arena(pbf_solver* s, char* b, const ArenaLayout& L, int capacity) {
  s->key0 = (uint32_t*)(b + L.key0);
  s->val0 = (uint32_t*)(b + L.val0);
  s->key1 = (uint32_t*)(b + L.key1);
  s->val1 = (uint32_t*)(b + L.val1);
  s->iid_s = (uint32_t*)(b + L.iid);
  s->hist = (uint32_t*)(b + L.hist);
  s->totals = (uint32_t*)(b + L.tot);
  s->xp = (float4*)(b + L.xp);
  s->x0s = (float4*)(b + L.x0s);
  s->xa = (float4*)(b + L.xa);
  s->vr = (float4*)(b + L.vr);
  s->rho = (float*)(b + L.rho);
  s->counters = (uint32_t*)(b + L.cnt);
  s->hl.counters = s->counters;
  s->hl.recs = (uint4*)(b + L.recs);
  s->hl.items = (uint4*)(b + L.items);
  s->hl.partials = (float*)(b + L.part);
  s->hl.item_cap = (uint32_t)L.icap;
  s->hl.masks = (uint64_t*)(b + L.mask);
  s->hl.mask_stride = (uint32_t)capacity;
  s->dens_err = (double*)(b + L.de);
  s->dens_partial = (double*)(b + L.dp);
  s->dens_out = (double*)(b + L.dout);
  s->cnt = (uint32_t*)(b + L.cnt2);
  s->cs = (uint32_t*)(b + L.cs);
  s->tile_sum = (uint32_t*)(b + L.ts);
  s->geo_d = (pbf::SceneGeo*)(b + L.geo);
  s->sid_d = (uint16_t*)(b + L.sid);
}

// Allocation of a solver for `capacity` particles whose initial grid has
// `cells` cells.  The caller fills in the boxes.
int create_impl(const pbf_params& params, int capacity, long long cells, long long max_cells,
                int layout, int nscenes, pbf_solver** out) {
  if (layout != PBF_LAYOUT_FLOAT3 && layout != PBF_LAYOUT_FLOAT4)
    return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "layout must be 3 or 4, got %d", layout);
  // Simulator.h:13-14 sizes the grid at 4x the initial box's cells.
  if (max_cells <= 0) max_cells = 4ll * cells;
  if (max_cells > kMaxCellsLimit) max_cells = kMaxCellsLimit;
  if (max_cells < cells)
    return fail(nullptr, PBF_ERR_CAPACITY, "max_cells %lld < cells of the box %lld", max_cells, cells);

  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev < 1)
    return fail(nullptr, PBF_ERR_NO_DEVICE, "no HIP device: %s", hipGetErrorString(e));
  int dev = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return fail_hip(nullptr, e, "hipGetDevice");

  pbf_solver* s = new (std::nothrow) pbf_solver();
  if (!s) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "out of host memory");
  s->device = dev;
  s->prm = params;
  s->capacity = capacity;
  s->layout = layout;
  s->max_cells = max_cells;

  const ArenaLayout lay = arena_layout(capacity, max_cells, nscenes);
  void* arena = nullptr;
  if ((e = hipMalloc(&arena, lay.total)) != hipSuccess) {
    delete s;
    return fail_hip(nullptr, e, "hipMalloc of solver scratch");
  }
  s->arena = arena;
  bind_arena(s, (char*)arena, lay, capacity);
  // the per-cell counts start (and stay, between steps) all zero; the heavy
  // threshold word starts at 0 (kLightMax)
  pbf::launch_fill_u32(s->cnt, max_cells, 0u, nullptr);
  pbf::launch_fill_u32(s->counters, pbf::kCounterWords, 0u, nullptr);
  if ((e = hipDeviceSynchronize()) != hipSuccess) {
    (void)hipFree(arena);
    delete s;
    return fail_hip(nullptr, e, "clearing the cell counts");
  }
  *out = s;
  return PBF_OK;
}

// ---- the allocation-free step ----------------------------------------------
// Scratch of pbf_step_noalloc: a file-scope device array, one copy per GPU
// (each device holds its own until the process ends).  Host code never uses
// its name as an address: the device address comes from hipGetSymbolAddress,
// taken once per device with that device current.
constexpr int kNoAllocParticles = 32768;
constexpr long long kNoAllocCells = 65536;
constexpr size_t kNoAllocBytes = (size_t)25 << 20;
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// One solver record per device whose arrays point into that device's
// scratch (bound on the first call on the device; no allocation).
struct NoAllocCtx {
  std::mutex mu;
  bool ready = false;
  pbf_solver s;
};
NoAllocCtx g_noalloc[kNoAllocDevices];

int noalloc_bind(NoAllocCtx& ctx, int dev) {
  if (ctx.ready) return PBF_OK;
  const ArenaLayout lay = arena_layout(kNoAllocParticles, kNoAllocCells, 0);
  if (lay.total > kNoAllocBytes)
    return fail(nullptr, PBF_ERR_CAPACITY, "no-alloc scratch needs %zu bytes > %zu", lay.total,
                kNoAllocBytes);
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr)
    return fail_hip(nullptr, e != hipSuccess ? e : hipErrorInvalidSymbol,
                    "hipGetSymbolAddress of the no-alloc scratch");
  pbf_solver& s = ctx.s;
  s.device = dev;
  s.capacity = kNoAllocParticles;
  s.layout = PBF_LAYOUT_FLOAT4;
  s.max_cells = kNoAllocCells;
  s.nscenes = 0;
  s.arena = nullptr;  // not owned: the scratch array
  bind_arena(&s, (char*)ptr, lay, kNoAllocParticles);
  ctx.ready = true;
  return PBF_OK;
}

}  // namespace

extern "C" {


int pbf_abi_version(void) { return PBF_ABI_VERSION; }

const char* pbf_status_string(int code) {
  switch (code) {
    case PBF_OK: return "PBF_OK";
    case PBF_ERR_INVALID_ARGUMENT: return "PBF_ERR_INVALID_ARGUMENT";
    case PBF_ERR_CAPACITY: return "PBF_ERR_CAPACITY";
    case PBF_ERR_HIP: return "PBF_ERR_HIP";
    case PBF_ERR_NO_DEVICE: return "PBF_ERR_NO_DEVICE";
    case PBF_ERR_STATE: return "PBF_ERR_STATE";
    default: return "PBF_ERR_UNKNOWN";
  }
}

const char* pbf_last_error(const pbf_solver* s) {
  return s ? s->err.c_str() : g_create_error.c_str();
}

int pbf_create(const pbf_params* params, pbf_vec3 ulim, pbf_vec3 llim, int capacity,
               long long max_cells, int layout, pbf_solver** out) {
  if (!out) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  int rc = validate_params(nullptr, params);
  if (rc) return rc;
  if (capacity < 1 || capacity > kMaxParticles)
    return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "capacity must be in [1, 2^30], got %d", capacity);
  pbf::StepConsts c;
  rc = make_consts(nullptr, *params, ulim, llim, c);
  if (rc) return rc;
  pbf_solver* s = nullptr;
  rc = create_impl(*params, capacity, c.cells, max_cells, layout, 0, &s);
  if (rc) return rc;
  s->ulim = ulim;
  s->llim = llim;
  *out = s;
  return PBF_OK;
}

int pbf_create_batched(const pbf_params* params, int nscenes, const pbf_vec3* ulim,
                       const pbf_vec3* llim, const int* counts, long long max_cells, int layout,
                       pbf_solver** out) {
  if (!out) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  int rc = validate_params(nullptr, params);
  if (rc) return rc;
  if (nscenes < 1 || nscenes > kMaxScenes)
    return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "nscenes must be in [1, %d], got %d", kMaxScenes,
                nscenes);
  if (!ulim || !llim || !counts)
    return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "ulim, llim and counts must not be NULL");
  long long total = 0;
  for (int k = 0; k < nscenes; ++k) {
    if (counts[k] < 0)
      return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "counts[%d] = %d is negative", k, counts[k]);
    total += counts[k];
  }
  if (total < 1 || total > kMaxParticles)
    return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "total particle count must be in [1, 2^30], got %lld",
                total);
  pbf_solver tmp;
  tmp.nscenes = nscenes;
  tmp.sc_ulim.assign(ulim, ulim + nscenes);
  tmp.sc_llim.assign(llim, llim + nscenes);
  std::vector<pbf::SceneGeo> geo;
  long long cells = 0;
  rc = make_batch_geo(&tmp, *params, geo, cells);
  if (rc) {
    g_create_error = tmp.err;
    return rc;
  }
  pbf_solver* s = nullptr;
  rc = create_impl(*params, (int)total, cells, max_cells, layout, nscenes, &s);
  if (rc) return rc;
  s->nscenes = nscenes;
  s->counts.assign(counts, counts + nscenes);
  s->sc_ulim = tmp.sc_ulim;
  s->sc_llim = tmp.sc_llim;
  s->ulim = ulim[0];
  s->llim = llim[0];
  s->batch_cells = cells;
  // scene table and per-index scene ids: set once (counts are fixed)
  std::vector<uint16_t> sid((size_t)total);
  size_t o = 0;
  for (int k = 0; k < nscenes; ++k)
    for (int q = 0; q < counts[k]; ++q) sid[o++] = (uint16_t)k;
  hipError_t e = hipMemcpy(s->sid_d, sid.data(), sid.size() * sizeof(uint16_t), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(s->geo_d, geo.data(), geo.size() * sizeof(pbf::SceneGeo), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    g_create_error = std::string("uploading the scene table: ") + hipGetErrorString(e);
    pbf_destroy(s);
    return PBF_ERR_HIP;
  }
  *out = s;
  return PBF_OK;
}

int pbf_get_scenes(const pbf_solver* s, int* nscenes, int* counts) {
  if (!s || !nscenes) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (s->nscenes == 0) {
    *nscenes = 1;
    if (counts) counts[0] = s->capacity;
  } else {
    *nscenes = s->nscenes;
    if (counts) std::memcpy(counts, s->counts.data(), sizeof(int) * (size_t)s->nscenes);
  }
  return PBF_OK;
}

void pbf_destroy(pbf_solver* s) {
  if (!s) return;
  {
    DeviceGuard g(s->device);
    drop_graphs(s);
    for (auto& r : s->timer.recs) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t e : s->timer.pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : s->timer.open)
      if (e) (void)hipEventDestroy(e);
    if (s->arena) (void)hipFree(s->arena);
  }
  delete s;
}

int pbf_set_params(pbf_solver* s, const pbf_params* params) {
  if (!s) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "solver is NULL");
  int rc = validate_params(s, params);
  if (rc) return rc;
  // the reference reloads its parameters every frame (FluidSystem.cpp:102):
  // unchanged parameters change nothing
  if (std::memcmp(params, &s->prm, sizeof(pbf_params)) == 0) return PBF_OK;
  pbf::StepConsts c;
  rc = make_consts(s, *params, s->ulim, s->llim, c);
  if (rc) return rc;
  if (s->nscenes) {
    // the grid dims follow h: re-derive and upload the scene table
    std::vector<pbf::SceneGeo> geo;
    long long cells = 0;
    if ((rc = make_batch_geo(s, *params, geo, cells))) return rc;
    if (cells > s->max_cells)
      return fail(s, PBF_ERR_CAPACITY, "h gives %lld cells > max_cells %lld", cells, s->max_cells);
    DeviceGuard g(s->device);
    hipError_t e = hipDeviceSynchronize();  // no step in flight reads the old table
    if (e == hipSuccess)
      e = hipMemcpy(s->geo_d, geo.data(), geo.size() * sizeof(pbf::SceneGeo),
                    hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail_hip(s, e, "uploading the scene table");
    s->batch_cells = cells;
  } else if (c.cells > s->max_cells) {
    return fail(s, PBF_ERR_CAPACITY, "h gives %d cells > max_cells %lld", c.cells, s->max_cells);
  }
  s->prm = *params;
  return PBF_OK;
}

int pbf_get_params(const pbf_solver* s, pbf_params* out) {
  if (!s || !out) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = s->prm;
  return PBF_OK;
}

int pbf_set_lim(pbf_solver* s, pbf_vec3 ulim, pbf_vec3 llim) {
  if (!s) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "solver is NULL");
  if (s->nscenes) return fail(s, PBF_ERR_STATE, "the boxes of a batched solver are fixed");
  pbf::StepConsts c;
  int rc = make_consts(s, s->prm, ulim, llim, c);
  if (rc) return rc;
  if (c.cells > s->max_cells)
    return fail(s, PBF_ERR_CAPACITY, "box needs %d cells > max_cells %lld", c.cells, s->max_cells);
  s->ulim = ulim;
  s->llim = llim;
  return PBF_OK;
}

int pbf_get_lim(const pbf_solver* s, pbf_vec3* ulim, pbf_vec3* llim) {
  if (!s || !ulim || !llim) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "NULL argument");
  if (s->nscenes)
    return fail(const_cast<pbf_solver*>(s), PBF_ERR_STATE, "a batched solver has one box per scene");
  *ulim = s->ulim;
  *llim = s->llim;
  return PBF_OK;
}

int pbf_step(pbf_solver* s, float* pos, float* npos, float* vel, float* nvel, uint32_t* iid,
             int n, void* stream) {
  if (!s) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "solver is NULL");
  if (n < 0) return fail(s, PBF_ERR_INVALID_ARGUMENT, "n must be >= 0, got %d", n);
  if (n > s->capacity)
    return fail(s, PBF_ERR_CAPACITY, "n = %d exceeds capacity %d", n, s->capacity);
  if (s->nscenes && n != s->capacity)
    return fail(s, PBF_ERR_INVALID_ARGUMENT,
                "a batched solver steps all its particles: n must be %d, got %d", s->capacity, n);
  if (n == 0) return PBF_OK;
  int rc;
  if ((rc = check_vec_ptr(s, pos, "pos")) || (rc = check_vec_ptr(s, npos, "npos")) ||
      (rc = check_vec_ptr(s, vel, "vel")) || (rc = check_vec_ptr(s, nvel, "nvel")))
    return rc;
  if (!iid) return fail(s, PBF_ERR_INVALID_ARGUMENT, "iid is NULL");
  {
    const void* p[5] = {pos, npos, vel, nvel, iid};
    for (int a = 0; a < 5; ++a)
      for (int b = a + 1; b < 5; ++b)
        if (p[a] == p[b]) return fail(s, PBF_ERR_INVALID_ARGUMENT, "particle arrays must be distinct");
  }
  pbf::StepConsts c;
  if ((rc = step_consts(s, c))) return rc;
  DeviceGuard g(s->device);
  if (!g.ok) return fail(s, PBF_ERR_HIP, "hipSetDevice(%d) failed", s->device);
  hipStream_t st = (hipStream_t)stream;
  const StepArrays arr{pos, npos, vel, nvel, iid, n};

  if (s->cnt_dirty) pbf::launch_fill_u32(s->cnt, s->max_cells, 0u, st);
  s->cnt_dirty = true;
  if (s->graph_on && st != nullptr) {
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    hipError_t e = hipStreamIsCapturing(st, &cap);
    if (e != hipSuccess) return fail_hip(s, e, "hipStreamIsCapturing");
    if (cap == hipStreamCaptureStatusNone) {
      GraphEntry* ge = nullptr;
      if ((rc = find_or_record(s, arr, c, st, &ge))) return rc;
      tbegin(s, PBF_PHASE_STEP, st);
      e = hipGraphLaunch(ge->exec, st);
      tend(s, PBF_PHASE_STEP, st);
      if (e != hipSuccess) return fail_hip(s, e, "hipGraphLaunch");
      ++s->ginfo.replayed;
      s->cnt_dirty = false;
      return PBF_OK;
    }
  }
  enqueue_step(s, arr, c, st, true);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(s, e, "kernel launch");
  s->cnt_dirty = false;
  return PBF_OK;
}

int pbf_step_batched(pbf_solver* s, float* pos, float* npos, float* vel, float* nvel,
                     uint32_t* iid, void* stream) {
  if (!s) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "solver is NULL");
  if (!s->nscenes) return fail(s, PBF_ERR_STATE, "not a batched solver");
  return pbf_step(s, pos, npos, vel, nvel, iid, s->capacity, stream);
}

int pbf_set_graph(pbf_solver* s, int enable) {
  if (!s) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "solver is NULL");
  DeviceGuard g(s->device);
  drop_graphs(s);
  s->graph_on = enable != 0;
  return PBF_OK;
}

int pbf_get_graph_info(const pbf_solver* s, pbf_graph_info* out) {
  if (!s || !out) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = s->ginfo;
  out->live = (int)s->graphs.size();
  return PBF_OK;
}

int pbf_set_shapes(pbf_solver* s, const pbf_shapes* sh) {
  if (!s) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "solver is NULL");
  if (!sh) return fail(s, PBF_ERR_INVALID_ARGUMENT, "shapes is NULL");
  if (sh->light_block != 64 && sh->light_block != 128 && sh->light_block != 256)
    return fail(s, PBF_ERR_INVALID_ARGUMENT, "light_block must be 64, 128 or 256, got %d",
                sh->light_block);
  if (sh->chunk_blocks < 64 || sh->chunk_blocks > 65536)
    return fail(s, PBF_ERR_INVALID_ARGUMENT, "chunk_blocks must be in [64, 65536], got %d",
                sh->chunk_blocks);
  s->shapes.light_block = sh->light_block;
  s->shapes.chunk_blocks = sh->chunk_blocks;
  return PBF_OK;
}

int pbf_set_heavy_threshold(pbf_solver* s, int threshold) {
  if (!s) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "solver is NULL");
  if (threshold != 0 && (threshold < 64 || threshold > (1 << 30)))
    return fail(s, PBF_ERR_INVALID_ARGUMENT, "heavy threshold must be 0 or in [64, 2^30], got %d",
                threshold);
  s->shapes.heavy = threshold;
  return PBF_OK;
}

int pbf_get_heavy_threshold(const pbf_solver* s, int* out) {
  if (!s || !out) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = s->shapes.heavy;
  return PBF_OK;
}

int pbf_get_shapes(const pbf_solver* s, pbf_shapes* out) {
  if (!s || !out) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "NULL argument");
  out->light_block = s->shapes.light_block;
  out->chunk_blocks = s->shapes.chunk_blocks;
  return PBF_OK;
}

int pbf_autotune(pbf_solver* s, const float* pos, const float* vel, const uint32_t* iid, int n,
                 int reps, void* stream, pbf_shapes* best, double* best_ms) {
  if (!s) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "solver is NULL");
  if (n < 1 || n > s->capacity)
    return fail(s, PBF_ERR_INVALID_ARGUMENT, "n must be in [1, capacity = %d], got %d", s->capacity,
                n);
  if (s->nscenes && n != s->capacity)
    return fail(s, PBF_ERR_INVALID_ARGUMENT, "a batched solver needs n = %d, got %d", s->capacity,
                n);
  if (reps < 1 || reps > 100) return fail(s, PBF_ERR_INVALID_ARGUMENT, "reps must be in [1, 100]");
  int rc;
  if ((rc = check_vec_ptr(s, pos, "pos")) || (rc = check_vec_ptr(s, vel, "vel"))) return rc;
  if (!iid) return fail(s, PBF_ERR_INVALID_ARGUMENT, "iid is NULL");
  DeviceGuard g(s->device);
  if (!g.ok) return fail(s, PBF_ERR_HIP, "hipSetDevice(%d) failed", s->device);
  hipStream_t st = (hipStream_t)stream;
  const size_t vb = (size_t)n * (s->layout == PBF_LAYOUT_FLOAT4 ? 16 : 12);
  const size_t ib = (size_t)n * 4;
  const size_t vba = align_up(vb), iba = align_up(ib);
  char* work = nullptr;
  hipError_t e = hipMalloc((void**)&work, 4 * vba + iba);
  if (e != hipSuccess) return fail_hip(s, e, "hipMalloc of autotune arrays");
  float *wp = (float*)work, *wnp = (float*)(work + vba), *wv = (float*)(work + 2 * vba),
        *wnv = (float*)(work + 3 * vba);
  uint32_t* wi = (uint32_t*)(work + 4 * vba);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const pbf::Shapes saved = s->shapes;
  const bool graph_saved = s->graph_on, timer_saved = s->timer.on;
  s->graph_on = false;
  s->timer.on = false;
  std::vector<char> ref, out;
  pbf::Shapes win = saved;
  double win_ms = 1e300;
  rc = PBF_OK;
  if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
    rc = fail_hip(s, e, "hipEventCreate");
  }
  const int blocks[3] = {64, 128, 256};
  const int grids[4] = {512, 1024, 2048, 4096};
  bool first = true;
  for (int bi = 0; bi < 3 && rc == PBF_OK; ++bi) {
    for (int gi = 0; gi < 4 && rc == PBF_OK; ++gi) {
      s->shapes.light_block = blocks[bi];
      s->shapes.chunk_blocks = grids[gi];
      std::vector<double> t;
      for (int r = 0; r < reps + (first ? 1 : 0) && rc == PBF_OK; ++r) {
        if ((e = hipMemcpyAsync(wp, pos, vb, hipMemcpyDeviceToDevice, st)) != hipSuccess ||
            (e = hipMemcpyAsync(wv, vel, vb, hipMemcpyDeviceToDevice, st)) != hipSuccess ||
            (e = hipMemcpyAsync(wi, iid, ib, hipMemcpyDeviceToDevice, st)) != hipSuccess ||
            (e = hipEventRecord(e0, st)) != hipSuccess) {
          rc = fail_hip(s, e, "autotune copy");
          break;
        }
        if ((rc = pbf_step(s, wp, wnp, wv, wnv, wi, n, stream))) break;
        float ms = 0.0f;
        if ((e = hipEventRecord(e1, st)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess ||
            (e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) {
          rc = fail_hip(s, e, "autotune timing");
          break;
        }
        if (!(first && r == 0)) t.push_back(ms);  // the very first run warms up
      }
      if (rc) break;
      // every shape must give the same bits (the canonical summation order)
      out.resize(2 * vb + ib);
      if ((e = hipMemcpy(out.data(), wnp, vb, hipMemcpyDeviceToHost)) != hipSuccess ||
          (e = hipMemcpy(out.data() + vb, wnv, vb, hipMemcpyDeviceToHost)) != hipSuccess ||
          (e = hipMemcpy(out.data() + 2 * vb, wi, ib, hipMemcpyDeviceToHost)) != hipSuccess) {
        rc = fail_hip(s, e, "autotune result copy");
        break;
      }
      if (first) {
        ref.swap(out);
      } else if (std::memcmp(ref.data(), out.data(), ref.size()) != 0) {
        rc = fail(s, PBF_ERR_STATE, "launch shape (%d, %d) changed the result", blocks[bi],
                  grids[gi]);
        break;
      }
      first = false;
      std::sort(t.begin(), t.end());
      const double med = t[t.size() / 2];
      if (med < win_ms) {
        win_ms = med;
        win = s->shapes;
      }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(st);
  (void)hipFree(work);
  s->graph_on = graph_saved;
  s->timer.on = timer_saved;
  drop_graphs(s);
  if (rc) {
    s->shapes = saved;
    return rc;
  }
  s->shapes = win;
  if (best) {
    best->light_block = win.light_block;
    best->chunk_blocks = win.chunk_blocks;
  }
  if (best_ms) *best_ms = win_ms;
  return PBF_OK;
}

int pbf_sweep_box(const pbf_sweep* sw, int frame, pbf_vec3* ulim, pbf_vec3* llim) {
  if (!sw || !ulim || !llim) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "NULL argument");
  float phi = 0.0f;
  if (frame >= sw->start_frame) {
    // FluidSystem::stepSimulate (fluids/FluidSystem.cpp:107-109), FP32
    const float t = sw->w * (float)(frame - sw->start_frame);
    phi = std::sin(t);
  }
  ulim->x = sw->ulim0.x + sw->a_ulim.x * phi;
  ulim->y = sw->ulim0.y + sw->a_ulim.y * phi;
  ulim->z = sw->ulim0.z + sw->a_ulim.z * phi;
  llim->x = sw->llim0.x + sw->a_llim.x * phi;
  llim->y = sw->llim0.y + sw->a_llim.y * phi;
  llim->z = sw->llim0.z + sw->a_llim.z * phi;
  return PBF_OK;
}

int pbf_run_frames(pbf_solver* s, float* pos_a, float* vel_a, float* pos_b, float* vel_b,
                   uint32_t* iid, int n, int first_frame, int nframes, const pbf_sweep* sweep,
                   void* stream, int* which) {
  if (!s) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "solver is NULL");
  if (!which) return fail(s, PBF_ERR_INVALID_ARGUMENT, "which is NULL");
  *which = 0;
  if (nframes < 0) return fail(s, PBF_ERR_INVALID_ARGUMENT, "nframes must be >= 0");
  if (sweep && s->nscenes) return fail(s, PBF_ERR_STATE, "the boxes of a batched solver are fixed");
  for (int k = 0; k < nframes; ++k) {
    const int frame = first_frame + k;
    if (sweep) {
      pbf_vec3 ul, ll;
      pbf_sweep_box(sweep, frame, &ul, &ll);
      const int rc = pbf_set_lim(s, ul, ll);
      if (rc) return rc;
    }
    // tic-toc of fluids/FluidSystem.cpp:112-117
    const bool tic = (k & 1) == 0;
    const int rc = tic ? pbf_step(s, pos_a, pos_b, vel_a, vel_b, iid, n, stream)
                       : pbf_step(s, pos_b, pos_a, vel_b, vel_a, iid, n, stream);
    if (rc) return rc;
    *which = tic ? 1 : 0;
  }
  return PBF_OK;
}

int pbf_density_error(pbf_solver* s, const float* pos, int n, void* stream, double out[2]) {
  if (!s) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "solver is NULL");
  if (!out) return fail(s, PBF_ERR_INVALID_ARGUMENT, "out is NULL");
  out[0] = out[1] = 0.0;
  if (n < 0) return fail(s, PBF_ERR_INVALID_ARGUMENT, "n must be >= 0, got %d", n);
  if (n > s->capacity)
    return fail(s, PBF_ERR_CAPACITY, "n = %d exceeds capacity %d", n, s->capacity);
  if (s->nscenes && n != s->capacity)
    return fail(s, PBF_ERR_INVALID_ARGUMENT, "a batched solver needs n = %d, got %d", s->capacity,
                n);
  if (n == 0) return PBF_OK;
  int rc;
  if ((rc = check_vec_ptr(s, pos, "pos"))) return rc;
  pbf::StepConsts c;
  if ((rc = step_consts(s, c))) return rc;
  DeviceGuard g(s->device);
  if (!g.ok) return fail(s, PBF_ERR_HIP, "hipSetDevice(%d) failed", s->device);
  hipStream_t st = (hipStream_t)stream;
  const pbf::Scenes scn = scenes_of(s);
  if (s->cnt_dirty) pbf::launch_fill_u32(s->cnt, s->max_cells, 0u, st);
  s->cnt_dirty = true;
  pbf::launch_key_only(s->layout, pos, n, c, scn, s->xp, s->key0, s->val0, s->counters, st);
  const int which = pbf::radix_sort_pairs(s->key0, s->val0, s->key1, s->val1, n,
                                          key_bits_for(c.cells), s->hist, s->totals, st);
  pbf::launch_permute_pos(which ? s->key1 : s->key0, which ? s->val1 : s->val0, s->xp, n, s->xa,
                          s->cnt, st);
  pbf::launch_cell_starts(s->cnt, c.cells, s->tile_sum, s->cs, st);
  pbf::launch_density_error(s->xa, s->cs, n, c, scn, s->dens_err, s->dens_partial, s->dens_out,
                            s->hl, s->shapes, st);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(s, e, "kernel launch");
  s->cnt_dirty = false;
  double host[2];
  if ((e = hipMemcpyAsync(host, s->dens_out, sizeof(host), hipMemcpyDeviceToHost, st)) != hipSuccess)
    return fail_hip(s, e, "hipMemcpyAsync");
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail_hip(s, e, "hipStreamSynchronize");
  out[0] = host[0];
  out[1] = host[1];
  return PBF_OK;
}

int pbf_set_timing(pbf_solver* s, int enable) {
  if (!s) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "solver is NULL");
  s->timer.on = enable != 0;
  return PBF_OK;
}

int pbf_get_timing(pbf_solver* s, double ms_total[PBF_NUM_PHASES],
                   long long count[PBF_NUM_PHASES]) {
  if (!s || !ms_total || !count) return fail(s, PBF_ERR_INVALID_ARGUMENT, "NULL argument");
  Timer& t = s->timer;
  DeviceGuard g(s->device);
  if (!t.recs.empty()) {
    hipError_t e = hipEventSynchronize(t.recs.back().b);
    if (e != hipSuccess) return fail_hip(s, e, "hipEventSynchronize");
    for (auto& r : t.recs) {
      float ms = 0.0f;
      if ((e = hipEventElapsedTime(&ms, r.a, r.b)) != hipSuccess)
        return fail_hip(s, e, "hipEventElapsedTime");
      t.ms[r.phase] += ms;
      t.cnt[r.phase] += 1;
      t.pool.push_back(r.a);
      t.pool.push_back(r.b);
    }
    t.recs.clear();
  }
  for (int p = 0; p < PBF_NUM_PHASES; ++p) {
    ms_total[p] = t.ms[p];
    count[p] = t.cnt[p];
    t.ms[p] = 0.0;
    t.cnt[p] = 0;
  }
  return PBF_OK;
}


int pbf_tape_create(pbf_solver* s, int n, int nframes, pbf_tape** out) {
  return pbf_tape_create_ex(s, n, nframes, 0, out);
}

int pbf_tape_create_ex(pbf_solver* s, int n, int nframes, int flags, pbf_tape** out) {
  if (!out) return fail(s, PBF_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (!s) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "solver is NULL");
  if (s->nscenes) return fail(s, PBF_ERR_STATE, "tapes are for single-scene solvers");
  if (n < 1 || n > s->capacity)
    return fail(s, PBF_ERR_INVALID_ARGUMENT, "n must be in [1, capacity = %d], got %d", s->capacity,
                n);
  if (nframes < 1 || nframes > 100000)
    return fail(s, PBF_ERR_INVALID_ARGUMENT, "nframes must be in [1, 100000], got %d", nframes);
  if (flags & ~(PBF_TAPE_STORE | PBF_TAPE_GATHER | PBF_TAPE_SCATTER | PBF_TAPE_RESCALE))
    return fail(s, PBF_ERR_INVALID_ARGUMENT, "unknown tape flags %d", flags);
  if ((flags & PBF_TAPE_GATHER) && (flags & PBF_TAPE_SCATTER))
    return fail(s, PBF_ERR_INVALID_ARGUMENT, "PBF_TAPE_GATHER and PBF_TAPE_SCATTER exclude each other");
  DeviceGuard g(s->device);
  if (!g.ok) return fail(s, PBF_ERR_HIP, "hipSetDevice(%d) failed", s->device);
  pbf_tape* t = new (std::nothrow) pbf_tape();
  if (!t) return fail(s, PBF_ERR_INVALID_ARGUMENT, "out of host memory");
  t->device = s->device;
  t->n = n;
  t->nframes = nframes;
  t->niter = s->prm.niter;
  t->layout = s->layout;
  t->store = (flags & PBF_TAPE_STORE) != 0;
  t->adjoint_flag = flags & (PBF_TAPE_GATHER | PBF_TAPE_SCATTER);
  t->rescale = (flags & PBF_TAPE_RESCALE) != 0;
  const size_t N = (size_t)n, L = (size_t)s->layout, NI = (size_t)t->niter;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    const size_t o = off;
    off = align_up(off + bytes);
    return o;
  };
  std::vector<size_t> o_cp, o_cv, o_ci;
  for (int f = 0; f <= nframes; ++f) {
    o_cp.push_back(carve(4 * L * N));
    o_cv.push_back(carve(4 * L * N));
    o_ci.push_back(carve(4 * N));
  }
  const size_t o_wp = carve(4 * L * N), o_wv = carve(4 * L * N), o_wnp = carve(4 * L * N),
               o_wnv = carve(4 * L * N), o_wi = carve(4 * N);
  std::vector<size_t> o_xl, o_sd, o_dp, o_rho;
  for (size_t k = 0; k < NI; ++k) {
    o_xl.push_back(carve(16 * N));
    o_sd.push_back(carve(16 * N));
    o_dp.push_back(carve(16 * N));
    o_rho.push_back(carve(4 * N));
  }
  size_t o_f4[12];
  for (int q = 0; q < 12; ++q)  // ex_x and ex_v (2, 3) hold kScatterCopies copies
    o_f4[q] = carve(16 * N * (q == 2 || q == 3 ? (size_t)pbf::kScatterCopies : 1));
  const size_t o_qk0 = carve(4 * N), o_qv0 = carve(4 * N), o_qk1 = carve(4 * N),
               o_qv1 = carve(4 * N), o_qhist = carve(4 * 256 * (size_t)pbf::radix_tiles(n)),
               o_qtot = carve(4 * 256), o_pc = carve(4 * N), o_scale = carve(8);
  const size_t o_grx = carve(4 * N), o_qc = carve(4 * N), o_gp0 = carve(4 * L * N), o_gp1 = carve(4 * L * N),
               o_gv0 = carve(4 * L * N), o_gv1 = carve(4 * L * N), o_r0 = carve(8 * N),
               o_r0s = carve(8 * (1 + 1024));
  std::vector<size_t> o_fxl, o_fsd, o_fdp, o_frho, o_fxa, o_fvr, o_fkey, o_fval, o_fcs;
  if (t->store) {
    const size_t cs_words = (size_t)s->max_cells + 1;
    for (int f = 0; f < nframes; ++f) {
      for (size_t k = 0; k < NI; ++k) {
        o_fxl.push_back(carve(16 * N));
        o_fsd.push_back(carve(16 * N));
        o_fdp.push_back(carve(16 * N));
        o_frho.push_back(carve(4 * N));
      }
      o_fxa.push_back(carve(16 * N));
      o_fvr.push_back(carve(16 * N));
      o_fkey.push_back(carve(4 * N));
      o_fval.push_back(carve(4 * N));
      o_fcs.push_back(carve(4 * cs_words));
    }
  }
  hipError_t e = hipMalloc(&t->arena, off);
  if (e != hipSuccess) {
    delete t;
    return fail_hip(s, e, "hipMalloc of the tape");
  }
  char* b = (char*)t->arena;
  for (int f = 0; f <= nframes; ++f) {
    t->cpos.push_back((float*)(b + o_cp[f]));
    t->cvel.push_back((float*)(b + o_cv[f]));
    t->ciid.push_back((uint32_t*)(b + o_ci[f]));
  }
  t->wpos = (float*)(b + o_wp);
  t->wvel = (float*)(b + o_wv);
  t->wnpos = (float*)(b + o_wnp);
  t->wnvel = (float*)(b + o_wnv);
  t->wiid = (uint32_t*)(b + o_wi);
  for (size_t k = 0; k < NI; ++k) {
    t->xl.push_back((float4*)(b + o_xl[k]));
    t->sd.push_back((float4*)(b + o_sd[k]));
    t->dp.push_back((float4*)(b + o_dp[k]));
    t->rho.push_back((float*)(b + o_rho[k]));
  }
  float4** f4[12] = {&t->out_x, &t->out_v, &t->ex_x, &t->ex_v, &t->gx, &t->gxn,
                     &t->gx0,   &t->A,     &t->xr,   &t->sdb,  &t->gxo, &t->gvo};
  for (int q = 0; q < 12; ++q) *f4[q] = (float4*)(b + o_f4[q]);
  t->grho_x = (float*)(b + o_grx);
  t->qc = (uint32_t*)(b + o_qc);
  t->qk0 = (uint32_t*)(b + o_qk0);
  t->qv0 = (uint32_t*)(b + o_qv0);
  t->qk1 = (uint32_t*)(b + o_qk1);
  t->qv1 = (uint32_t*)(b + o_qv1);
  t->qhist = (uint32_t*)(b + o_qhist);
  t->qtotals = (uint32_t*)(b + o_qtot);
  t->pc = (uint32_t*)(b + o_pc);
  t->maxbits = (uint32_t*)(b + o_scale);
  t->escale = (int*)(b + o_scale + 4);
  t->gp[0] = (float*)(b + o_gp0);
  t->gp[1] = (float*)(b + o_gp1);
  t->gv[0] = (float*)(b + o_gv0);
  t->gv[1] = (float*)(b + o_gv1);
  t->r0 = (double*)(b + o_r0);
  t->r0_sum = (double*)(b + o_r0s);
  for (size_t q = 0; q < o_fxl.size(); ++q) {
    t->f_xl.push_back((float4*)(b + o_fxl[q]));
    t->f_sd.push_back((float4*)(b + o_fsd[q]));
    t->f_dp.push_back((float4*)(b + o_fdp[q]));
    t->f_rho.push_back((float*)(b + o_frho[q]));
  }
  for (size_t q = 0; q < o_fxa.size(); ++q) {
    t->f_xa.push_back((float4*)(b + o_fxa[q]));
    t->f_vr.push_back((float4*)(b + o_fvr[q]));
    t->f_key.push_back((uint32_t*)(b + o_fkey[q]));
    t->f_val.push_back((uint32_t*)(b + o_fval[q]));
    t->f_cs.push_back((uint32_t*)(b + o_fcs[q]));
  }
  // the scatter arrays start (and stay between kernels) all zero
  pbf::launch_fill_f4(t->ex_x, n * pbf::kScatterCopies, nullptr);
  pbf::launch_fill_f4(t->ex_v, n * pbf::kScatterCopies, nullptr);
  if ((e = hipDeviceSynchronize()) != hipSuccess) {
    (void)hipFree(t->arena);
    delete t;
    return fail_hip(s, e, "clearing the tape");
  }
  *out = t;
  return PBF_OK;
}

int pbf_tape_gradient_scale(const pbf_tape* t, int* log2_scale) {
  if (!t || !log2_scale) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "NULL argument");
  *log2_scale = t->log2_scale;
  return PBF_OK;
}

void pbf_tape_destroy(pbf_tape* t) {
  if (!t) return;
  {
    DeviceGuard g(t->device);
    if (t->arena) (void)hipFree(t->arena);
  }
  delete t;
}

int pbf_tape_state(const pbf_tape* t, int frame, float** pos, float** vel, uint32_t** iid) {
  if (!t) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "tape is NULL");
  if (frame < 0 || frame > t->nframes)
    return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "frame must be in [0, %d], got %d", t->nframes,
                frame);
  if (pos) *pos = t->cpos[(size_t)frame];
  if (vel) *vel = t->cvel[(size_t)frame];
  if (iid) *iid = t->ciid[(size_t)frame];
  return PBF_OK;
}

int pbf_tape_read(const pbf_tape* t, int frame, float* pos, float* vel, uint32_t* iid,
                  void* stream) {
  if (!t) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "tape is NULL");
  if (frame < 0 || frame > t->nframes)
    return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "frame must be in [0, %d], got %d", t->nframes,
                frame);
  DeviceGuard g(t->device);
  hipStream_t st = (hipStream_t)stream;
  const size_t vb = (size_t)t->n * (size_t)t->layout * 4, ib = (size_t)t->n * 4;
  hipError_t e = hipSuccess;
  if (pos) e = hipMemcpyAsync(pos, t->cpos[(size_t)frame], vb, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess && vel)
    e = hipMemcpyAsync(vel, t->cvel[(size_t)frame], vb, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess && iid)
    e = hipMemcpyAsync(iid, t->ciid[(size_t)frame], ib, hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return fail_hip(nullptr, e, "tape read");
  return PBF_OK;
}

int pbf_tape_forward(pbf_solver* s, pbf_tape* t, const float* pos, const float* vel,
                     const uint32_t* iid, int first_frame, const pbf_sweep* sweep,
                     void* stream) {
  if (!s) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "solver is NULL");
  if (!t) return fail(s, PBF_ERR_INVALID_ARGUMENT, "tape is NULL");
  if (t->device != s->device || t->layout != s->layout || t->niter != s->prm.niter)
    return fail(s, PBF_ERR_STATE, "tape made for another solver state (device, layout or niter)");
  int rc;
  if ((rc = check_vec_ptr(s, pos, "pos")) || (rc = check_vec_ptr(s, vel, "vel"))) return rc;
  if (!iid) return fail(s, PBF_ERR_INVALID_ARGUMENT, "iid is NULL");
  DeviceGuard g(s->device);
  if (!g.ok) return fail(s, PBF_ERR_HIP, "hipSetDevice(%d) failed", s->device);
  hipStream_t st = (hipStream_t)stream;
  const size_t vb = (size_t)t->n * (size_t)t->layout * 4, ib = (size_t)t->n * 4;
  hipError_t e;
  t->forward_done = false;
  if ((e = hipMemcpyAsync(t->cpos[0], pos, vb, hipMemcpyDeviceToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(t->cvel[0], vel, vb, hipMemcpyDeviceToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(t->ciid[0], iid, ib, hipMemcpyDeviceToDevice, st)) != hipSuccess)
    return fail_hip(s, e, "tape checkpoint copy");
  t->ul.assign((size_t)t->nframes, pbf_vec3{});
  t->ll.assign((size_t)t->nframes, pbf_vec3{});
  for (int f = 0; f < t->nframes; ++f) {
    if (sweep) {
      pbf_vec3 u, l;
      pbf_sweep_box(sweep, first_frame + f, &u, &l);
      if ((rc = pbf_set_lim(s, u, l))) return rc;
    }
    t->ul[(size_t)f] = s->ulim;
    t->ll[(size_t)f] = s->llim;
    // the step clobbers its inputs: step a copy of checkpoint f
    if ((e = hipMemcpyAsync(t->wpos, t->cpos[f], vb, hipMemcpyDeviceToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(t->wvel, t->cvel[f], vb, hipMemcpyDeviceToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(t->ciid[f + 1], t->ciid[f], ib, hipMemcpyDeviceToDevice, st)) !=
            hipSuccess)
      return fail_hip(s, e, "tape checkpoint copy");
    if (!t->store) {
      if ((rc = pbf_step(s, t->wpos, t->cpos[f + 1], t->wvel, t->cvel[f + 1], t->ciid[f + 1],
                         t->n, stream)))
        return rc;
      continue;
    }
    // stored tape: the step keeps its iterations in frame f's arrays, then
    // the grid, the sort and the pre-XSPH state the adjoint reads are kept
    pbf::StepConsts c;
    if ((rc = step_consts(s, c))) return rc;
    const size_t q0 = (size_t)f * (size_t)t->niter;
    Record rec;
    rec.xl = &t->f_xl[q0];
    rec.sd = &t->f_sd[q0];
    rec.dp = &t->f_dp[q0];
    rec.rho = &t->f_rho[q0];
    if (s->cnt_dirty) pbf::launch_fill_u32(s->cnt, s->max_cells, 0u, st);
    s->cnt_dirty = true;
    const StepArrays arr{t->wpos, t->cpos[f + 1], t->wvel, t->cvel[f + 1], t->ciid[f + 1], t->n};
    enqueue_step(s, arr, c, st, false, &rec);
    const size_t fb = (size_t)f, v4 = (size_t)t->n * 16;
    if ((e = hipMemcpyAsync(t->f_key[fb], rec.skey, ib, hipMemcpyDeviceToDevice, st)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(t->f_val[fb], rec.sval, ib, hipMemcpyDeviceToDevice, st)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(t->f_cs[fb], s->cs, ((size_t)c.cells + 1) * 4,
                            hipMemcpyDeviceToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(t->f_xa[fb], s->xa, v4, hipMemcpyDeviceToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(t->f_vr[fb], s->vr, v4, hipMemcpyDeviceToDevice, st)) != hipSuccess)
      return fail_hip(s, e, "tape frame copy");
    if ((e = hipGetLastError()) != hipSuccess) return fail_hip(s, e, "kernel launch");
    s->cnt_dirty = false;
  }
  t->prm = s->prm;
  t->forward_done = true;
  return PBF_OK;
}

int pbf_tape_backward(pbf_solver* s, pbf_tape* t, const float* gpos, const float* gvel,
                      float* gpos0, float* gvel0, double* grho0, void* stream) {
  if (!s) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "solver is NULL");
  if (!t) return fail(s, PBF_ERR_INVALID_ARGUMENT, "tape is NULL");
  if (!t->forward_done) return fail(s, PBF_ERR_STATE, "pbf_tape_forward has not run on this tape");
  if (t->device != s->device) return fail(s, PBF_ERR_STATE, "tape made on another device");
  if (std::memcmp(&t->prm, &s->prm, sizeof(pbf_params)) != 0 || t->layout != s->layout)
    return fail(s, PBF_ERR_STATE, "parameters changed since the forward run");
  int rc;
  if ((rc = check_vec_ptr(s, gpos0, "gpos0")) || (rc = check_vec_ptr(s, gvel0, "gvel0"))) return rc;
  if ((gpos && (rc = check_vec_ptr(s, gpos, "gpos"))) || (gvel && (rc = check_vec_ptr(s, gvel, "gvel"))))
    return rc;
  DeviceGuard g(s->device);
  if (!g.ok) return fail(s, PBF_ERR_HIP, "hipSetDevice(%d) failed", s->device);
  hipStream_t st = (hipStream_t)stream;
  const int n = t->n, L = t->layout, NI = t->niter;
  // every frame below runs under its own box; the caller's box is restored
  // on every return
  struct BoxRestore {
    pbf_solver* s;
    pbf_vec3 ul, ll;
    ~BoxRestore() {
      s->ulim = ul;
      s->llim = ll;
    }
  } box_restore{s, s->ulim, s->llim};
  const size_t vb = (size_t)n * (size_t)L * 4, ib = (size_t)n * 4;
  hipError_t e;
  // seeds: dL/d(final state), caller layout
  if (gpos) {
    e = hipMemcpyAsync(t->gp[0], gpos, vb, hipMemcpyDeviceToDevice, st);
  } else {
    e = hipMemsetAsync(t->gp[0], 0, vb, st);
  }
  if (e == hipSuccess)
    e = gvel ? hipMemcpyAsync(t->gv[0], gvel, vb, hipMemcpyDeviceToDevice, st)
             : hipMemsetAsync(t->gv[0], 0, vb, st);
  if (e != hipSuccess) return fail_hip(s, e, "seeding the backward pass");
  pbf::launch_fill_f64(t->r0, n, st);
  if (t->rescale) {
    pbf::launch_fill_u32(reinterpret_cast<uint32_t*>(t->escale), 1, 0u, st);
    pbf::launch_rescale(t->gp[0], n * L, t->gv[0], n * L, t->maxbits, t->escale, st);
  }
  // Adjoint mode: gather (each pass over a grid of its own positions, no
  // atomics) from kGatherMinParticles particles or on request, when every
  // frame's grid has at most 1024 cells per axis; else scatter.
  bool gather = t->adjoint_flag != PBF_TAPE_SCATTER &&
                (t->adjoint_flag == PBF_TAPE_GATHER || n >= kGatherMinParticles);
  for (int f = 0; gather && f < t->nframes; ++f) {
    pbf::SceneGeo geo;
    long long cells = 0;
    if ((rc = make_geo(s, s->prm, t->ul[(size_t)f], t->ll[(size_t)f], 0, geo, cells))) return rc;
    gather = geo.dx <= 1024 && geo.dy <= 1024 && geo.dz <= 1024;
  }
  if (t->adjoint_flag == PBF_TAPE_GATHER && !gather)
    return fail(s, PBF_ERR_INVALID_ARGUMENT,
                "PBF_TAPE_GATHER needs grids of at most 1024 cells per axis");
  pbf::BwdQuery bq{};
  if (gather) {
    // the query-order arrays take the scatter sums' memory
    t->ex_dirty = true;
    char* ex = (char*)t->ex_x;
    char* ev = (char*)t->ex_v;
    const size_t v16 = 16 * (size_t)n;
    bq.k0 = t->qk0;
    bq.v0 = t->qv0;
    bq.k1 = t->qk1;
    bq.v1 = t->qv1;
    bq.hist = t->qhist;
    bq.totals = t->qtotals;
    bq.cnt = s->cnt;
    bq.tile_sum = s->tile_sum;
    bq.cs = s->cs;
    bq.pc = t->pc;
    bq.xq = (float4*)ex;
    bq.aq = (float4*)(ex + v16);
    bq.xr = (float4*)(ex + 2 * v16);
    bq.sdb = (float4*)(ex + 3 * v16);
    bq.cq = (uint2*)(ex + 4 * v16);
    bq.vrq = (float4*)ev;
    bq.gvq = (float4*)(ev + v16);
    if (s->cnt_dirty) {
      pbf::launch_fill_u32(s->cnt, s->max_cells, 0u, st);
      s->cnt_dirty = false;
    }
  } else if (t->ex_dirty) {
    pbf::launch_fill_f4(t->ex_x, n * pbf::kScatterCopies, st);
    pbf::launch_fill_f4(t->ex_v, n * pbf::kScatterCopies, st);
    t->ex_dirty = false;
  }
  mark_push("pbf.tape_backward");
  struct PopOnExit {
    ~PopOnExit() { mark_pop(); }
  } pop_on_exit;
  std::vector<float4*> xl(t->xl), sd(t->sd), dp(t->dp);
  std::vector<float*> rho(t->rho);
  Record rec;
  rec.skip_xsph = true;
  rec.xl = xl.data();
  rec.sd = sd.data();
  rec.dp = dp.data();
  rec.rho = rho.data();
  int cur = 0;
  // Neighbour passes of the adjoint kernels: after a recompute they follow
  // its passes 0 .. 2 niter - 1 (it leaves out XSPH); a stored tape numbers
  // them on from the forward run's last pass (any full step leaves the odd
  // parity's counters cleared).
  int pass = 2 * NI + 1;
  for (int f = t->nframes - 1; f >= 0; --f) {
    // frame f under its own box
    s->ulim = t->ul[(size_t)f];
    s->llim = t->ll[(size_t)f];
    pbf::StepConsts c;
    if ((rc = step_consts(s, c))) return rc;
    const uint32_t* cs = s->cs;
    const float4 *x_final = s->xa, *vr = s->vr;
    const uint32_t *skey = nullptr, *sval = nullptr;
    std::vector<pbf::BwdIter> iters((size_t)NI);
    if (t->store) {
      const size_t q0 = (size_t)f * (size_t)NI, fb = (size_t)f;
      for (size_t k = 0; k < (size_t)NI; ++k)
        iters[k] = pbf::BwdIter{t->f_xl[q0 + k], t->f_sd[q0 + k], t->f_rho[q0 + k],
                                t->f_dp[q0 + k]};
      cs = t->f_cs[fb];
      x_final = t->f_xa[fb];
      vr = t->f_vr[fb];
      skey = t->f_key[fb];
      sval = t->f_val[fb];
    } else {
      // recompute frame f from its checkpoint, keeping its iterations
      if ((e = hipMemcpyAsync(t->wpos, t->cpos[f], vb, hipMemcpyDeviceToDevice, st)) !=
              hipSuccess ||
          (e = hipMemcpyAsync(t->wvel, t->cvel[f], vb, hipMemcpyDeviceToDevice, st)) !=
              hipSuccess ||
          (e = hipMemcpyAsync(t->wiid, t->ciid[f], ib, hipMemcpyDeviceToDevice, st)) != hipSuccess)
        return fail_hip(s, e, "tape checkpoint copy");
      if (s->cnt_dirty) pbf::launch_fill_u32(s->cnt, s->max_cells, 0u, st);
      s->cnt_dirty = true;
      const StepArrays arr{t->wpos, t->wnpos, t->wvel, t->wnvel, t->wiid, n};
      enqueue_step(s, arr, c, st, false, &rec);
      for (size_t k = 0; k < (size_t)NI; ++k)
        iters[k] = pbf::BwdIter{t->xl[k], t->sd[k], t->rho[k], t->dp[k]};
      skey = rec.skey;
      sval = rec.sval;
      // the recompute's last neighbour pass was the last delta-p pass (2 NI -
      // 1), whose light kernel cleared the counters of pass 2 NI
      pass = 2 * NI;
    }
    // adjoint of the frame
    pbf::launch_to_f4(L, t->gp[cur], n, t->gxo, st);
    pbf::launch_to_f4(L, t->gv[cur], n, t->gvo, st);
    pbf::BwdFrame bf{};
    bf.x_final = x_final;
    bf.vr = vr;
    bf.key = skey;
    bf.gx_out = t->gxo;
    bf.gv_out = t->gvo;
    bf.out_x = t->out_x;
    bf.out_v = t->out_v;
    bf.ex_x = t->ex_x;
    bf.ex_v = t->ex_v;
    bf.gx = t->gx;
    bf.gxn = t->gxn;
    bf.gx0 = t->gx0;
    bf.A = t->A;
    bf.xr = t->xr;
    bf.sdb = t->sdb;
    bf.grho_x = t->grho_x;
    bf.qc = t->qc;
    bf.q = gather ? &bq : nullptr;
    bf.escale = t->rescale ? t->escale : nullptr;
    bq.key_bits = key_bits_for(c.cells);
    pbf::launch_bwd_frame(bf, n, c, st);
    bf.r0 = t->r0;
    pbf::launch_bwd_xsph(bf, n, c, cs, s->hl, pass++, s->shapes, st);
    for (int k = NI - 1; k >= 0; --k) {
      pbf::launch_bwd_iteration(bf, iters[(size_t)k], k == NI - 1, n, c, cs, s->hl, pass,
                                s->shapes, st);
      pass += 2;
    }
    float* op = f == 0 ? gpos0 : t->gp[cur ^ 1];
    float* ov = f == 0 ? gvel0 : t->gv[cur ^ 1];
    pbf::launch_bwd_predict(L, bf, sval, n, c, op, ov, st);
    if (t->rescale && f > 0) pbf::launch_rescale(op, n * L, ov, n * L, t->maxbits, t->escale, st);
    cur ^= 1;
    if ((e = hipGetLastError()) != hipSuccess) return fail_hip(s, e, "kernel launch");
    s->cnt_dirty = false;
  }
  pbf::launch_sum_f64(t->r0, n, t->r0_sum + 1, t->r0_sum, st);
  double host = 0.0;
  int escale = 0;
  if ((e = hipMemcpyAsync(&host, t->r0_sum, sizeof(double), hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (t->rescale &&
       (e = hipMemcpyAsync(&escale, t->escale, sizeof(int), hipMemcpyDeviceToHost, st)) != hipSuccess))
    return fail_hip(s, e, "hipMemcpyAsync");
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail_hip(s, e, "hipStreamSynchronize");
  t->log2_scale = escale;
  if (grho0) *grho0 = host;
  return PBF_OK;
}

int pbf_step_noalloc(const pbf_params* params, pbf_vec3 ulim, pbf_vec3 llim, float* pos,
                     float* npos, float* vel, float* nvel, uint32_t* iid, int n, void* stream) {
  int rc = validate_params(nullptr, params);
  if (rc) return rc;
  if (n < 0 || n > kNoAllocParticles)
    return fail(nullptr, PBF_ERR_CAPACITY, "n must be in [0, %d], got %d", kNoAllocParticles, n);
  if (n == 0) return PBF_OK;
  if (!pos || !npos || !vel || !nvel || !iid)
    return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "NULL particle array");
  {
    const void* p[5] = {pos, npos, vel, nvel, iid};
    for (int a = 0; a < 5; ++a)
      for (int b = a + 1; b < 5; ++b)
        if (p[a] == p[b])
          return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "particle arrays must be distinct");
  }
  pbf::StepConsts c;
  if ((rc = make_consts(nullptr, *params, ulim, llim, c))) return rc;
  if (c.cells > kNoAllocCells)
    return fail(nullptr, PBF_ERR_CAPACITY, "box needs %d cells > %lld", c.cells, kNoAllocCells);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail_hip(nullptr, e, "hipGetDevice");
  if (dev < 0 || dev >= kNoAllocDevices)
    return fail(nullptr, PBF_ERR_STATE, "device %d beyond the %d the no-alloc step serves", dev,
                kNoAllocDevices);
  NoAllocCtx& ctx = g_noalloc[dev];
  std::lock_guard<std::mutex> lock(ctx.mu);
  if ((rc = noalloc_bind(ctx, dev))) return rc;
  pbf_solver& s = ctx.s;
  s.prm = *params;
  s.ulim = ulim;
  s.llim = llim;
  hipStream_t st = (hipStream_t)stream;
  // the scratch keeps the previous call's contents: the per-cell counts and
  // the work-list counters must start at zero (fill kernels, so the call can
  // be captured into a graph)
  pbf::launch_fill_u32(s.cnt, kNoAllocCells, 0u, st);
  pbf::launch_fill_u32(s.counters, pbf::kCounterWords, 0u, st);
  const StepArrays arr{pos, npos, vel, nvel, iid, n};
  enqueue_step(&s, arr, c, st, false);
  if ((e = hipGetLastError()) != hipSuccess) return fail_hip(nullptr, e, "kernel launch");
  return PBF_OK;
}

}  // extern "C"
