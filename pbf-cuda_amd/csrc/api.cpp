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
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "pbf.h"
#include "pbf_launch.h"

namespace {

// fluids/helper.h:4-9 (M_PI is the reference's literal).
constexpr double kPi = 3.14159265359;
constexpr double kLimEps = 1e-3;
constexpr float kKernelEps = 1e-4f;
constexpr float kMaxDp = 0.1f;
constexpr int kMaxIter = 1 << 16;
constexpr long long kMaxCellsLimit = 1ll << 31;

thread_local std::string g_create_error;

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

struct pbf_solver {
  int device = 0;
  pbf_params prm{};
  pbf_vec3 ulim{}, llim{};
  int capacity = 0;
  int layout = PBF_LAYOUT_FLOAT4;
  long long max_cells = 0;

  void* arena = nullptr;
  uint32_t *key0 = nullptr, *val0 = nullptr, *key1 = nullptr, *val1 = nullptr;
  uint32_t *iid_s = nullptr, *hist = nullptr, *totals = nullptr;
  float4 *xp = nullptr, *x0s = nullptr, *xa = nullptr;
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

// Grid dims and every per-step constant, computed on the host exactly once
// per step (the reference recomputes the dims in every buildGridHash,
// fluids/Simulator.cu:187-188, because the box may move).
int make_consts(pbf_solver* s, const pbf_params& p, pbf_vec3 ul, pbf_vec3 ll,
                pbf::StepConsts& c) {
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
  if (cells > kMaxCellsLimit) return fail(s, PBF_ERR_CAPACITY, "grid has %lld cells (> 2^31)", cells);

  std::memset(&c, 0, sizeof(c));
  c.llx = ll.x;
  c.lly = ll.y;
  c.llz = ll.z;
  c.ulx = ul.x;
  c.uly = ul.y;
  c.ulz = ul.z;
  c.dx = d[0];
  c.dy = d[1];
  c.dz = d[2];
  c.cells = (int)cells;
  c.h = p.h;
  c.h2 = p.h * p.h;
  c.h2_outer = c.h2 * 1.0001f;
  c.keps = kKernelEps;
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
  // Box clamp of computetpos (fluids/Simulator_kernel.cuh:190-192): the bound
  // ulim - LIM_EPS is formed in double there; rounding it to FP32 once gives
  // the same clamped FP32 value.
  c.lox = (float)((double)ll.x + kLimEps);
  c.loy = (float)((double)ll.y + kLimEps);
  c.loz = (float)((double)ll.z + kLimEps);
  c.hix = (float)((double)ul.x - kLimEps);
  c.hiy = (float)((double)ul.y - kLimEps);
  c.hiz = (float)((double)ul.z - kLimEps);
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
  if (capacity < 1 || capacity > (1 << 30))
    return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "capacity must be in [1, 2^30], got %d", capacity);
  if (layout != PBF_LAYOUT_FLOAT3 && layout != PBF_LAYOUT_FLOAT4)
    return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "layout must be 3 or 4, got %d", layout);
  pbf::StepConsts c;
  rc = make_consts(nullptr, *params, ulim, llim, c);
  if (rc) return rc;
  // Simulator.h:13-14 sizes the grid at 4x the initial box's cells.
  if (max_cells <= 0) max_cells = 4ll * c.cells;
  if (max_cells > kMaxCellsLimit) max_cells = kMaxCellsLimit;
  if (max_cells < c.cells)
    return fail(nullptr, PBF_ERR_CAPACITY, "max_cells %lld < cells of the box %d", max_cells, c.cells);

  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev < 1)
    return fail(nullptr, PBF_ERR_NO_DEVICE, "no HIP device: %s", hipGetErrorString(e));
  int dev = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return fail_hip(nullptr, e, "hipGetDevice");

  pbf_solver* s = new (std::nothrow) pbf_solver();
  if (!s) return fail(nullptr, PBF_ERR_INVALID_ARGUMENT, "out of host memory");
  s->device = dev;
  s->prm = *params;
  s->ulim = ulim;
  s->llim = llim;
  s->capacity = capacity;
  s->layout = layout;
  s->max_cells = max_cells;

  const size_t N = (size_t)capacity;
  const size_t tiles = (size_t)pbf::radix_tiles(capacity);
  const size_t dblk = (size_t)pbf::density_partials(capacity);
  // chunk items of crowded targets: 2 per particle covers piles holding a
  // tenth of the particles; beyond that the targets are summed in place
  const size_t ICAP = 2 * N + 1024;
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    const size_t o = off;
    off = align_up(off + bytes);
    return o;
  };
  const size_t o_key0 = carve(4 * N), o_val0 = carve(4 * N), o_key1 = carve(4 * N),
               o_val1 = carve(4 * N), o_iid = carve(4 * N), o_hist = carve(4 * 256 * tiles),
               o_tot = carve(4 * 256), o_xp = carve(16 * N), o_x0s = carve(16 * N),
               o_xa = carve(16 * N), o_rho = carve(4 * N), o_recs = carve(16 * N),
               o_items = carve(8 * ICAP), o_part = carve(4 * pbf::kPartialWords * ICAP),
               o_cnt = carve(4 * 8), o_de = carve(8 * N), o_dp = carve(16 * dblk),
               o_do = carve(16), o_cnt2 = carve(4 * (size_t)max_cells),
               o_cs = carve(4 * ((size_t)max_cells + 1)),
               o_ts = carve(4 * ((size_t)pbf::scan_tiles(max_cells) + 1));
  void* arena = nullptr;
  if ((e = hipMalloc(&arena, off)) != hipSuccess) {
    delete s;
    return fail_hip(nullptr, e, "hipMalloc of solver scratch");
  }
  char* b = (char*)arena;
  s->arena = arena;
  s->key0 = (uint32_t*)(b + o_key0);
  s->val0 = (uint32_t*)(b + o_val0);
  s->key1 = (uint32_t*)(b + o_key1);
  s->val1 = (uint32_t*)(b + o_val1);
  s->iid_s = (uint32_t*)(b + o_iid);
  s->hist = (uint32_t*)(b + o_hist);
  s->totals = (uint32_t*)(b + o_tot);
  s->xp = (float4*)(b + o_xp);
  s->x0s = (float4*)(b + o_x0s);
  s->xa = (float4*)(b + o_xa);
  s->rho = (float*)(b + o_rho);
  s->counters = (uint32_t*)(b + o_cnt);
  s->hl.counters = s->counters;
  s->hl.recs = (uint4*)(b + o_recs);
  s->hl.items = (uint2*)(b + o_items);
  s->hl.partials = (float*)(b + o_part);
  s->hl.item_cap = (uint32_t)ICAP;
  s->dens_err = (double*)(b + o_de);
  s->dens_partial = (double*)(b + o_dp);
  s->dens_out = (double*)(b + o_do);
  s->cnt = (uint32_t*)(b + o_cnt2);
  s->cs = (uint32_t*)(b + o_cs);
  s->tile_sum = (uint32_t*)(b + o_ts);
  // the per-cell counts start (and stay, between steps) all zero
  pbf::launch_fill_u32(s->cnt, max_cells, 0u, nullptr);
  if ((e = hipDeviceSynchronize()) != hipSuccess) {
    (void)hipFree(arena);
    delete s;
    return fail_hip(nullptr, e, "clearing the cell counts");
  }
  *out = s;
  return PBF_OK;
}

void pbf_destroy(pbf_solver* s) {
  if (!s) return;
  {
    DeviceGuard g(s->device);
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
  pbf::StepConsts c;
  rc = make_consts(s, *params, s->ulim, s->llim, c);
  if (rc) return rc;
  if (c.cells > s->max_cells)
    return fail(s, PBF_ERR_CAPACITY, "h gives %d cells > max_cells %lld", c.cells, s->max_cells);
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
  if ((rc = make_consts(s, s->prm, s->ulim, s->llim, c))) return rc;
  if (c.cells > s->max_cells)
    return fail(s, PBF_ERR_CAPACITY, "box needs %d cells > max_cells %lld", c.cells, s->max_cells);
  DeviceGuard g(s->device);
  if (!g.ok) return fail(s, PBF_ERR_HIP, "hipSetDevice(%d) failed", s->device);
  hipStream_t st = (hipStream_t)stream;
  const int L = s->layout;

  tbegin(s, PBF_PHASE_STEP, st);
  // advect() + getGridId (Simulator.cpp:50, Simulator.cu:165-193)
  tbegin(s, PBF_PHASE_PREDICT, st);
  if (s->cnt_dirty) pbf::launch_fill_u32(s->cnt, s->max_cells, 0u, st);
  s->cnt_dirty = true;
  pbf::launch_predict_key(L, pos, vel, n, c, s->xp, s->key0, s->val0, s->cnt, s->counters, st);
  tend(s, PBF_PHASE_PREDICT, st);
  // thrust::sort_by_key (Simulator.cu:196-198)
  tbegin(s, PBF_PHASE_SORT, st);
  const int which = pbf::radix_sort_pairs(s->key0, s->val0, s->key1, s->val1, n,
                                          key_bits_for(c.cells), s->hist, s->totals, st);
  tend(s, PBF_PHASE_SORT, st);
  const uint32_t* sval = which ? s->val1 : s->val0;
  // memset + computeGridRange (Simulator.cu:200-204), fused with the payload gather
  tbegin(s, PBF_PHASE_GRID, st);
  pbf::launch_cell_starts(s->cnt, c.cells, s->tile_sum, s->cs, st);
  pbf::launch_permute(L, sval, pos, s->xp, iid, n, s->x0s, s->xa, s->iid_s, st);
  tend(s, PBF_PHASE_GRID, st);
  // niter x correctDensity (Simulator.cpp:62-66, Simulator.cu:213-249);
  // xl (x, lambda) aliases xp, which is dead after the permute.
  float4* xl = s->xp;
  int pass = 0;
  for (int it = 0; it < s->prm.niter; ++it) {
    tbegin(s, PBF_PHASE_LAMBDA, st);
    pbf::launch_lambda(s->xa, s->cs, n, c, xl, it + 1 == s->prm.niter ? s->rho : nullptr,
                       s->hl, pass++, st);
    tend(s, PBF_PHASE_LAMBDA, st);
    tbegin(s, PBF_PHASE_DELTA, st);
    pbf::launch_delta(xl, s->cs, n, c, s->xa, s->hl, pass++, st);
    tend(s, PBF_PHASE_DELTA, st);
  }
  // updateVelocity (Simulator.cpp:72, Simulator.cu:267-274); vr aliases xp.
  float4* vr = s->xp;
  tbegin(s, PBF_PHASE_VEL, st);
  pbf::launch_velocity(L, s->xa, s->x0s, s->iid_s, s->rho, n, c, vr, pos, vel, iid, st);
  tend(s, PBF_PHASE_VEL, st);
  // correctVelocity (Simulator.cpp:77, Simulator.cu:251-265)
  tbegin(s, PBF_PHASE_XSPH, st);
  pbf::launch_xsph(L, s->xa, vr, s->cs, n, c, npos, nvel, s->hl, pass, st);
  tend(s, PBF_PHASE_XSPH, st);
  tend(s, PBF_PHASE_STEP, st);

  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(s, e, "kernel launch");
  s->cnt_dirty = false;
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
  if (n == 0) return PBF_OK;
  int rc;
  if ((rc = check_vec_ptr(s, pos, "pos"))) return rc;
  pbf::StepConsts c;
  if ((rc = make_consts(s, s->prm, s->ulim, s->llim, c))) return rc;
  if (c.cells > s->max_cells)
    return fail(s, PBF_ERR_CAPACITY, "box needs %d cells > max_cells %lld", c.cells, s->max_cells);
  DeviceGuard g(s->device);
  if (!g.ok) return fail(s, PBF_ERR_HIP, "hipSetDevice(%d) failed", s->device);
  hipStream_t st = (hipStream_t)stream;
  if (s->cnt_dirty) pbf::launch_fill_u32(s->cnt, s->max_cells, 0u, st);
  s->cnt_dirty = true;
  pbf::launch_key_only(s->layout, pos, n, c, s->xp, s->key0, s->val0, s->cnt, s->counters, st);
  const int which = pbf::radix_sort_pairs(s->key0, s->val0, s->key1, s->val1, n,
                                          key_bits_for(c.cells), s->hist, s->totals, st);
  pbf::launch_cell_starts(s->cnt, c.cells, s->tile_sum, s->cs, st);
  pbf::launch_permute_pos(which ? s->val1 : s->val0, s->xp, n, s->xa, st);
  pbf::launch_density_error(s->xa, s->cs, n, c, s->dens_err, s->dens_partial, s->dens_out,
                            s->hl, st);
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

}  // extern "C"
