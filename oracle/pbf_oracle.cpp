// pbf_oracle.cpp — TEST INFRASTRUCTURE, not product code.
//
// Host C++ restatement of the reference PBF step (naeioi/PBF-CUDA).  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
// library; the product path (libpbf.so) never links or calls it.
//
// One templated solver, two instantiations:
//  * T = double — the FP64 oracle: the reference algorithm in FP64.
//  * T = float  — the FP32 mirror: the same algorithm with a fixed FP32
//    operation sequence (IEEE +,-,*,/,sqrt and explicit fused multiply-adds,
//    no other contraction).  The GPU kernels use the same sequence, so their
//    results must equal the mirror's bit for bit; the FP64 instantiation
//    bounds how far that sequence is from exact arithmetic.
// Integer-producing parts (predicted positions feeding the keys, grid dims,
// keys, query cells) are FP32 in both instantiations, exactly as the GPU.
// Each function cites the reference lines it restates.
#include "pbf_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// fluids/helper.h:4-9 (M_PI is the reference's 12-digit literal).
constexpr double kPi = 3.14159265359;
constexpr double kLimEps = 1e-3;
constexpr double kKernelEps = 1e-4;
constexpr double kMaxDp = 0.1;

int nthreads(int t) {
#ifdef _OPENMP
  return t > 0 ? t : omp_get_max_threads();
#else
  (void)t;
  return 1;
#endif
}

// One axis of getGridxyz / getGridId (fluids/Simulator.cu:30-36, :64-71):
// (int)((p - llim) / h) clamped to [0, dim-1], all in FP32.  Clamping the
// FP32 quotient before truncation gives the same integer for every finite
// quotient and maps NaN to 0.
inline int cell_coord_f32(float p, float llim, float h, int dim) {
  const float q = (p - llim) / h;
  const float c = std::fmin(std::fmax(q, 0.0f), (float)(dim - 1));
  return (int)c;
}

inline int cell_coord_f64(double p, double llim, double h, int dim) {
  const double q = (p - llim) / h;
  const double c = std::fmin(std::fmax(q, 0.0), (double)(dim - 1));
  return (int)c;
}

inline float fma_(float a, float b, float c) { return std::fmaf(a, b, c); }
inline double fma_(double a, double b, double c) { return std::fma(a, b, c); }

struct Grid {
  int d[3] = {0, 0, 0};
  long long cells = 0;
  std::vector<uint32_t> start, end;
};

// The 27-cell walk of computeLambda / computetpos / computeXSPH
// (fluids/Simulator_kernel.cuh:78-104, :153-170, :214-233): dx, then dy, then
// dz from -1 to +1, cells outside [0, D) skipped, j ascending inside a cell.
template <class F>
inline void for_each_candidate(const Grid& g, const int q[3], F&& f) {
  for (int ox = -1; ox <= 1; ++ox) {
    const int x = q[0] + ox;
    if (x < 0 || x >= g.d[0]) continue;
    for (int oy = -1; oy <= 1; ++oy) {
      const int y = q[1] + oy;
      if (y < 0 || y >= g.d[1]) continue;
      for (int oz = -1; oz <= 1; ++oz) {
        const int z = q[2] + oz;
        if (z < 0 || z >= g.d[2]) continue;
        // xyzToId, fluids/Simulator.cu:52: x-major, z fastest.
        const long long c = ((long long)x * g.d[1] + y) * g.d[2] + z;
        for (uint32_t j = g.start[c]; j < g.end[c]; ++j) f(j);
      }
    }
  }
}

// Step constants.  For T = float they are formed exactly as the GPU host
// code forms them (FP64 formula rounded once to FP32); for T = double they
// are the FP64 values.
template <class T>
struct Consts {
  T h, h2, h2_outer, keps, poly6, spiky, spiky_rho0, pho0, lambda_eps, coef_corr, n_corr;
  int n_corr_int;
  T k_boundary, wall_c, lo[3], hi[3], ul[3], ll[3], max_dp, dt, inv_dt, c_xsph;
};

template <class T>
Consts<T> make_consts(const orc_params* p, const orc_box* box) {
  Consts<T> c;
  const double h = (double)p->h;
  // getPoly6 / getSpikyGrad coefficients (fluids/Simulator.cu:77-83, :94-98).
  const double poly6 = 315.0 / (64.0 * kPi * std::pow(h, 9));
  const double spiky = -45.0 / (kPi * std::pow(h, 6));
  const double dq = (double)p->delta_q;
  const double wd = h * h - dq * dq;
  // coef_corr = -k_corr / W(delta_q)^n_corr (fluids/Simulator.cu:235).
  const double coef_corr = -(double)p->k_corr / std::pow(poly6 * wd * wd * wd, (double)p->n_corr);
  c.h = (T)p->h;
  c.h2 = c.h * c.h;
  c.h2_outer = c.h2 * (T)1.0001;
  c.keps = (T)kKernelEps;
  c.poly6 = (T)poly6;
  c.spiky = (T)spiky;
  c.spiky_rho0 = (T)(spiky / (double)p->pho0);
  c.pho0 = (T)p->pho0;
  c.lambda_eps = (T)p->lambda_eps;
  c.coef_corr = (T)coef_corr;
  c.n_corr = (T)p->n_corr;
  c.n_corr_int = (p->n_corr >= 0.0f && p->n_corr <= 16.0f && std::floor(p->n_corr) == p->n_corr)
                     ? (int)p->n_corr
                     : -1;
  c.k_boundary = (T)p->k_boundaryDensity;
  c.wall_c = (T)(2.0 * kPi / 3.0);
  for (int a = 0; a < 3; ++a) {
    // Box clamp, fluids/Simulator_kernel.cuh:190-192 (bound formed in double).
    c.lo[a] = (T)((double)box->llim[a] + kLimEps);
    c.hi[a] = (T)((double)box->ulim[a] - kLimEps);
    c.ul[a] = (T)box->ulim[a];
    c.ll[a] = (T)box->llim[a];
  }
  c.max_dp = (T)kMaxDp;
  c.dt = (T)p->dt;
  c.inv_dt = (T)1 / c.dt;  // h_updateVelocity, fluids/Simulator.cu:129
  c.c_xsph = (T)p->c_XSPH;
  return c;
}

// x^n for the s_corr exponent (powf in computetpos,
// fluids/Simulator_kernel.cuh:165): binary exponentiation for small integer
// n, pow otherwise.
template <class T>
inline T pow_corr(T w, const Consts<T>& c) {
  if (c.n_corr_int >= 0) {
    T r = (T)1, b = w;
    int e = c.n_corr_int;
    while (e) {
      if (e & 1) r = r * b;
      b = b * b;
      e >>= 1;
    }
    return r;
  }
  return std::pow(w, c.n_corr);
}

// DensityBoundary, fluids/Simulator.cu:139-161.
template <class T>
inline T wall_density_at(T d, const Consts<T>& c) {
  if (d > c.h) return (T)0;
  if (d <= (T)0) return c.wall_c;
  return ((c.wall_c * (c.h - d)) * (c.h - d)) * (c.h + d);
}

template <class T>
inline T boundary_density(const T* x, const Consts<T>& c) {
  T s = wall_density_at(c.ul[0] - x[0], c);
  s = s + wall_density_at(x[0] - c.ll[0], c);
  s = s + wall_density_at(c.ul[1] - x[1], c);
  s = s + wall_density_at(x[1] - c.ll[1], c);
  s = s + wall_density_at(c.ul[2] - x[2], c);
  s = s + wall_density_at(x[2] - c.ll[2], c);
  return s;
}

template <class T>
inline T dist2(const T* xi, const T* xj, T d[3]) {
  d[0] = xi[0] - xj[0];
  d[1] = xi[1] - xj[1];
  d[2] = xi[2] - xj[2];
  T r2 = d[0] * d[0];
  r2 = fma_(d[1], d[1], r2);
  r2 = fma_(d[2], d[2], r2);
  return r2;
}

// Canonical summation order (shared with the GPU kernels, which must
// reproduce it bit for bit): a target's candidates, in the reference's walk
// order, are summed in consecutive chunks of kChunk candidates, each chunk
// summed sequentially from zero and the chunk sums added in chunk order (one
// chunk: the plain sequential sum).  The chunked form lets the GPU spread a
// crowded target's candidates over many work-items.
constexpr size_t kChunk = 64;

template <class T, int K>
struct Acc {
  T v[K] = {};
};

template <class T, int K, class F>
Acc<T, K> canonical_sum(const std::vector<uint32_t>& cand, F&& add) {
  Acc<T, K> total;
  const size_t m = cand.size();
  for (size_t c0 = 0; c0 < m; c0 += kChunk) {
    Acc<T, K> part;
    const size_t c1 = std::min(m, c0 + kChunk);
    for (size_t u = c0; u < c1; ++u) add(part, cand[u]);
    for (int k = 0; k < K; ++k) total.v[k] = total.v[k] + part.v[k];
  }
  return total;
}

bool finite_params(const orc_params* p) {
  const float v[] = {p->pho0, p->g, p->h, p->dt, p->lambda_eps, p->delta_q,
                     p->k_corr, p->n_corr, p->k_boundaryDensity, p->c_XSPH};
  for (float f : v)
    if (!std::isfinite(f)) return false;
  return p->h > 0.0f && p->dt > 0.0f && p->pho0 > 0.0f && p->niter >= 1 &&
         p->lambda_eps > 0.0f && p->delta_q >= 0.0f && p->delta_q < p->h;
}

void build_grid(const std::vector<uint32_t>& skeys, int n, const int dims[3], Grid& g) {
  g.d[0] = dims[0];
  g.d[1] = dims[1];
  g.d[2] = dims[2];
  g.cells = (long long)dims[0] * dims[1] * dims[2];
  g.start.assign((size_t)g.cells, 0u);
  g.end.assign((size_t)g.cells, 0u);
  orc_cell_ranges(skeys.data(), n, g.cells, g.start.data(), g.end.data());
}

template <class T>
void query_cell(const T* x, const orc_box* box, float h, const int dims[3], bool f64, int q[3]) {
  for (int a = 0; a < 3; ++a) {
    q[a] = f64 ? cell_coord_f64((double)x[a], (double)box->llim[a], (double)h, dims[a])
               : cell_coord_f32((float)x[a], box->llim[a], h, dims[a]);
  }
}

// Simulator::step, fluids/Simulator.cpp:10-99, with the phases of
// fluids/Simulator.cu:165-274 and the kernels of
// fluids/Simulator_kernel.cuh:7-239.
template <class T>
int step_impl(const orc_params* p, const orc_box* box, int n, float* pos, float* npos, float* vel,
              float* nvel, uint32_t* iid, const orc_options* opt, orc_diag* diag) {
  if (!p || !box || n < 0 || !finite_params(p)) return -1;
  if (n > 0 && (!pos || !npos || !vel || !nvel || !iid)) return -1;
  const int stride = opt ? opt->stride : 4;
  if (stride != 3 && stride != 4) return -1;
  const bool brute = opt && opt->brute_force;
  const bool q64 = opt && opt->query_cell_f64;
  const int TH = nthreads(opt ? opt->threads : 0);
  int dims[3];
  if (orc_grid_dims(box, p->h, dims) != 0) return -1;
  if (n == 0) return 0;

  const float dt = p->dt;
  const float gvec[3] = {0.0f, 0.0f, -p->g};  // Simulator.cu:174, z is up

  // advect_kernel, fluids/Simulator_kernel.cuh:14-15: vel += dt*g,
  // npos = pos + dt*vel, as FP32 fused multiply-adds (what the GPU does, so
  // the cell keys agree bit for bit).
  std::vector<float> xs((size_t)3 * n);
  for (int i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      const float v1 = std::fmaf(dt, gvec[a], vel[(size_t)i * stride + a]);
      xs[(size_t)3 * i + a] = std::fmaf(dt, v1, pos[(size_t)i * stride + a]);
    }
  }

  // buildGridHash, fluids/Simulator.cu:178-211.
  std::vector<uint32_t> keys(n), perm(n), skeys(n);
  orc_cell_keys(xs.data(), 3, n, box, p->h, dims, keys.data());
  orc_stable_order(keys.data(), n, perm.data());
  for (int s = 0; s < n; ++s) skeys[s] = keys[perm[s]];
  Grid g;
  build_grid(skeys, n, dims, g);

  std::vector<T> x0((size_t)3 * n), x((size_t)3 * n), xn((size_t)3 * n);
  std::vector<uint32_t> id(n);
  for (int s = 0; s < n; ++s) {
    const uint32_t src = perm[s];
    for (int a = 0; a < 3; ++a) {
      x0[(size_t)3 * s + a] = (T)pos[(size_t)src * stride + a];
      x[(size_t)3 * s + a] = (T)xs[(size_t)3 * src + a];
    }
    id[s] = iid[src];
  }

  const Consts<T> c = make_consts<T>(p, box);
  std::vector<T> lam(n, (T)0), rho(n, (T)0);

  // The candidate list of a target in the reference's walk order.
  auto collect = [&](const T* xi, std::vector<uint32_t>& cand) {
    cand.clear();
    if (brute) {
      for (int j = 0; j < n; ++j) cand.push_back((uint32_t)j);
    } else {
      int q[3];
      query_cell(xi, box, p->h, dims, q64, q);
      for_each_candidate(g, q, [&](uint32_t j) { cand.push_back(j); });
    }
  };

  for (int it = 0; it < p->niter; ++it) {
    // computeLambda, fluids/Simulator_kernel.cuh:52-129: rho sums poly6 over
    // every candidate, self included (:91); the spiky gradient / pho0 is
    // summed into grad_i (:93) and, for j != i, its squared norm into
    // gradj_l2 (:98-100); wall term (:119-120); lambda (:122-123).
#pragma omp parallel num_threads(TH)
    {
      std::vector<uint32_t> cand;
#pragma omp for schedule(dynamic, 256)
      for (int i = 0; i < n; ++i) {
        const T* xi = &x[(size_t)3 * i];
        collect(xi, cand);
        // v = {rho, grad_i x, y, z, gradj_l2}
        Acc<T, 5> a = canonical_sum<T, 5>(cand, [&](Acc<T, 5>& s, uint32_t j) {
          T d[3];
          const T r2 = dist2(xi, &x[(size_t)3 * j], d);
          if (r2 < c.h2_outer) {
            if (r2 < c.h2) {
              const T t = c.h2 - r2;
              s.v[0] = s.v[0] + ((c.poly6 * t) * t) * t;
            }
            const T rl = std::sqrt(r2);
            if (rl >= c.keps && rl < c.h) {
              const T t = c.h - rl;
              const T f = ((c.spiky_rho0 * t) * t) / rl;
              const T ax = f * d[0], ay = f * d[1], az = f * d[2];
              s.v[1] = s.v[1] + ax;
              s.v[2] = s.v[2] + ay;
              s.v[3] = s.v[3] + az;
              if (j != (uint32_t)i) {
                s.v[4] = fma_(ax, ax, s.v[4]);
                s.v[4] = fma_(ay, ay, s.v[4]);
                s.v[4] = fma_(az, az, s.v[4]);
              }
            }
          }
        });
        T r = a.v[0];
        if (c.k_boundary != (T)0) r = r + c.k_boundary * boundary_density(xi, c);
        T den = fma_(a.v[1], a.v[1], a.v[4]);
        den = fma_(a.v[2], a.v[2], den);
        den = fma_(a.v[3], a.v[3], den);
        den = den + c.lambda_eps;
        lam[i] = -(r / c.pho0 - (T)1) / den;
        rho[i] = r;
      }
    }
    // computetpos, fluids/Simulator_kernel.cuh:131-194: sum over j != i of
    // (lambda_i + lambda_j + s_corr) * spikyGrad (:163-166), / pho0, clamp to
    // +-MAX_DP (:184), add (:188), clamp into the shrunk box (:190-192).
    // Jacobi: all reads from x, writes to xn, then x <- xn (thrust::copy_n,
    // fluids/Simulator.cu:247-248).
#pragma omp parallel num_threads(TH)
    {
      std::vector<uint32_t> cand;
#pragma omp for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
      const T* xi = &x[(size_t)3 * i];
      collect(xi, cand);
      Acc<T, 3> a = canonical_sum<T, 3>(cand, [&](Acc<T, 3>& acc, uint32_t j) {
        if (j == (uint32_t)i) return;
        T d[3];
        const T r2 = dist2(xi, &x[(size_t)3 * j], d);
        if (r2 < c.h2_outer) {
          const T rl = std::sqrt(r2);
          if (rl >= c.keps && rl < c.h) {
            T w = 0;
            if (r2 < c.h2) {
              const T t2 = c.h2 - r2;
              w = ((c.poly6 * t2) * t2) * t2;
            }
            const T corr = c.coef_corr * pow_corr(w, c);
            const T t = c.h - rl;
            const T gs = ((c.spiky * t) * t) / rl;
            const T s = ((lam[i] + lam[j]) + corr) * gs;
            acc.v[0] = fma_(s, d[0], acc.v[0]);
            acc.v[1] = fma_(s, d[1], acc.v[1]);
            acc.v[2] = fma_(s, d[2], acc.v[2]);
          }
        }
      });
      const T px = a.v[0], py = a.v[1], pz = a.v[2];
      if (it == 0 && diag && diag->dp_raw_first) {
        diag->dp_raw_first[(size_t)3 * i + 0] = (double)px;
        diag->dp_raw_first[(size_t)3 * i + 1] = (double)py;
        diag->dp_raw_first[(size_t)3 * i + 2] = (double)pz;
      }
      const T dp[3] = {px, py, pz};
      for (int a = 0; a < 3; ++a) {
        T v = dp[a] / c.pho0;
        v = std::fmin(std::fmax(v, -c.max_dp), c.max_dp);
        T y = xi[a] + v;
        y = std::fmax(std::fmin(y, c.hi[a]), c.lo[a]);
        xn[(size_t)3 * i + a] = y;
      }
    }
    }
    x.swap(xn);
  }

  // updateVelocity, fluids/Simulator.cu:267-274 / :127-137:
  // v = (npos - pos) * (1/dt).
  std::vector<T> v((size_t)3 * n), vf((size_t)3 * n);
  for (size_t k = 0; k < (size_t)3 * n; ++k) v[k] = (x[k] - x0[k]) * c.inv_dt;

  // computeXSPH, fluids/Simulator_kernel.cuh:196-239 (j == i included, rho
  // of the last lambda pass).  A pair with W = 0 adds exactly 0 to the
  // reference's sum whenever rho_i + rho_j != 0, so it is skipped.
#pragma omp parallel num_threads(TH)
  {
    std::vector<uint32_t> cand;
#pragma omp for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
      const T* xi = &x[(size_t)3 * i];
      const T* vi = &v[(size_t)3 * i];
      collect(xi, cand);
      Acc<T, 3> a = canonical_sum<T, 3>(cand, [&](Acc<T, 3>& acc, uint32_t j) {
        T d[3];
        const T r2 = dist2(xi, &x[(size_t)3 * j], d);
        if (r2 < c.h2) {
          const T t = c.h2 - r2;
          const T w = ((c.poly6 * t) * t) * t;
          const T* vj = &v[(size_t)3 * j];
          const T f = ((T)2 * w) / (rho[i] + rho[j]);
          acc.v[0] = fma_(vj[0] - vi[0], f, acc.v[0]);
          acc.v[1] = fma_(vj[1] - vi[1], f, acc.v[1]);
          acc.v[2] = fma_(vj[2] - vi[2], f, acc.v[2]);
        }
      });
      vf[(size_t)3 * i + 0] = fma_(c.c_xsph, a.v[0], vi[0]);  // :235
      vf[(size_t)3 * i + 1] = fma_(c.c_xsph, a.v[1], vi[1]);
      vf[(size_t)3 * i + 2] = fma_(c.c_xsph, a.v[2], vi[2]);
    }
  }

  // Outputs in sorted order, FP32: the contract of Simulator::step.
  for (int s = 0; s < n; ++s) {
    for (int a = 0; a < 3; ++a) {
      pos[(size_t)s * stride + a] = (float)x0[(size_t)3 * s + a];
      vel[(size_t)s * stride + a] = (float)v[(size_t)3 * s + a];
      npos[(size_t)s * stride + a] = (float)x[(size_t)3 * s + a];
      nvel[(size_t)s * stride + a] = (float)vf[(size_t)3 * s + a];
    }
    if (stride == 4) {
      pos[(size_t)s * 4 + 3] = vel[(size_t)s * 4 + 3] = 0.0f;
      npos[(size_t)s * 4 + 3] = nvel[(size_t)s * 4 + 3] = 0.0f;
    }
    iid[s] = id[s];
  }
  if (diag) {
    for (size_t k = 0; k < (size_t)3 * n; ++k) {
      if (diag->x_final) diag->x_final[k] = (double)x[k];
      if (diag->v_final) diag->v_final[k] = (double)vf[k];
      if (diag->v_pre) diag->v_pre[k] = (double)v[k];
    }
    for (int s = 0; s < n; ++s) {
      if (diag->lambda) diag->lambda[s] = (double)lam[s];
      if (diag->rho) diag->rho[s] = (double)rho[s];
    }
    if (diag->perm) std::memcpy(diag->perm, perm.data(), sizeof(uint32_t) * n);
    if (diag->keys) std::memcpy(diag->keys, skeys.data(), sizeof(uint32_t) * n);
  }
  return 0;
}

// Density error on a fresh grid: rho as in computeLambda
// (fluids/Simulator_kernel.cuh:77-120), self included, plus the wall term;
// per-particle rho in T, the error statistic in FP64.
template <class T>
int density_impl(const orc_params* p, const orc_box* box, int n, const float* xin, int stride,
                 int threads, double out[2], double* err_out) {
  if (!p || !box || !out || n < 0 || !finite_params(p) || (stride != 3 && stride != 4)) return -1;
  out[0] = out[1] = 0.0;
  int dims[3];
  if (orc_grid_dims(box, p->h, dims) != 0) return -1;
  if (n == 0) return 0;
  std::vector<uint32_t> keys(n), perm(n), skeys(n);
  orc_cell_keys(xin, stride, n, box, p->h, dims, keys.data());
  orc_stable_order(keys.data(), n, perm.data());
  std::vector<T> x((size_t)3 * n);
  for (int s = 0; s < n; ++s) {
    skeys[s] = keys[perm[s]];
    for (int a = 0; a < 3; ++a) x[(size_t)3 * s + a] = (T)xin[(size_t)perm[s] * stride + a];
  }
  Grid g;
  build_grid(skeys, n, dims, g);
  const Consts<T> c = make_consts<T>(p, box);
  std::vector<double> err(n);
  const int TH = nthreads(threads);
#pragma omp parallel num_threads(TH)
  {
    std::vector<uint32_t> cand;
#pragma omp for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
      const T* xi = &x[(size_t)3 * i];
      int q[3];
      query_cell(xi, box, p->h, dims, false, q);
      cand.clear();
      for_each_candidate(g, q, [&](uint32_t j) { cand.push_back(j); });
      Acc<T, 1> a = canonical_sum<T, 1>(cand, [&](Acc<T, 1>& acc, uint32_t j) {
        T d[3];
        const T r2 = dist2(xi, &x[(size_t)3 * j], d);
        if (r2 < c.h2) {
          const T t = c.h2 - r2;
          acc.v[0] = acc.v[0] + ((c.poly6 * t) * t) * t;
        }
      });
      T r = a.v[0];
      if (c.k_boundary != (T)0) r = r + c.k_boundary * boundary_density(xi, c);
      err[i] = std::fabs((double)r / (double)c.pho0 - 1.0);
    }
  }
  double sum = 0.0, mx = 0.0;
  for (int i = 0; i < n; ++i) {
    sum += err[i];
    mx = std::max(mx, err[i]);
  }
  out[0] = sum / n;
  out[1] = mx;
  if (err_out) {
    for (int s = 0; s < n; ++s) err_out[perm[s]] = err[s];
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Rollout over several frames and its adjoint (reverse-mode derivative).
//
// The step is differentiated with its discrete choices held fixed: the sort
// order, the stale grid, every target's candidate list, which pairs fall
// inside h / KERNAL_EPS, and which clamps are active (a clamped component
// passes no derivative).  The adjoint is written in scatter form: for every
// (target i, candidate j) pair of a pass, the pair term's derivative with
// respect to x_i is added to i and its derivative with respect to x_j to j,
// so it assumes nothing about the symmetry of candidate lists (the stale
// grid makes them asymmetric).  All derivative arithmetic is FP64; the
// forward state is T (double: the FP64 rollout; float: the GPU's FP32
// trajectory, operation for operation).
// ---------------------------------------------------------------------------

template <class T>
struct FrameTape {
  std::vector<uint32_t> perm;  // sorted slot -> input index
  Grid g;
  std::vector<T> X0;               // sorted start positions (3n)
  std::vector<std::vector<T>> x;   // x[k]: positions entering iteration k, x[niter] final (3n)
  std::vector<std::vector<T>> lam, rho, S, D, dp;  // per iteration: n, n, 3n, n, 3n
  std::vector<T> V;                // velocities before XSPH (3n)
};

template <class T>
void consts_with_pho0(Consts<T>& c, const orc_params* p, double pho0) {
  const double h = (double)p->h;
  const double spiky = -45.0 / (kPi * std::pow(h, 6));
  c.pho0 = (T)pho0;
  c.spiky_rho0 = (T)(spiky / pho0);
}

// One frame from (X, V, iid) in input order; outputs in sorted order.
template <class T>
void frame_forward(const orc_params* p, const orc_box* box, const Consts<T>& c, const int dims[3],
                   int TH, const std::vector<T>& X, const std::vector<T>& Vin,
                   const std::vector<uint32_t>& iid, FrameTape<T>& tp, std::vector<T>& Xout,
                   std::vector<T>& Vout, std::vector<uint32_t>& iid_out) {
  const int n = (int)iid.size();
  const T gz = (T)(-p->g);
  // advect_kernel (fluids/Simulator_kernel.cuh:14-15) as fused multiply-adds;
  // the keys come from the predicted positions rounded to FP32
  std::vector<T> xp((size_t)3 * n);
  std::vector<float> xs((size_t)3 * n);
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      const T v1 = fma_(c.dt, a == 2 ? gz : (T)0, Vin[(size_t)3 * i + a]);
      const T y = fma_(c.dt, v1, X[(size_t)3 * i + a]);
      xp[(size_t)3 * i + a] = y;
      xs[(size_t)3 * i + a] = (float)y;
    }
  std::vector<uint32_t> keys(n), skeys(n);
  tp.perm.resize(n);
  orc_cell_keys(xs.data(), 3, n, box, p->h, dims, keys.data());
  orc_stable_order(keys.data(), n, tp.perm.data());
  for (int s = 0; s < n; ++s) skeys[s] = keys[tp.perm[s]];
  build_grid(skeys, n, dims, tp.g);
  const int NI = p->niter;
  tp.X0.assign((size_t)3 * n, (T)0);
  tp.x.assign(NI + 1, std::vector<T>((size_t)3 * n));
  tp.lam.assign(NI, std::vector<T>(n));
  tp.rho.assign(NI, std::vector<T>(n));
  tp.D.assign(NI, std::vector<T>(n));
  tp.S.assign(NI, std::vector<T>((size_t)3 * n));
  tp.dp.assign(NI, std::vector<T>((size_t)3 * n));
  iid_out.resize(n);
  for (int s = 0; s < n; ++s) {
    const uint32_t src = tp.perm[s];
    for (int a = 0; a < 3; ++a) {
      tp.X0[(size_t)3 * s + a] = X[(size_t)3 * src + a];
      tp.x[0][(size_t)3 * s + a] = xp[(size_t)3 * src + a];
    }
    iid_out[s] = iid[src];
  }
  auto collect = [&](const T* xi, std::vector<uint32_t>& cand) {
    int q[3];
    query_cell(xi, box, p->h, dims, false, q);
    cand.clear();
    for_each_candidate(tp.g, q, [&](uint32_t j) { cand.push_back(j); });
  };
  for (int k = 0; k < NI; ++k) {
    const std::vector<T>& x = tp.x[k];
    std::vector<T>& lam = tp.lam[k];
#pragma omp parallel num_threads(TH)
    {
      std::vector<uint32_t> cand;
#pragma omp for schedule(dynamic, 256)
      for (int i = 0; i < n; ++i) {
        const T* xi = &x[(size_t)3 * i];
        collect(xi, cand);
        Acc<T, 5> a = canonical_sum<T, 5>(cand, [&](Acc<T, 5>& sacc, uint32_t j) {
          T d[3];
          const T r2 = dist2(xi, &x[(size_t)3 * j], d);
          if (r2 < c.h2_outer) {
            if (r2 < c.h2) {
              const T t = c.h2 - r2;
              sacc.v[0] = sacc.v[0] + ((c.poly6 * t) * t) * t;
            }
            const T rl = std::sqrt(r2);
            if (rl >= c.keps && rl < c.h) {
              const T t = c.h - rl;
              const T f = ((c.spiky_rho0 * t) * t) / rl;
              const T ax = f * d[0], ay = f * d[1], az = f * d[2];
              sacc.v[1] = sacc.v[1] + ax;
              sacc.v[2] = sacc.v[2] + ay;
              sacc.v[3] = sacc.v[3] + az;
              if (j != (uint32_t)i) {
                sacc.v[4] = fma_(ax, ax, sacc.v[4]);
                sacc.v[4] = fma_(ay, ay, sacc.v[4]);
                sacc.v[4] = fma_(az, az, sacc.v[4]);
              }
            }
          }
        });
        T r = a.v[0];
        if (c.k_boundary != (T)0) r = r + c.k_boundary * boundary_density(xi, c);
        T den = fma_(a.v[1], a.v[1], a.v[4]);
        den = fma_(a.v[2], a.v[2], den);
        den = fma_(a.v[3], a.v[3], den);
        den = den + c.lambda_eps;
        lam[i] = -(r / c.pho0 - (T)1) / den;
        tp.rho[k][i] = r;
        tp.D[k][i] = den;
        for (int q = 0; q < 3; ++q) tp.S[k][(size_t)3 * i + q] = a.v[1 + q];
      }
    }
    std::vector<T>& xn = tp.x[k + 1];
#pragma omp parallel num_threads(TH)
    {
      std::vector<uint32_t> cand;
#pragma omp for schedule(dynamic, 256)
      for (int i = 0; i < n; ++i) {
        const T* xi = &x[(size_t)3 * i];
        collect(xi, cand);
        Acc<T, 3> a = canonical_sum<T, 3>(cand, [&](Acc<T, 3>& acc, uint32_t j) {
          if (j == (uint32_t)i) return;
          T d[3];
          const T r2 = dist2(xi, &x[(size_t)3 * j], d);
          if (r2 < c.h2_outer) {
            const T rl = std::sqrt(r2);
            if (rl >= c.keps && rl < c.h) {
              T w = 0;
              if (r2 < c.h2) {
                const T t2 = c.h2 - r2;
                w = ((c.poly6 * t2) * t2) * t2;
              }
              const T corr = c.coef_corr * pow_corr(w, c);
              const T t = c.h - rl;
              const T gs = ((c.spiky * t) * t) / rl;
              const T sc = ((lam[i] + lam[j]) + corr) * gs;
              acc.v[0] = fma_(sc, d[0], acc.v[0]);
              acc.v[1] = fma_(sc, d[1], acc.v[1]);
              acc.v[2] = fma_(sc, d[2], acc.v[2]);
            }
          }
        });
        for (int q = 0; q < 3; ++q) {
          const T v = a.v[q] / c.pho0;
          tp.dp[k][(size_t)3 * i + q] = v;
          T y = xi[q] + std::fmin(std::fmax(v, -c.max_dp), c.max_dp);
          y = std::fmax(std::fmin(y, c.hi[q]), c.lo[q]);
          xn[(size_t)3 * i + q] = y;
        }
      }
    }
  }
  const std::vector<T>& xf = tp.x[NI];
  tp.V.resize((size_t)3 * n);
  for (size_t q = 0; q < (size_t)3 * n; ++q) tp.V[q] = (xf[q] - tp.X0[q]) * c.inv_dt;
  const std::vector<T>& rho = tp.rho[NI - 1];
  Vout.resize((size_t)3 * n);
#pragma omp parallel num_threads(TH)
  {
    std::vector<uint32_t> cand;
#pragma omp for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
      const T* xi = &xf[(size_t)3 * i];
      const T* vi = &tp.V[(size_t)3 * i];
      collect(xi, cand);
      Acc<T, 3> a = canonical_sum<T, 3>(cand, [&](Acc<T, 3>& acc, uint32_t j) {
        T d[3];
        const T r2 = dist2(xi, &xf[(size_t)3 * j], d);
        if (r2 < c.h2) {
          const T t = c.h2 - r2;
          const T w = ((c.poly6 * t) * t) * t;
          const T* vj = &tp.V[(size_t)3 * j];
          const T f = ((T)2 * w) / (rho[i] + rho[j]);
          acc.v[0] = fma_(vj[0] - vi[0], f, acc.v[0]);
          acc.v[1] = fma_(vj[1] - vi[1], f, acc.v[1]);
          acc.v[2] = fma_(vj[2] - vi[2], f, acc.v[2]);
        }
      });
      for (int q = 0; q < 3; ++q) Vout[(size_t)3 * i + q] = fma_(c.c_xsph, a.v[q], vi[q]);
    }
  }
  Xout = xf;
}

// Per-thread accumulation buffers of a scatter loop, summed in thread order
// (deterministic for a given thread count).
struct Scatter {
  int th, len;
  std::vector<std::vector<double>> buf;
  Scatter(int threads, int length) : th(threads), len(length), buf(threads) {
    for (auto& b : buf) b.assign((size_t)length, 0.0);
  }
  double* mine() {
#ifdef _OPENMP
    return buf[(size_t)omp_get_thread_num()].data();
#else
    return buf[0].data();
#endif
  }
  void add_into(std::vector<double>& out) const {
    for (int t = 0; t < th; ++t)
      for (int q = 0; q < len; ++q) out[(size_t)q] += buf[(size_t)t][(size_t)q];
  }
};

// Pair geometry of x_i - x_j in FP64, with the forward's branch decisions
// (taken in T exactly as the forward took them).
template <class T>
struct Pair {
  bool in_h2 = false;   // r2 < h2: poly6 terms
  bool spiky = false;   // r2 < h2_outer && keps <= r < h: gradient terms
  double d[3], r2 = 0, r = 0;
};

template <class T>
inline Pair<T> pair_of(const T* xi, const T* xj, const Consts<T>& c) {
  Pair<T> pr;
  T dT[3];
  const T r2 = dist2(xi, xj, dT);
  pr.in_h2 = r2 < c.h2;
  if (r2 < c.h2_outer) {
    const T rl = std::sqrt(r2);
    pr.spiky = rl >= c.keps && rl < c.h;
  }
  for (int a = 0; a < 3; ++a) pr.d[a] = (double)xi[a] - (double)xj[a];
  pr.r2 = pr.d[0] * pr.d[0] + pr.d[1] * pr.d[1] + pr.d[2] * pr.d[2];
  pr.r = std::sqrt(pr.r2);
  return pr;
}

// Derivative of the wall term DensityBoundary (fluids/Simulator.cu:139-161)
// at distance d from a wall: d/dd of (2pi/3)(h-d)^2(h+d) inside (0, h].
inline double wall_slope(double d, double h, double wall_c) {
  if (d > h || d <= 0.0) return 0.0;
  return -wall_c * (h - d) * (h + 3.0 * d);
}

// Backward pass of one frame: (gX, gV) = adjoints of the frame's outputs
// (x_final, v after XSPH) in sorted order; returns the adjoints of the
// frame's inputs (positions, velocities) in input order and adds dL/dpho0.
template <class T>
void frame_backward(const orc_params* p, const orc_box* box, const Consts<T>& c,
                    const int dims[3], int TH, const FrameTape<T>& tp,
                    const std::vector<double>& gX, const std::vector<double>& gV,
                    std::vector<double>& gXin, std::vector<double>& gVin, double& g_rho0) {
  const int n = (int)tp.perm.size();
  const int NI = p->niter;
  const double h = (double)c.h, h2 = h * h, poly6 = (double)c.poly6, spiky = (double)c.spiky;
  const double spiky_rho0 = (double)c.spiky_rho0, pho0 = (double)c.pho0;
  const double coef = (double)c.coef_corr, ncorr = (double)c.n_corr, eps = (double)c.lambda_eps;
  const double kb = (double)c.k_boundary, cx = (double)c.c_xsph, inv_dt = (double)c.inv_dt;
  const double dt = (double)c.dt, maxdp = (double)c.max_dp;
  auto collect = [&](const T* xi, std::vector<uint32_t>& cand) {
    int q[3];
    query_cell(xi, box, p->h, dims, false, q);
    cand.clear();
    for_each_candidate(tp.g, q, [&](uint32_t j) { cand.push_back(j); });
  };
  auto poly6_w = [&](double r2) {
    const double t = h2 - r2;
    return poly6 * t * t * t;
  };
  // dW/dx_i = -6 poly6 (h2 - r2)^2 d
  auto poly6_grad = [&](const double* d, double r2, double out[3]) {
    const double t = h2 - r2;
    const double f = -6.0 * poly6 * t * t;
    for (int a = 0; a < 3; ++a) out[a] = f * d[a];
  };
  // (coef, grad) of the spiky kernel: G = k (h-r)^2 d / r and Hess * A
  auto spiky_hess = [&](double k, const double* d, double r, const double* A, double out[3]) {
    const double phi = k * (h - r) * (h - r) / r;
    const double psi = -k * (h - r) * (h + r) / (r * r * r);
    const double dA = d[0] * A[0] + d[1] * A[1] + d[2] * A[2];
    for (int a = 0; a < 3; ++a) out[a] = phi * A[a] + psi * dA * d[a];
  };
  // d s_corr / d w for s = coef * w^n
  auto dscorr_dw = [&](double w) {
    if (c.n_corr_int == 0 || ncorr == 0.0) return 0.0;
    if (c.n_corr_int > 0) {
      double r = 1.0;
      for (int e = 1; e < c.n_corr_int; ++e) r *= w;
      return coef * ncorr * r;
    }
    return w > 0.0 ? coef * ncorr * std::pow(w, ncorr - 1.0) : 0.0;
  };
  auto scorr = [&](double w) {
    if (c.n_corr_int >= 0) {
      double r = 1.0;
      for (int e = 0; e < c.n_corr_int; ++e) r *= w;
      return coef * r;
    }
    return coef * std::pow(w, ncorr);
  };

  std::vector<double> gx(gX);           // adjoint of the current positions
  std::vector<double> gXs((size_t)3 * n, 0.0);  // adjoint of the sorted start positions
  std::vector<double> grho_x(n, 0.0);   // adjoint of rho of the last iteration, from XSPH
  // ---- XSPH (fluids/Simulator_kernel.cuh:196-239) and v = (x - x0)/dt
  {
    const std::vector<T>& xf = tp.x[NI];
    const std::vector<T>& rho = tp.rho[NI - 1];
    std::vector<double> gVp(gV);  // adjoint of V (identity path)
    Scatter sv(TH, 3 * n), sx(TH, 3 * n), sr(TH, n);
#pragma omp parallel num_threads(TH)
    {
      std::vector<uint32_t> cand;
      double* bv = sv.mine();
      double* bx = sx.mine();
      double* br = sr.mine();
#pragma omp for schedule(dynamic, 256)
      for (int i = 0; i < n; ++i) {
        const T* xi = &xf[(size_t)3 * i];
        collect(xi, cand);
        const double* B = &gV[(size_t)3 * i];
        for (uint32_t j : cand) {
          const Pair<T> pr = pair_of(xi, &xf[(size_t)3 * j], c);
          if (!pr.in_h2) continue;
          const double W = poly6_w(pr.r2);
          const double den = (double)rho[i] + (double)rho[j];
          const double f = 2.0 * cx * W / den;
          double u = 0.0;
          for (int a = 0; a < 3; ++a) {
            bv[3 * j + a] += f * B[a];
            bv[3 * i + a] -= f * B[a];
            u += B[a] * ((double)tp.V[3 * (size_t)j + a] - (double)tp.V[3 * (size_t)i + a]);
          }
          double P[3];
          poly6_grad(pr.d, pr.r2, P);
          const double fx = 2.0 * cx * u / den;
          for (int a = 0; a < 3; ++a) {
            bx[3 * i + a] += fx * P[a];
            bx[3 * j + a] -= fx * P[a];
          }
          const double fr = 2.0 * cx * W * u / (den * den);
          br[i] -= fr;
          br[j] -= fr;
        }
      }
    }
    sv.add_into(gVp);
    sx.add_into(gx);
    sr.add_into(grho_x);
    for (size_t q = 0; q < (size_t)3 * n; ++q) {
      gx[q] += gVp[q] * inv_dt;
      gXs[q] -= gVp[q] * inv_dt;
    }
  }
  // ---- solver iterations, last to first
  for (int k = NI - 1; k >= 0; --k) {
    const std::vector<T>& x = tp.x[k];
    const std::vector<T>& lam = tp.lam[k];
    std::vector<double> gxn((size_t)3 * n, 0.0), A((size_t)3 * n, 0.0), glam(n, 0.0);
    // position update and clamps (fluids/Simulator_kernel.cuh:184-192)
    for (int i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) {
        const size_t q = (size_t)3 * i + a;
        const T v = tp.dp[k][q];
        const T y = x[q] + std::fmin(std::fmax(v, -c.max_dp), c.max_dp);
        const double e = (y > c.lo[a] && y < c.hi[a]) ? gx[q] : 0.0;
        gxn[q] = e;
        const double dv = (double)v;
        const double db = (dv > -maxdp && dv < maxdp) ? e : 0.0;
        g_rho0 += db * (-dv / pho0);
        A[q] = db / pho0;
      }
    // delta-p pair terms (fluids/Simulator_kernel.cuh:153-170)
    {
      Scatter sx(TH, 3 * n), sl(TH, n);
#pragma omp parallel num_threads(TH)
      {
        std::vector<uint32_t> cand;
        double* bx = sx.mine();
        double* bl = sl.mine();
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < n; ++i) {
          const T* xi = &x[(size_t)3 * i];
          collect(xi, cand);
          const double* Ai = &A[(size_t)3 * i];
          for (uint32_t j : cand) {
            if (j == (uint32_t)i) continue;
            const Pair<T> pr = pair_of(xi, &x[(size_t)3 * j], c);
            if (!pr.spiky) continue;
            const double w = pr.in_h2 ? poly6_w(pr.r2) : 0.0;
            const double qc = (double)lam[i] + (double)lam[j] + scorr(w);
            const double phi = spiky * (h - pr.r) * (h - pr.r) / pr.r;
            double AG = 0.0;
            for (int a = 0; a < 3; ++a) AG += Ai[a] * phi * pr.d[a];
            double HA[3];
            spiky_hess(spiky, pr.d, pr.r, Ai, HA);
            double ds[3] = {0.0, 0.0, 0.0};
            if (pr.in_h2) {
              double P[3];
              poly6_grad(pr.d, pr.r2, P);
              const double k1 = dscorr_dw(w);
              for (int a = 0; a < 3; ++a) ds[a] = k1 * P[a];
            }
            bl[i] += AG;
            bl[j] += AG;
            for (int a = 0; a < 3; ++a) {
              const double t = qc * HA[a] + AG * ds[a];
              bx[3 * i + a] += t;
              bx[3 * j + a] -= t;
            }
          }
        }
      }
      sx.add_into(gxn);
      sl.add_into(glam);
    }
    // lambda = -(rho/pho0 - 1) / D (fluids/Simulator_kernel.cuh:119-123)
    std::vector<double> grho(n), gD(n);
    for (int i = 0; i < n; ++i) {
      const double D = (double)tp.D[k][i], rho = (double)tp.rho[k][i];
      const double C = rho / pho0 - 1.0;
      const double gC = -glam[i] / D;
      gD[i] = glam[i] * C / (D * D);
      grho[i] = gC / pho0 + (k == NI - 1 ? grho_x[i] : 0.0);
      g_rho0 += gC * (-rho / (pho0 * pho0)) + gD[i] * (-2.0 * (D - eps) / pho0);
      if (kb != 0.0) {
        const T* xi = &x[(size_t)3 * i];
        for (int a = 0; a < 3; ++a) {
          const double s = -wall_slope((double)c.ul[a] - (double)xi[a], h, (double)c.wall_c) +
                           wall_slope((double)xi[a] - (double)c.ll[a], h, (double)c.wall_c);
          gxn[(size_t)3 * i + a] += grho[i] * kb * s;
        }
      }
    }
    // density and gradient-norm pair terms (fluids/Simulator_kernel.cuh:78-104)
    {
      Scatter sx(TH, 3 * n);
#pragma omp parallel num_threads(TH)
      {
        std::vector<uint32_t> cand;
        double* bx = sx.mine();
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < n; ++i) {
          const T* xi = &x[(size_t)3 * i];
          collect(xi, cand);
          double Si[3];
          for (int a = 0; a < 3; ++a) Si[a] = (double)tp.S[k][(size_t)3 * i + a];
          for (uint32_t j : cand) {
            const Pair<T> pr = pair_of(xi, &x[(size_t)3 * j], c);
            if (pr.in_h2 && grho[i] != 0.0) {
              double P[3];
              poly6_grad(pr.d, pr.r2, P);
              for (int a = 0; a < 3; ++a) {
                bx[3 * i + a] += grho[i] * P[a];
                bx[3 * j + a] -= grho[i] * P[a];
              }
            }
            if (pr.spiky && j != (uint32_t)i) {
              const double phig = spiky_rho0 * (h - pr.r) * (h - pr.r) / pr.r;
              double Vv[3];
              for (int a = 0; a < 3; ++a) Vv[a] = phig * pr.d[a] + Si[a];
              double HV[3];
              spiky_hess(spiky_rho0, pr.d, pr.r, Vv, HV);
              for (int a = 0; a < 3; ++a) {
                const double t = 2.0 * gD[i] * HV[a];
                bx[3 * i + a] += t;
                bx[3 * j + a] -= t;
              }
            }
          }
        }
      }
      sx.add_into(gxn);
    }
    gx.swap(gxn);
  }
  // ---- predict: x[0] = X + dt (V + dt g), un-permuted to input order
  gXin.assign((size_t)3 * n, 0.0);
  gVin.assign((size_t)3 * n, 0.0);
  for (int s = 0; s < n; ++s) {
    const uint32_t src = tp.perm[s];
    for (int a = 0; a < 3; ++a) {
      gXin[(size_t)3 * src + a] = gXs[(size_t)3 * s + a] + gx[(size_t)3 * s + a];
      gVin[(size_t)3 * src + a] = dt * gx[(size_t)3 * s + a];
    }
  }
}

template <class T>
int rollout_impl(const orc_params* p, const orc_box* box, const orc_box* boxes, int n,
                 const double* pos0, const double* vel0, const uint32_t* iid0, int nframes,
                 double pho0, int threads, const double* target, double* loss, double* x_final,
                 double* g_pos0, double* g_vel0, double* g_rho0) {
  if (!p || !box || n < 1 || nframes < 0 || !pos0 || !vel0 || !iid0 || !finite_params(p))
    return -1;
  const bool want_grad = g_pos0 || g_vel0 || g_rho0;
  if ((want_grad || loss) && !target) return -1;
  for (int s = 0; s < n; ++s)
    if (iid0[s] >= (uint32_t)n) return -1;
  const int TH = nthreads(threads);
  // per frame: its box, grid dims and constants
  std::vector<orc_box> bx((size_t)nframes, *box);
  if (boxes)
    for (int f = 0; f < nframes; ++f) bx[(size_t)f] = boxes[f];
  std::vector<std::array<int, 3>> dims((size_t)nframes);
  std::vector<Consts<T>> cs;
  for (int f = 0; f < nframes; ++f) {
    if (orc_grid_dims(&bx[(size_t)f], p->h, dims[(size_t)f].data()) != 0) return -1;
    cs.push_back(make_consts<T>(p, &bx[(size_t)f]));
    if (pho0 > 0.0) consts_with_pho0(cs.back(), p, pho0);
  }
  std::vector<T> X((size_t)3 * n), V((size_t)3 * n);
  for (size_t q = 0; q < (size_t)3 * n; ++q) {
    X[q] = (T)pos0[q];
    V[q] = (T)vel0[q];
  }
  std::vector<uint32_t> iid(iid0, iid0 + n);
  std::vector<FrameTape<T>> tapes(want_grad ? nframes : 1);
  std::vector<T> Xo, Vo;
  std::vector<uint32_t> io;
  for (int f = 0; f < nframes; ++f) {
    FrameTape<T>& tp = tapes[want_grad ? f : 0];
    frame_forward(p, &bx[(size_t)f], cs[(size_t)f], dims[(size_t)f].data(), TH, X, V, iid, tp, Xo,
                  Vo, io);
    X.swap(Xo);
    V.swap(Vo);
    iid.swap(io);
  }
  std::vector<double> gX((size_t)3 * n, 0.0), gV((size_t)3 * n, 0.0);
  double L = 0.0;
  for (int s = 0; s < n; ++s)
    for (int a = 0; a < 3; ++a) {
      const double xv = (double)X[(size_t)3 * s + a];
      if (x_final) x_final[(size_t)3 * iid[s] + a] = xv;
      if (target) {
        const double dlt = xv - target[(size_t)3 * iid[s] + a];
        L += dlt * dlt;
        gX[(size_t)3 * s + a] = 2.0 * dlt / n;
      }
    }
  if (loss) *loss = L / n;
  if (!want_grad) return 0;
  double gr = 0.0;
  std::vector<double> gXin, gVin;
  for (int f = nframes - 1; f >= 0; --f) {
    frame_backward(p, &bx[(size_t)f], cs[(size_t)f], dims[(size_t)f].data(), TH, tapes[f], gX, gV,
                   gXin, gVin, gr);
    gX.swap(gXin);
    gV.swap(gVin);
  }
  for (size_t q = 0; q < (size_t)3 * n; ++q) {
    if (g_pos0) g_pos0[q] = gX[q];
    if (g_vel0) g_vel0[q] = gV[q];
  }
  if (g_rho0) *g_rho0 = gr;
  return 0;
}

}  // namespace

extern "C" {

// Simulator::buildGridHash grid dims, fluids/Simulator.cu:187-188:
// (int)ceilf(diff / h) with diff = ulim - llim, all FP32.
int orc_grid_dims(const orc_box* box, float h, int dims[3]) {
  if (!box || !dims || !(h > 0.0f)) return -1;
  for (int a = 0; a < 3; ++a) {
    const float diff = box->ulim[a] - box->llim[a];
    const float q = std::ceil(diff / h);
    if (!(diff > 0.0f) || !(q >= 1.0f) || q > 1073741824.0f) return -1;
    dims[a] = (int)q;
  }
  return 0;
}

// getGridId, fluids/Simulator.cu:56-73.
void orc_cell_keys(const float* x, int stride, int n, const orc_box* box, float h,
                   const int dims[3], uint32_t* keys) {
  for (int i = 0; i < n; ++i) {
    const float* p = x + (size_t)i * stride;
    const int cx = cell_coord_f32(p[0], box->llim[0], h, dims[0]);
    const int cy = cell_coord_f32(p[1], box->llim[1], h, dims[1]);
    const int cz = cell_coord_f32(p[2], box->llim[2], h, dims[2]);
    keys[i] = (uint32_t)(((long long)cx * dims[1] + cy) * dims[2] + cz);
  }
}

// thrust::sort_by_key on uint keys (fluids/Simulator.cu:196-198) is a stable
// radix sort; std::stable_sort defines the same order.
void orc_stable_order(const uint32_t* keys, int n, uint32_t* perm) {
  std::iota(perm, perm + n, 0u);
  std::stable_sort(perm, perm + n, [keys](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
}

// cudaMemset + computeGridRange (fluids/Simulator.cu:201-204,
// fluids/Simulator_kernel.cuh:21-50), as a linear scan.
void orc_cell_ranges(const uint32_t* sk, int n, long long cells, uint32_t* start, uint32_t* end) {
  std::fill(start, start + cells, 0u);
  std::fill(end, end + cells, 0u);
  for (int i = 0; i < n; ++i) {
    if (i == 0 || sk[i] != sk[i - 1]) {
      start[sk[i]] = (uint32_t)i;
      if (i > 0) end[sk[i - 1]] = (uint32_t)i;
    }
    if (i == n - 1) end[sk[i]] = (uint32_t)n;
  }
}

// getPoly6, fluids/Simulator.cu:75-90, in FP64: 315/(64 pi h^9)(h^2-r^2)^3.
double orc_poly6(double h, double r2) {
  const double h2 = h * h;
  if (r2 >= h2) return 0.0;
  const double d = h2 - r2;
  return 315.0 / (64.0 * kPi * std::pow(h, 9)) * d * d * d;
}

// getSpikyGrad, fluids/Simulator.cu:92-107, in FP64:
// -45/(pi h^6) (h-|r|)^2 r/|r| for KERNAL_EPS <= |r| < h, else 0.
void orc_spiky_grad(double h, const double r[3], double grad[3]) {
  const double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  if (rl >= h || rl < kKernelEps) {
    grad[0] = grad[1] = grad[2] = 0.0;
    return;
  }
  const double d = h - rl;
  const double s = -45.0 / (kPi * std::pow(h, 6)) * d * d / rl;
  grad[0] = s * r[0];
  grad[1] = s * r[1];
  grad[2] = s * r[2];
}

// DensityBoundary, fluids/Simulator.cu:139-161, in FP64.
double orc_boundary_density(const orc_box* box, double h, const double p[3]) {
  Consts<double> c{};
  c.h = h;
  c.wall_c = 2.0 * kPi / 3.0;
  for (int a = 0; a < 3; ++a) {
    c.ul[a] = box->ulim[a];
    c.ll[a] = box->llim[a];
  }
  return boundary_density(p, c);
}

int orc_step(const orc_params* p, const orc_box* box, int n, float* pos, float* npos, float* vel,
             float* nvel, uint32_t* iid, const orc_options* opt, orc_diag* diag) {
  return step_impl<double>(p, box, n, pos, npos, vel, nvel, iid, opt, diag);
}

int orc_step_f32(const orc_params* p, const orc_box* box, int n, float* pos, float* npos,
                 float* vel, float* nvel, uint32_t* iid, const orc_options* opt, orc_diag* diag) {
  return step_impl<float>(p, box, n, pos, npos, vel, nvel, iid, opt, diag);
}

int orc_density_error(const orc_params* p, const orc_box* box, int n, const float* x, int stride,
                      int threads, double out[2], double* err_out) {
  return density_impl<double>(p, box, n, x, stride, threads, out, err_out);
}

int orc_density_error_f32(const orc_params* p, const orc_box* box, int n, const float* x,
                          int stride, int threads, double out[2], double* err_out) {
  return density_impl<float>(p, box, n, x, stride, threads, out, err_out);
}

int orc_max_threads(void) { return nthreads(0); }

int orc_rollout(const orc_params* p, const orc_box* box, const orc_box* boxes, int n,
                const double* pos0, const double* vel0, const uint32_t* iid0, int nframes,
                double pho0, int fp32, int threads, const double* target, double* loss,
                double* x_final, double* g_pos0, double* g_vel0, double* g_rho0) {
  return fp32 ? rollout_impl<float>(p, box, boxes, n, pos0, vel0, iid0, nframes, pho0, threads,
                                    target, loss, x_final, g_pos0, g_vel0, g_rho0)
              : rollout_impl<double>(p, box, boxes, n, pos0, vel0, iid0, nframes, pho0, threads,
                                     target, loss, x_final, g_pos0, g_vel0, g_rho0);
}

}  // extern "C"
