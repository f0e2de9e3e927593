// arith_check.hip — a check of the pair terms' range-specialised square root
// and quotient (sqrt_pair, div_pair in pbf_device.h) against the IEEE
// operations (this file is built with -fhip-fp32-correctly-rounded-divide-sqrt,
// so sqrtf and '/' are correctly rounded).  Built into its own library
// (pbf/libpbf_arith_check.so), used by the tests only.
//
// pbf_check_pair_arith(lo_bits, hi_bits, small_hi_bits, keps, div_samples,
// out) runs, on the current device:
//   out[0] = inputs x with bit patterns in [lo_bits, hi_bits) (positive
//            floats, every one of them) where sqrt_pair(x) != sqrtf(x);
//   out[1] = inputs x in [0, small_hi_bits) where sqrt_pair(x) >= keps
//            (below the pair terms' range only "< KERNAL_EPS" matters);
//   out[2] = quotients a / b, div_samples of them with a log-uniform in
//            [1e-18, 1e6] and b log-uniform in [1e-4, 1e4] (both signs of a),
//            where div_pair(a, b) != a / b;
//   out[3] = inputs tested in total.
// Returns 0, or a hipError_t value.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pbf_device.h"

namespace {

__global__ void k_sqrt(uint32_t lo, uint32_t hi, unsigned long long* bad) {
  unsigned long long n = 0;
  for (uint64_t b = (uint64_t)lo + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; b < hi;
       b += (uint64_t)gridDim.x * blockDim.x) {
    const float x = __uint_as_float((uint32_t)b);
    n += __float_as_uint(pbf::sqrt_pair(x)) != __float_as_uint(sqrtf(x)) ? 1u : 0u;
  }
  if (n) atomicAdd(bad, n);
}

__global__ void k_sqrt_small(uint32_t hi, float keps, unsigned long long* bad) {
  unsigned long long n = 0;
  for (uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; b < hi;
       b += (uint64_t)gridDim.x * blockDim.x) {
    const float x = __uint_as_float((uint32_t)b);
    n += pbf::sqrt_pair(x) >= keps ? 1u : 0u;
  }
  if (n) atomicAdd(bad, n);
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ float log_uniform(uint64_t r, float lg_lo, float lg_hi) {
  const float u = (float)(r >> 40) * (1.0f / 16777216.0f);
  return exp2f(lg_lo + u * (lg_hi - lg_lo));
}

__global__ void k_div(uint64_t samples, unsigned long long* bad) {
  unsigned long long n = 0;
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < samples;
       k += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r1 = mix64(0x9E3779B97F4A7C15ull * (2 * k + 1));
    const uint64_t r2 = mix64(0x9E3779B97F4A7C15ull * (2 * k + 2));
    float a = log_uniform(r1, -59.79f, 19.93f);  // 1e-18 .. 1e6
    const float b = log_uniform(r2, -13.29f, 13.29f);  // 1e-4 .. 1e4
    if (r1 & 1) a = -a;
    n += __float_as_uint(pbf::div_pair(a, b)) != __float_as_uint(a / b) ? 1u : 0u;
  }
  if (n) atomicAdd(bad, n);
}

}  // namespace

extern "C" int pbf_check_pair_arith(uint32_t lo_bits, uint32_t hi_bits, uint32_t small_hi_bits,
                                    float keps, unsigned long long div_samples,
                                    unsigned long long out[4]) {
  unsigned long long* d = nullptr;
  hipError_t e = hipMalloc(&d, 3 * sizeof(unsigned long long));
  if (e != hipSuccess) return (int)e;
  e = hipMemset(d, 0, 3 * sizeof(unsigned long long));
  if (e == hipSuccess) {
    k_sqrt<<<4096, 256>>>(lo_bits, hi_bits, d);
    k_sqrt_small<<<4096, 256>>>(small_hi_bits, keps, d + 1);
    k_div<<<4096, 256>>>(div_samples, d + 2);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, d, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return (int)e;
  out[3] = (unsigned long long)(hi_bits - lo_bits) + small_hi_bits + div_samples;
  return 0;
}
