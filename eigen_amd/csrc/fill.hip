// fill.hip -- seeded synthetic inputs generated on the device.
//
// Element i of a stream with seed s is derived from splitmix64 at counter
// (s, 2i) (and (s, 2i+1) where a second variate is needed):
//     z = s + (c+1) * 0x9E3779B97F4A7C15;  z = mix(z)
// so every element is independent and the whole matrix is produced in one
// pass over HBM.  Distributions (computed in double, then rounded once):
//   0 uniform[p0,p1]       v = fma(p1-p0, u, p0),  u = (z >> 11) * 2^-53
//   1 normal(p0, p1)       Box-Muller on (u1 in (0,1], u2)
//   2 sign*U(1,2)*2^U(p0,p1)
//   3 integers in [p0,p1]
#include "common.hpp"
#include "kernels.hpp"

namespace eh {

__device__ inline uint64_t splitmix64_at(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ inline double unit53(uint64_t u) { return (double)(u >> 11) * 0x1.0p-53; }

__device__ inline double gen_value(uint64_t seed, uint64_t i, int dist, double p0, double p1) {
  const uint64_t ua = splitmix64_at(seed, 2 * i);
  if (dist == 0) return __builtin_fma(p1 - p0, unit53(ua), p0);
  if (dist == 1) {
    const uint64_t ub = splitmix64_at(seed, 2 * i + 1);
    const double r1 = ((double)(ua >> 11) + 1.0) * 0x1.0p-53;
    const double r2 = unit53(ub);
    return p0 + p1 * sqrt(-2.0 * log(r1)) * cos(6.283185307179586 * r2);
  }
  if (dist == 2) {
    const uint64_t ub = splitmix64_at(seed, 2 * i + 1);
    const double sgn = (ua & 1) ? -1.0 : 1.0;
    const double mant = 1.0 + unit53(ua);
    const double e = __builtin_fma(p1 - p0, unit53(ub), p0);
    return sgn * mant * exp2(e);
  }
  return p0 + floor(unit53(ua) * (p1 - p0 + 1.0));
}

template <typename T>
__global__ void __launch_bounds__(256) fill_kernel(T* x, int64_t n, uint64_t seed, int dist, double p0, double p1) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    x[i] = (T)gen_value(seed, (uint64_t)i, dist, p0, p1);
}

hipError_t launch_fill(int dtype, void* x, int64_t n, uint64_t seed, int dist, double p0, double p1,
                       hipStream_t s) {
  int64_t blocks = (n + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  if (blocks < 1) blocks = 1;
  if (dtype == 0)
    hipLaunchKernelGGL(fill_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, s, (float*)x, n, seed, dist, p0, p1);
  else
    hipLaunchKernelGGL(fill_kernel<double>, dim3((unsigned)blocks), dim3(256), 0, s, (double*)x, n, seed, dist, p0,
                       p1);
  return hipGetLastError();
}

}  // namespace eh
