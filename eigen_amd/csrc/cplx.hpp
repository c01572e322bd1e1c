// cplx.hpp -- complex scalar shared by host code and kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace eh {

// Complex scalar with interleaved (re, im) storage, the layout of
// std::complex<R>, for the complex level-3 kernels.  Host and device.
template <typename R>
struct cplx {
  R re, im;
  cplx() = default;
  __host__ __device__ cplx(R r) : re(r), im(0) {}
  __host__ __device__ cplx(R r, R i) : re(r), im(i) {}
};
template <typename R>
__host__ __device__ inline cplx<R> operator+(cplx<R> a, cplx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R>
__host__ __device__ inline cplx<R> operator-(cplx<R> a, cplx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R>
__host__ __device__ inline cplx<R> operator-(cplx<R> a) { return {-a.re, -a.im}; }
template <typename R>
__host__ __device__ inline cplx<R> operator*(cplx<R> a, cplx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R>
__host__ __device__ inline bool operator==(cplx<R> a, cplx<R> b) { return a.re == b.re && a.im == b.im; }
// 1 / d by Smith's scaling (no overflow for |d| near the range limits)
template <typename R>
__host__ __device__ inline cplx<R> recip(cplx<R> d) {
  if ((d.re < 0 ? -d.re : d.re) >= (d.im < 0 ? -d.im : d.im)) {
    const R r = d.im / d.re, den = d.re + d.im * r;
    return {R(1) / den, -r / den};
  }
  const R r = d.re / d.im, den = d.re * r + d.im;
  return {r / den, R(-1) / den};
}
__host__ __device__ inline float recip(float d) { return 1.0f / d; }
__host__ __device__ inline double recip(double d) { return 1.0 / d; }
template <typename R>
__host__ __device__ inline cplx<R> conj_if(cplx<R> a, bool c) { return c ? cplx<R>{a.re, -a.im} : a; }
__host__ __device__ inline float conj_if(float a, bool) { return a; }
__host__ __device__ inline double conj_if(double a, bool) { return a; }
template <typename R>
__host__ __device__ inline cplx<R> real_part(cplx<R> a) { return {a.re, R(0)}; }
__host__ __device__ inline float real_part(float a) { return a; }
__host__ __device__ inline double real_part(double a) { return a; }
}  // namespace eh
