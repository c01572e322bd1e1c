// complex_kernels.hip -- repacking kernels of the complex product.
//
// A complex product C = alpha*op(A)*op(B) + beta*C runs as ONE real product
// on the MFMA kernels.  Each complex number a = x + iy acts on a column
// vector (Re, Im) as the 2x2 real block [x -y; y x], so with
//   A3 (2m x 2k real):  rows 2i, 2i+1 and columns 2k', 2k'+1 hold the block
//                       of alpha*op(A)(i, k'),
//   B3 (2k x n real):   rows 2k', 2k'+1 of column j hold Re, Im of op(B)(k', j),
//   C3 (2m x n real):   rows 2i, 2i+1 of column j are Re, Im of C(i, j),
// C3 = A3*B3 + beta*C3 is the complex product with alpha folded in and a real
// beta.  C3 is C itself viewed as a real matrix (leading dimension 2*ldc), and
// so is B3 when op(B) = B (leading dimension 2*ldb): only A, and B when
// transposed or conjugated, are repacked.  This is the 4M scheme (four real
// products per complex one, the same 8mnk flops) fused into a single large
// product; Eigen's complex GEBP instead keeps separate real and imaginary
// accumulators per micro-panel (GeneralBlockPanelKernel.h:631-1050).
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace eh {

template <typename T>
struct Cplx {
  T re, im;
};

// Threads walk the contiguous index of op(A) (coalesced reads).
template <typename T>
__global__ void __launch_bounds__(256) complex_pack_a_kernel(const T* A, int64_t lda, int m, int k, int ta, int conj,
                                                             T ar, T ai, T* A3, int64_t ld3) {
  const int64_t total = (int64_t)m * k;
  const int contig = ta ? k : m;
  for (int64_t idx = blockIdx.x * 256ll + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
    const int64_t c = idx % contig, o = idx / contig;
    const int64_t i = ta ? o : c, kk = ta ? c : o;  // op(A)(i, kk)
    const Cplx<T> a = reinterpret_cast<const Cplx<T>*>(A)[ta ? kk + i * lda : i + kk * lda];
    const T y = conj ? -a.im : a.im;
    const T re = ar * a.re - ai * y, im = ar * y + ai * a.re;  // alpha * op(A)(i, kk)
    T* col0 = A3 + 2 * i + (2 * kk) * ld3;
    col0[0] = re;
    col0[1] = im;
    col0[ld3] = -im;
    col0[ld3 + 1] = re;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) complex_pack_b_kernel(const T* B, int64_t ldb, int k, int n, int tb, int conj,
                                                             T* B3, int64_t ld3) {
  const int64_t total = (int64_t)k * n;
  const int contig = tb ? n : k;
  for (int64_t idx = blockIdx.x * 256ll + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
    const int64_t c = idx % contig, o = idx / contig;
    const int64_t kk = tb ? o : c, j = tb ? c : o;  // op(B)(kk, j)
    const Cplx<T> b = reinterpret_cast<const Cplx<T>*>(B)[tb ? j + kk * ldb : kk + j * ldb];
    T* dst = B3 + 2 * kk + j * ld3;
    dst[0] = b.re;
    dst[1] = conj ? -b.im : b.im;
  }
}

// C = beta*C on m x n complex elements; beta == 0 writes zeros without reading.
template <typename T>
__global__ void __launch_bounds__(256) complex_scale_kernel(T* C, int m, int n, int64_t ldc, T br, T bi) {
  const int64_t total = (int64_t)m * n;
  for (int64_t idx = blockIdx.x * 256ll + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
    const int64_t i = idx % m, j = idx / m;
    Cplx<T>* p = reinterpret_cast<Cplx<T>*>(C) + i + j * ldc;
    if (br == T(0) && bi == T(0)) {
      *p = Cplx<T>{T(0), T(0)};
    } else {
      const Cplx<T> c = *p;
      *p = Cplx<T>{br * c.re - bi * c.im, br * c.im + bi * c.re};
    }
  }
}

// Complex matrix-vector product y = alpha*op(A)*x + beta*y on the real view R
// of A (2m x n real, leading dimension 2*lda: rows 2i, 2i+1 are Re, Im of
// row i).  gemv_gather builds the two real right-hand sides X (len x 2,
// column-major, ld = len2) from x (element i at x[i*incx], already based at
// the element BLAS calls first):
//   op = N (len = n, len2 = n):  X = [Re x, Im x]               P = R X   (2m x 2)
//   op = T (len = m, len2 = 2m): X(2i) = (Re, Im), X(2i+1) = (-Im, Re)
//   op = C:                      X(2i) = (Re, Im), X(2i+1) = (Im, -Re)  P = R^T X (n x 2)
// and gemv_combine forms y from P:
//   op = N:  y_i = (P(2i,0) - P(2i+1,1)) + i (P(2i,1) + P(2i+1,0))
//   op = T/C: y_j = P(j,0) + i P(j,1)
// scaled by alpha, plus beta*y (beta == 0: y not read).
template <typename T>
__global__ void __launch_bounds__(256) gemv_gather_kernel(const T* x, int64_t incx, int len, int op, T* X,
                                                          int64_t ld) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < len; i += (int64_t)gridDim.x * 256) {
    const Cplx<T> v = reinterpret_cast<const Cplx<T>*>(x)[i * incx];
    if (op == 0) {
      X[i] = v.re;
      X[i + ld] = v.im;
    } else {
      X[2 * i] = v.re;
      X[2 * i + ld] = v.im;
      X[2 * i + 1] = op == 2 ? v.im : -v.im;
      X[2 * i + 1 + ld] = op == 2 ? -v.re : v.re;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) gemv_combine_kernel(const T* P, int64_t ldp, int len, int op, T ar, T ai,
                                                           T br, T bi, T* y, int64_t incy) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < len; i += (int64_t)gridDim.x * 256) {
    T pr, pi;
    if (op == 0) {
      pr = P[2 * i] - P[2 * i + 1 + ldp];
      pi = P[2 * i + ldp] + P[2 * i + 1];
    } else {
      pr = P[i];
      pi = P[i + ldp];
    }
    Cplx<T>* dst = reinterpret_cast<Cplx<T>*>(y) + i * incy;
    T re = ar * pr - ai * pi, im = ar * pi + ai * pr;
    if (br != T(0) || bi != T(0)) {
      const Cplx<T> c = *dst;
      re += br * c.re - bi * c.im;
      im += br * c.im + bi * c.re;
    }
    *dst = Cplx<T>{re, im};
  }
}

static unsigned grid_for(int64_t total) {
  int64_t b = (total + 255) / 256;
  if (b > 65536) b = 65536;
  return (unsigned)(b < 1 ? 1 : b);
}

template <typename T>
hipError_t launch_complex_pack_a(const T* A, int64_t lda, int m, int k, bool ta, bool conj, T ar, T ai, T* A3,
                                 int64_t ld3, hipStream_t s) {
  hipLaunchKernelGGL(complex_pack_a_kernel<T>, dim3(grid_for((int64_t)m * k)), dim3(256), 0, s, A, lda, m, k, (int)ta,
                     (int)conj, ar, ai, A3, ld3);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_complex_pack_b(const T* B, int64_t ldb, int k, int n, bool tb, bool conj, T* B3, int64_t ld3,
                                 hipStream_t s) {
  hipLaunchKernelGGL(complex_pack_b_kernel<T>, dim3(grid_for((int64_t)k * n)), dim3(256), 0, s, B, ldb, k, n, (int)tb,
                     (int)conj, B3, ld3);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_complex_scale(T* C, int m, int n, int64_t ldc, T br, T bi, hipStream_t s) {
  hipLaunchKernelGGL(complex_scale_kernel<T>, dim3(grid_for((int64_t)m * n)), dim3(256), 0, s, C, m, n, ldc, br, bi);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_complex_gemv_gather(const T* x, int64_t incx, int len, int op, T* X, int64_t ld, hipStream_t s) {
  hipLaunchKernelGGL(gemv_gather_kernel<T>, dim3(grid_for(len)), dim3(256), 0, s, x, incx, len, op, X, ld);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_complex_gemv_combine(const T* P, int64_t ldp, int len, int op, T ar, T ai, T br, T bi, T* y,
                                       int64_t incy, hipStream_t s) {
  hipLaunchKernelGGL(gemv_combine_kernel<T>, dim3(grid_for(len)), dim3(256), 0, s, P, ldp, len, op, ar, ai, br, bi, y,
                     incy);
  return hipGetLastError();
}

template hipError_t launch_complex_gemv_gather<float>(const float*, int64_t, int, int, float*, int64_t, hipStream_t);
template hipError_t launch_complex_gemv_gather<double>(const double*, int64_t, int, int, double*, int64_t,
                                                       hipStream_t);
template hipError_t launch_complex_gemv_combine<float>(const float*, int64_t, int, int, float, float, float, float,
                                                       float*, int64_t, hipStream_t);
template hipError_t launch_complex_gemv_combine<double>(const double*, int64_t, int, int, double, double, double,
                                                        double, double*, int64_t, hipStream_t);
template hipError_t launch_complex_pack_a<float>(const float*, int64_t, int, int, bool, bool, float, float, float*,
                                                 int64_t, hipStream_t);
template hipError_t launch_complex_pack_a<double>(const double*, int64_t, int, int, bool, bool, double, double,
                                                  double*, int64_t, hipStream_t);
template hipError_t launch_complex_pack_b<float>(const float*, int64_t, int, int, bool, bool, float*, int64_t,
                                                 hipStream_t);
template hipError_t launch_complex_pack_b<double>(const double*, int64_t, int, int, bool, bool, double*, int64_t,
                                                  hipStream_t);
template hipError_t launch_complex_scale<float>(float*, int, int, int64_t, float, float, hipStream_t);
template hipError_t launch_complex_scale<double>(double*, int, int, int64_t, double, double, hipStream_t);

}  // namespace eh
