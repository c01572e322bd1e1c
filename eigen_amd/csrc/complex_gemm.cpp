// complex_gemm.cpp -- eh_gemm_complex: complex products on the real MFMA
// kernels (the real-block form described in complex_kernels.hip).
//
// Host steps: argument checks with ?gemm_'s rules (blas/level3_impl.h:61-78,
// complex elements), the quick returns (m == 0 or n == 0; alpha == 0 or
// k == 0 leaves beta*C), a complex beta applied by a scaling pass (the real
// product can only apply a real beta), the repack of op(A) with alpha folded
// in (and of op(B) unless it is B itself), then ONE real product
// (2m) x n x (2k) through the same planner and kernels as eh_gemm, writing C
// in place.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <type_traits>

#include "../../include/eigen_amd.h"
#include "internal.hpp"
#include "kernels.hpp"

namespace eh {
namespace {


template <typename T>
eh_status run(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, const T* alpha, const T* A, int64_t lda,
              const T* B, int64_t ldb, const T* beta, T* C, int64_t ldc, int prec, hipStream_t s) {
  const T ar = alpha[0], ai = alpha[1], br = beta[0], bi = beta[1];
  const bool vanishes = k == 0 || (ar == T(0) && ai == T(0));
  T beff = br;
  if (vanishes || bi != T(0)) {  // C = beta*C as a complex scaling
    if (!(br == T(1) && bi == T(0))) {
      hipError_t e = launch_complex_scale<T>(C, (int)m, (int)n, ldc, br, bi, s);
      if (e != hipSuccess) return hip_fail(e, "complex scale launch");
    }
    if (vanishes) return EH_SUCCESS;
    beff = T(1);
  }
  const bool pack_b = opb != EH_OP_N;
  StreamTemp t;
  const size_t a3 = (size_t)4 * m * k, b3 = pack_b ? (size_t)2 * k * n : 0;
  eh_status st = t.get((a3 + b3) * sizeof(T), s, "complex product temporary");
  if (st != EH_SUCCESS) return st;
  T* A3 = static_cast<T*>(t.ptr);
  hipError_t e = launch_complex_pack_a<T>(A, lda, (int)m, (int)k, opa != EH_OP_N, opa == EH_OP_C, ar, ai, A3, 2 * m, s);
  if (e != hipSuccess) return hip_fail(e, "complex pack launch");
  const T* B3 = B;
  int64_t ldb3 = 2 * ldb;
  if (pack_b) {
    T* P = A3 + a3;
    e = launch_complex_pack_b<T>(B, ldb, (int)k, (int)n, true, opb == EH_OP_C, P, 2 * k, s);
    if (e != hipSuccess) return hip_fail(e, "complex pack launch");
    B3 = P;
    ldb3 = 2 * k;
  }
  const T one = T(1);
  return gemm_dispatch(dtype, EH_OP_N, EH_OP_N, 2 * m, n, 2 * k, &one, A3, 2 * m, B3, ldb3, &beff, C, 2 * ldc, prec, -1,
                       s);
}

// y = alpha*op(A)*x + beta*y, complex, on the real view of A: one real
// product with two right-hand sides (A read once) between a gather and a
// combine pass (complex_kernels.hip).
template <typename T>
eh_status gemv_run(int dtype, int op, int64_t m, int64_t n, const T* alpha, const T* A, int64_t lda, const T* x,
                   int64_t incx, const T* beta, T* y, int64_t incy, hipStream_t s) {
  const int64_t lenx = op == EH_OP_N ? n : m, leny = op == EH_OP_N ? m : n;
  if (alpha[0] == T(0) && alpha[1] == T(0)) {  // y = beta*y
    if (beta[0] == T(1) && beta[1] == T(0)) return EH_SUCCESS;
    hipError_t e = launch_complex_scale<T>(y, 1, (int)leny, incy, beta[0], beta[1], s);
    return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "complex scale launch");
  }
  const int opc = op == EH_OP_N ? 0 : (op == EH_OP_T ? 1 : 2);
  const int64_t ldx = op == EH_OP_N ? n : 2 * m, ldp = op == EH_OP_N ? 2 * m : n;
  StreamTemp t;
  eh_status st = t.get((size_t)(2 * ldx + 2 * ldp) * sizeof(T), s, "complex product temporary");
  if (st != EH_SUCCESS) return st;
  T* X = static_cast<T*>(t.ptr);
  T* P = X + 2 * ldx;
  hipError_t e = launch_complex_gemv_gather<T>(x, incx, (int)lenx, opc, X, ldx, s);
  if (e != hipSuccess) return hip_fail(e, "complex gather launch");
  const T one = T(1), zero = T(0);
  if (op == EH_OP_N)
    st = gemm_dispatch(dtype, EH_OP_N, EH_OP_N, 2 * m, 2, n, &one, A, 2 * lda, X, ldx, &zero, P, ldp, EH_PREC_PLAIN, -1,
                       s);
  else
    st = gemm_dispatch(dtype, EH_OP_T, EH_OP_N, n, 2, 2 * m, &one, A, 2 * lda, X, ldx, &zero, P, ldp, EH_PREC_PLAIN, -1,
                       s);
  if (st != EH_SUCCESS) return st;
  e = launch_complex_gemv_combine<T>(P, ldp, (int)leny, opc, alpha[0], alpha[1], beta[0], beta[1], y, incy, s);
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "complex combine launch");
}

}  // namespace

template <typename R>
eh_status complex_product(int opa, int opb, int64_t m, int64_t n, int64_t k, cplx<R> alpha, const cplx<R>* A,
                          int64_t lda, const cplx<R>* B, int64_t ldb, cplx<R> beta, cplx<R>* C, int64_t ldc,
                          hipStream_t s) {
  if (m == 0 || n == 0) return EH_SUCCESS;
  if (m > INT32_MAX / 2 || k > INT32_MAX / 2) return fail(EH_ERR_UNSUPPORTED, "complex m, k must be < 2^30");
  return run<R>(std::is_same<R, double>::value ? EH_F64 : EH_F32, opa, opb, m, n, k, &alpha.re,
                reinterpret_cast<const R*>(A), lda, reinterpret_cast<const R*>(B), ldb, &beta.re,
                reinterpret_cast<R*>(C), ldc, EH_PREC_PLAIN, s);
}
template eh_status complex_product<float>(int, int, int64_t, int64_t, int64_t, cplx<float>, const cplx<float>*,
                                          int64_t, const cplx<float>*, int64_t, cplx<float>, cplx<float>*, int64_t,
                                          hipStream_t);
template eh_status complex_product<double>(int, int, int64_t, int64_t, int64_t, cplx<double>, const cplx<double>*,
                                           int64_t, const cplx<double>*, int64_t, cplx<double>, cplx<double>*, int64_t,
                                           hipStream_t);

}  // namespace eh

using namespace eh;

extern "C" eh_status eh_gemv_complex(eh_dtype dtype, eh_op trans, int64_t m, int64_t n, const void* alpha,
                                     const void* A, int64_t lda, const void* x, int64_t incx, const void* beta,
                                     void* y, int64_t incy, void* stream) {
  if (dtype != EH_F32 && dtype != EH_F64) return fail(EH_ERR_INVALID, "dtype must be EH_F32 or EH_F64");
  if (trans != EH_OP_N && trans != EH_OP_T && trans != EH_OP_C) return fail(EH_ERR_INVALID, "trans must be N, T or C");
  if (m < 0 || n < 0) return fail(EH_ERR_INVALID, "m, n must be >= 0");
  if (m > INT32_MAX / 2 || n > INT32_MAX / 2) return fail(EH_ERR_UNSUPPORTED, "m, n must be below 2^30");
  if (lda < std::max<int64_t>(1, m)) return fail(EH_ERR_INVALID, "lda too small");
  if (incx == 0 || incy == 0) return fail(EH_ERR_INVALID, "incx and incy must be nonzero");
  if (!alpha || !beta) return fail(EH_ERR_INVALID, "alpha and beta must point to host scalars");
  if (m == 0 || n == 0) return EH_SUCCESS;
  if (!A || !x || !y) return fail(EH_ERR_INVALID, "A, x or y is NULL");
  const int64_t lenx = trans == EH_OP_N ? n : m, leny = trans == EH_OP_N ? m : n;
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto run = [&](auto zero) -> eh_status {
    using T = decltype(zero);
    // a negative increment walks the vector from its last element (BLAS); 2 reals per element
    const T* xp = static_cast<const T*>(x) + 2 * (incx < 0 ? (1 - lenx) * incx : 0);
    T* yp = static_cast<T*>(y) + 2 * (incy < 0 ? (1 - leny) * incy : 0);
    return gemv_run<T>(dtype, trans, m, n, static_cast<const T*>(alpha), static_cast<const T*>(A), lda, xp, incx,
                       static_cast<const T*>(beta), yp, incy, s);
  };
  return dtype == EH_F64 ? run(0.0) : run(0.0f);
}

extern "C" eh_status eh_gemm_complex(eh_dtype dtype, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k,
                                     const void* alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
                                     const void* beta, void* C, int64_t ldc, eh_precision prec, void* stream) {
  eh_status st = check_gemm_args(dtype, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  if (st != EH_SUCCESS) return st;
  if (m > INT32_MAX / 2 || k > INT32_MAX / 2) return fail(EH_ERR_UNSUPPORTED, "complex m, k must be < 2^30");
  if (prec < EH_PREC_AUTO || prec > EH_PREC_SPLIT3) return fail(EH_ERR_INVALID, "bad precision");
  if (m == 0 || n == 0) return EH_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == EH_F64)
    return run<double>(dtype, opa, opb, m, n, k, static_cast<const double*>(alpha), static_cast<const double*>(A), lda,
                       static_cast<const double*>(B), ldb, static_cast<const double*>(beta), static_cast<double*>(C),
                       ldc, prec, s);
  return run<float>(dtype, opa, opb, m, n, k, static_cast<const float*>(alpha), static_cast<const float*>(A), lda,
                    static_cast<const float*>(B), ldb, static_cast<const float*>(beta), static_cast<float*>(C), ldc,
                    prec, s);
}
