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

#include "../../include/eigen_amd.h"
#include "internal.hpp"
#include "kernels.hpp"

namespace eh {
namespace {

struct Temp {
  void* ptr = nullptr;
  hipStream_t s = nullptr;
  eh_status get(size_t bytes, hipStream_t stream) {
    s = stream;
    hipError_t e = hipMallocAsync(&ptr, std::max<size_t>(bytes, 256), stream);
    if (e != hipSuccess) {
      ptr = nullptr;
      return fail(EH_ERR_ALLOC, std::string("complex product temporary: ") + hipGetErrorString(e));
    }
    return EH_SUCCESS;
  }
  ~Temp() {
    if (ptr) (void)hipFreeAsync(ptr, s);
  }
};

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
  Temp t;
  const size_t a3 = (size_t)4 * m * k, b3 = pack_b ? (size_t)2 * k * n : 0;
  eh_status st = t.get((a3 + b3) * sizeof(T), s);
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

}  // namespace
}  // namespace eh

using namespace eh;

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
