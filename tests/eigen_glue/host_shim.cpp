// Host stand-ins for the device runtime and the C-ABI routines that
// include/eigen_amd_eigen.h calls, so the Eigen drop-in programs can be linked
// and run on a machine without a GPU: "device" buffers are host memory, and
// each eh_* routine is the BLAS definition of its operation written with
// Eigen on the host.  Linked in place of libeigen_amd and libamdhip64, they
// check the header's mapping of Eigen expressions onto routine arguments
// (side, triangle, op, leading dimensions, scalars, storage order) against
// Eigen's own evaluation of the same expressions; the device kernels behind
// the routines are checked by the GPU tests.
#include <Eigen/Core>
#include <complex>
#include <cstdlib>
#include <cstring>

#include "../../include/eigen_amd.h"

extern "C" {
// the HIP runtime calls the test programs make (common.hpp)
int hipMalloc(void** p, size_t n) {
  *p = std::malloc(n ? n : 1);
  return *p ? 0 : 2;
}
int hipFree(void* p) {
  std::free(p);
  return 0;
}
int hipMemcpy(void* d, const void* s, size_t n, int) {
  std::memcpy(d, s, n);
  return 0;
}
const char* hipGetErrorString(int) { return "host stand-in"; }
}

namespace {

template <typename T>
using Mat = Eigen::Matrix<T, Eigen::Dynamic, Eigen::Dynamic>;
template <typename T>
using MapM = Eigen::Map<Mat<T>, 0, Eigen::OuterStride<>>;
template <typename T>
using CMapM = Eigen::Map<const Mat<T>, 0, Eigen::OuterStride<>>;

template <typename T>
T sc(const void* p) {
  return *static_cast<const T*>(p);
}

template <typename T>
Mat<T> op_of(const Mat<T>& a, eh_op op) {
  if (op == EH_OP_N) return a;
  if (op == EH_OP_T) return a.transpose();
  return a.adjoint();
}

// stored r x c matrix at p with leading dimension ld
template <typename T>
Mat<T> load(const void* p, int64_t r, int64_t c, int64_t ld) {
  return CMapM<T>(static_cast<const T*>(p), r, c, Eigen::OuterStride<>(ld));
}

// the triangle of the n x n matrix at A: zeros off it, 1s on the diagonal if unit
template <typename T>
Mat<T> tri(const void* A, int64_t n, int64_t lda, eh_uplo uplo, eh_diag diag) {
  Mat<T> a = load<T>(A, n, n, lda), t = Mat<T>::Zero(n, n);
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < n; ++i)
      if (i == j)
        t(i, j) = diag == EH_UNIT ? T(1) : a(i, j);
      else if ((i > j) == (uplo == EH_LOWER))
        t(i, j) = a(i, j);
  return t;
}

// the self-adjoint n x n matrix stored in the uplo triangle (conjugated
// mirror and real diagonal if hermitian)
template <typename T>
Mat<T> selfadjoint(const void* A, int64_t n, int64_t lda, eh_uplo uplo, bool herm) {
  Mat<T> a = load<T>(A, n, n, lda), s(n, n);
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < n; ++i) {
      const bool stored = i == j || (i > j) == (uplo == EH_LOWER);
      const T v = stored ? a(i, j) : a(j, i);
      s(i, j) = (!stored && herm) ? Eigen::numext::conj(v) : v;
      if (i == j && herm) s(i, j) = T(Eigen::numext::real(v));
    }
  return s;
}

// write `full` into the uplo triangle of the n x n C; hermitian results
// get a real diagonal
template <typename T>
void store_triangle(void* C, int64_t n, int64_t ldc, eh_uplo uplo, const Mat<T>& full, bool herm) {
  MapM<T> c(static_cast<T*>(C), n, n, Eigen::OuterStride<>(ldc));
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < n; ++i)
      if (i == j || (i > j) == (uplo == EH_LOWER))
        c(i, j) = (i == j && herm) ? T(Eigen::numext::real(full(i, j))) : full(i, j);
}

template <typename T>
eh_status gemm(eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, const void* alpha, const void* A, int64_t lda,
               const void* B, int64_t ldb, const void* beta, void* C, int64_t incc, int64_t ldc) {
  const Mat<T> a = op_of<T>(load<T>(A, opa == EH_OP_N ? m : k, opa == EH_OP_N ? k : m, lda), opa);
  const Mat<T> b = op_of<T>(load<T>(B, opb == EH_OP_N ? k : n, opb == EH_OP_N ? n : k, ldb), opb);
  const Mat<T> p = sc<T>(alpha) * (a * b);
  T* c = static_cast<T*>(C);
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < m; ++i) {
      T& x = c[i * incc + j * ldc];
      x = sc<T>(beta) == T(0) ? p(i, j) : p(i, j) + sc<T>(beta) * x;
    }
  return EH_SUCCESS;
}

template <typename T>
eh_status trsm(eh_side side, eh_uplo uplo, eh_op transa, eh_diag diag, int64_t m, int64_t n, const void* alpha,
               const void* A, int64_t lda, void* B, int64_t ldb) {
  const int64_t na = side == EH_LEFT ? m : n;
  const Mat<T> t = op_of<T>(tri<T>(A, na, lda, uplo, diag), transa);
  const bool lower = (uplo == EH_LOWER) != (transa != EH_OP_N);
  MapM<T> b(static_cast<T*>(B), m, n, Eigen::OuterStride<>(ldb));
  Mat<T> x = sc<T>(alpha) * b;
  if (side == EH_LEFT) {
    if (lower)
      t.template triangularView<Eigen::Lower>().solveInPlace(x);
    else
      t.template triangularView<Eigen::Upper>().solveInPlace(x);
  } else {
    if (lower)
      t.template triangularView<Eigen::Lower>().template solveInPlace<Eigen::OnTheRight>(x);
    else
      t.template triangularView<Eigen::Upper>().template solveInPlace<Eigen::OnTheRight>(x);
  }
  b = x;
  return EH_SUCCESS;
}

template <typename T>
eh_status trmm(eh_side side, eh_uplo uplo, eh_op transa, eh_diag diag, int64_t m, int64_t n, const void* alpha,
               const void* A, int64_t lda, void* B, int64_t ldb) {
  const int64_t na = side == EH_LEFT ? m : n;
  const Mat<T> t = op_of<T>(tri<T>(A, na, lda, uplo, diag), transa);
  MapM<T> b(static_cast<T*>(B), m, n, Eigen::OuterStride<>(ldb));
  const Mat<T> x = side == EH_LEFT ? Mat<T>(sc<T>(alpha) * (t * b)) : Mat<T>(sc<T>(alpha) * (b * t));
  b = x;
  return EH_SUCCESS;
}

template <typename T>
eh_status symm(eh_side side, eh_uplo uplo, int64_t m, int64_t n, const void* alpha, const void* A, int64_t lda,
               const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc, bool herm) {
  const int64_t na = side == EH_LEFT ? m : n;
  const Mat<T> s = selfadjoint<T>(A, na, lda, uplo, herm);
  const Mat<T> b = load<T>(B, m, n, ldb);
  MapM<T> c(static_cast<T*>(C), m, n, Eigen::OuterStride<>(ldc));
  const Mat<T> p = side == EH_LEFT ? Mat<T>(s * b) : Mat<T>(b * s);
  const Mat<T> r = sc<T>(beta) == T(0) ? Mat<T>(sc<T>(alpha) * p) : Mat<T>(sc<T>(alpha) * p + sc<T>(beta) * c);
  c = r;
  return EH_SUCCESS;
}

// C = alpha * op(A) * op(B) + beta * C on the uplo triangle (op(A) n x k)
template <typename T>
Mat<T> tri_product(int64_t n, int64_t k, const void* A, int64_t lda, eh_op opa, const void* B, int64_t ldb, eh_op opb,
                   const void* C, int64_t ldc, T alpha, T beta) {
  const Mat<T> a = op_of<T>(load<T>(A, opa == EH_OP_N ? n : k, opa == EH_OP_N ? k : n, lda), opa);
  const Mat<T> b = op_of<T>(load<T>(B, opb == EH_OP_N ? k : n, opb == EH_OP_N ? n : k, ldb), opb);
  const Mat<T> c = load<T>(C, n, n, ldc);
  return beta == T(0) ? Mat<T>(alpha * (a * b)) : Mat<T>(alpha * (a * b) + beta * c);
}

template <typename T>
eh_op partner(eh_op trans) {  // op of the second factor of a rank update
  return trans == EH_OP_N ? (Eigen::NumTraits<T>::IsComplex ? EH_OP_C : EH_OP_T) : EH_OP_N;
}

template <typename T>
eh_status rank2(eh_uplo uplo, eh_op trans, int64_t n, int64_t k, T alpha, const void* A, int64_t lda,
                       const void* B, int64_t ldb, T beta, void* C, int64_t ldc, bool herm) {
  const eh_op o1 = trans, o2 = partner<T>(trans);
  const T zero(0);
  const Mat<T> p1 = tri_product<T>(n, k, A, lda, o1, B, ldb, o2, C, ldc, alpha, zero);
  const Mat<T> p2 = tri_product<T>(n, k, B, ldb, o1, A, lda, o2, C, ldc, herm ? Eigen::numext::conj(alpha) : alpha,
                                   beta);
  store_triangle<T>(C, n, ldc, uplo, p1 + p2, herm);
  return EH_SUCCESS;
}

}  // namespace

#define EH_DISPATCH_REAL(dtype, F, ...) \
  ((dtype) == EH_F64 ? F<double>(__VA_ARGS__) : F<float>(__VA_ARGS__))
#define EH_DISPATCH_CPLX(dtype, F, ...) \
  ((dtype) == EH_F64 ? F<std::complex<double>>(__VA_ARGS__) : F<std::complex<float>>(__VA_ARGS__))

extern "C" {

eh_status eh_gemm(eh_dtype dtype, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, const void* alpha,
                  const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc,
                  eh_precision, void*) {
  return EH_DISPATCH_REAL(dtype, gemm, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, 1, ldc);
}
eh_status eh_gemm_strided(eh_dtype dtype, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, const void* alpha,
                          const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C,
                          int64_t incc, int64_t ldc, eh_precision, void*) {
  return EH_DISPATCH_REAL(dtype, gemm, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, incc, ldc);
}
eh_status eh_gemm_complex(eh_dtype dtype, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, const void* alpha,
                          const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C,
                          int64_t ldc, eh_precision, void*) {
  return EH_DISPATCH_CPLX(dtype, gemm, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, 1, ldc);
}
eh_status eh_trsm(eh_dtype dtype, eh_side side, eh_uplo uplo, eh_op transa, eh_diag diag, int64_t m, int64_t n,
                  const void* alpha, const void* A, int64_t lda, void* B, int64_t ldb, void*) {
  return EH_DISPATCH_REAL(dtype, trsm, side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb);
}
eh_status eh_trsm_complex(eh_dtype dtype, eh_side side, eh_uplo uplo, eh_op transa, eh_diag diag, int64_t m,
                          int64_t n, const void* alpha, const void* A, int64_t lda, void* B, int64_t ldb, void*) {
  return EH_DISPATCH_CPLX(dtype, trsm, side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb);
}
eh_status eh_trmm(eh_dtype dtype, eh_side side, eh_uplo uplo, eh_op transa, eh_diag diag, int64_t m, int64_t n,
                  const void* alpha, const void* A, int64_t lda, void* B, int64_t ldb, void*) {
  return EH_DISPATCH_REAL(dtype, trmm, side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb);
}
eh_status eh_trmm_complex(eh_dtype dtype, eh_side side, eh_uplo uplo, eh_op transa, eh_diag diag, int64_t m,
                          int64_t n, const void* alpha, const void* A, int64_t lda, void* B, int64_t ldb, void*) {
  return EH_DISPATCH_CPLX(dtype, trmm, side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb);
}
eh_status eh_symm(eh_dtype dtype, eh_side side, eh_uplo uplo, int64_t m, int64_t n, const void* alpha, const void* A,
                  int64_t lda, const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc, void*) {
  return EH_DISPATCH_REAL(dtype, symm, side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc, false);
}
eh_status eh_hemm(eh_dtype dtype, eh_side side, eh_uplo uplo, int64_t m, int64_t n, const void* alpha, const void* A,
                  int64_t lda, const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc, void*) {
  return EH_DISPATCH_CPLX(dtype, symm, side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc, true);
}

#define EH_TRI_BODY(T, herm)                                                                                  \
  store_triangle<T>(C, n, ldc, uplo,                                                                          \
                    tri_product<T>(n, k, A, lda, opa, B, ldb, opb, C, ldc, sc<T>(alpha), sc<T>(beta)), herm); \
  return EH_SUCCESS

eh_status eh_gemmtr(eh_dtype dtype, eh_uplo uplo, eh_op opa, eh_op opb, int64_t n, int64_t k, const void* alpha,
                    const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc,
                    void*) {
  if (dtype == EH_F64) {
    EH_TRI_BODY(double, false);
  }
  EH_TRI_BODY(float, false);
}
eh_status eh_gemmtr_complex(eh_dtype dtype, eh_uplo uplo, eh_op opa, eh_op opb, int64_t n, int64_t k,
                            const void* alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
                            const void* beta, void* C, int64_t ldc, void*) {
  if (dtype == EH_F64) {
    EH_TRI_BODY(std::complex<double>, false);
  }
  EH_TRI_BODY(std::complex<float>, false);
}

eh_status eh_syrk(eh_dtype dtype, eh_uplo uplo, eh_op trans, int64_t n, int64_t k, const void* alpha, const void* A,
                  int64_t lda, const void* beta, void* C, int64_t ldc, void*) {
  const eh_op opa = trans, opb = trans == EH_OP_N ? EH_OP_T : EH_OP_N;
  const void* B = A;
  const int64_t ldb = lda;
  if (dtype == EH_F64) {
    EH_TRI_BODY(double, false);
  }
  EH_TRI_BODY(float, false);
}
eh_status eh_herk(eh_dtype dtype, eh_uplo uplo, eh_op trans, int64_t n, int64_t k, const void* alpha_r,
                  const void* A, int64_t lda, const void* beta_r, void* C, int64_t ldc, void*) {
  const eh_op opa = trans, opb = trans == EH_OP_N ? EH_OP_C : EH_OP_N;
  const void* B = A;
  const int64_t ldb = lda;
  if (dtype == EH_F64) {
    const std::complex<double> al(sc<double>(alpha_r)), be(sc<double>(beta_r));
    const void *alpha = &al, *beta = &be;
    EH_TRI_BODY(std::complex<double>, true);
  }
  const std::complex<float> al(sc<float>(alpha_r)), be(sc<float>(beta_r));
  const void *alpha = &al, *beta = &be;
  EH_TRI_BODY(std::complex<float>, true);
}

eh_status eh_syr2k(eh_dtype dtype, eh_uplo uplo, eh_op trans, int64_t n, int64_t k, const void* alpha, const void* A,
                   int64_t lda, const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc, void*) {
  if (dtype == EH_F64)
    return rank2<double>(uplo, trans, n, k, sc<double>(alpha), A, lda, B, ldb, sc<double>(beta), C, ldc, false);
  return rank2<float>(uplo, trans, n, k, sc<float>(alpha), A, lda, B, ldb, sc<float>(beta), C, ldc, false);
}
eh_status eh_her2k(eh_dtype dtype, eh_uplo uplo, eh_op trans, int64_t n, int64_t k, const void* alpha, const void* A,
                   int64_t lda, const void* B, int64_t ldb, const void* beta_r, void* C, int64_t ldc, void*) {
  if (dtype == EH_F64)
    return rank2<std::complex<double>>(uplo, trans, n, k, sc<std::complex<double>>(alpha), A, lda, B, ldb,
                                       std::complex<double>(sc<double>(beta_r)), C, ldc, true);
  return rank2<std::complex<float>>(uplo, trans, n, k, sc<std::complex<float>>(alpha), A, lda, B, ldb,
                                    std::complex<float>(sc<float>(beta_r)), C, ldc, true);
}

eh_status eh_synchronize(void*) { return EH_SUCCESS; }
const char* eh_last_error(void) { return ""; }

}  // extern "C"
