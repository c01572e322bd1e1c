// eigen_ref.cpp -- thin C entry points over Eigen's own CPU product path.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// unmodified Eigen headers of the reference tree into oracle/_ref/, used by
// tests/ as the "real reference" the C restatement (oracle.c) is pinned
// against, and by the benchmark's CPU-baseline leg as the timed Eigen GEBP.
//
// The product runs exactly as a user's code would run it:
//   C.noalias() += alpha * op(A) * op(B)   on column-major host Maps
// (Eigen/src/Core/NoAlias.h:48-53 -> GeneralMatrixMatrix.h:459-514 ->
//  general_matrix_matrix_product::run -> gemm_pack_lhs/rhs + gebp_kernel),
// threaded by Eigen's own OpenMP GEMM parallelizer (Parallelizer.h:172-278).
#include <Eigen/Core>
#include <cstdint>
#include <omp.h>

namespace {

template <typename T>
using ColMap = Eigen::Map<Eigen::Matrix<T, Eigen::Dynamic, Eigen::Dynamic, Eigen::ColMajor>, 0, Eigen::OuterStride<>>;
template <typename T>
using ConstColMap =
    Eigen::Map<const Eigen::Matrix<T, Eigen::Dynamic, Eigen::Dynamic, Eigen::ColMajor>, 0, Eigen::OuterStride<>>;

// C.noalias() += alpha * op(A) * op(B); op(X) = X or X^T as stored column-major.
template <typename T>
void noalias_addmul(int ta, int tb, long m, long n, long k, T alpha, const T* A, long lda, const T* B, long ldb, T* C,
                    long ldc) {
  ColMap<T> c(C, m, n, Eigen::OuterStride<>(ldc));
  ConstColMap<T> a(A, ta ? k : m, ta ? m : k, Eigen::OuterStride<>(lda));
  ConstColMap<T> b(B, tb ? n : k, tb ? k : n, Eigen::OuterStride<>(ldb));
  if (!ta && !tb)
    c.noalias() += alpha * a * b;
  else if (ta && !tb)
    c.noalias() += alpha * a.transpose() * b;
  else if (!ta && tb)
    c.noalias() += alpha * a * b.transpose();
  else
    c.noalias() += alpha * a.transpose() * b.transpose();
}

// dgemm_-style semantics expressed through Eigen's expression API:
// beta==0 overwrites, beta!=1 scales, then the noalias product accumulates
// (blas/level3_impl.h:80-95).
template <typename T>
void gemm_beta(int ta, int tb, long m, long n, long k, T alpha, const T* A, long lda, const T* B, long ldb, T beta,
               T* C, long ldc) {
  if (m == 0 || n == 0) return;
  ColMap<T> c(C, m, n, Eigen::OuterStride<>(ldc));
  if (beta != T(1)) {
    if (beta == T(0))
      c.setZero();
    else
      c *= beta;
  }
  if (k == 0) return;
  noalias_addmul<T>(ta, tb, m, n, k, alpha, A, lda, B, ldb, C, ldc);
}

}  // namespace

extern "C" {

__attribute__((visibility("default"))) int ref_max_threads() { return omp_get_max_threads(); }
__attribute__((visibility("default"))) void ref_set_threads(int n) { Eigen::setNbThreads(n); }
__attribute__((visibility("default"))) int ref_get_threads() { return Eigen::nbThreads(); }
__attribute__((visibility("default"))) const char* ref_simd() { return Eigen::SimdInstructionSetsInUse(); }

__attribute__((visibility("default"))) void ref_gemm_f32(int ta, int tb, long m, long n, long k, float alpha,
                                                         const float* A, long lda, const float* B, long ldb, float* C,
                                                         long ldc) {
  noalias_addmul<float>(ta, tb, m, n, k, alpha, A, lda, B, ldb, C, ldc);
}

__attribute__((visibility("default"))) void ref_gemm_f64(int ta, int tb, long m, long n, long k, double alpha,
                                                         const double* A, long lda, const double* B, long ldb,
                                                         double* C, long ldc) {
  noalias_addmul<double>(ta, tb, m, n, k, alpha, A, lda, B, ldb, C, ldc);
}

__attribute__((visibility("default"))) void ref_gemm_beta_f32(int ta, int tb, long m, long n, long k, float alpha,
                                                              const float* A, long lda, const float* B, long ldb,
                                                              float beta, float* C, long ldc) {
  gemm_beta<float>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

__attribute__((visibility("default"))) void ref_gemm_beta_f64(int ta, int tb, long m, long n, long k, double alpha,
                                                              const double* A, long lda, const double* B, long ldb,
                                                              double beta, double* C, long ldc) {
  gemm_beta<double>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

// Many independent products (the batched workload on the CPU):
// dtype[p] 0=float 1=double; each product C_p = A_p * B_p (noalias, beta=0),
// column-major, lda=m, ldb=k, ldc=m.  `parallel` != 0 runs an OpenMP dynamic
// loop over products with Eigen's own threading switched off inside each
// product; 0 runs them in sequence with Eigen's internal threading.
__attribute__((visibility("default"))) void ref_grouped(int64_t count, const int32_t* dtype, const int32_t* ms,
                                                        const int32_t* ns, const int32_t* ks, void* const* As,
                                                        void* const* Bs, void* const* Cs, int parallel) {
  auto one = [&](int64_t p) {
    long m = ms[p], n = ns[p], k = ks[p];
    if (dtype[p] == 0) {
      ColMap<float> c(static_cast<float*>(Cs[p]), m, n, Eigen::OuterStride<>(m));
      ConstColMap<float> a(static_cast<const float*>(As[p]), m, k, Eigen::OuterStride<>(m));
      ConstColMap<float> b(static_cast<const float*>(Bs[p]), k, n, Eigen::OuterStride<>(k));
      c.noalias() = a * b;
    } else {
      ColMap<double> c(static_cast<double*>(Cs[p]), m, n, Eigen::OuterStride<>(m));
      ConstColMap<double> a(static_cast<const double*>(As[p]), m, k, Eigen::OuterStride<>(m));
      ConstColMap<double> b(static_cast<const double*>(Bs[p]), k, n, Eigen::OuterStride<>(k));
      c.noalias() = a * b;
    }
  };
  if (parallel) {
    int saved = Eigen::nbThreads();
    Eigen::setNbThreads(1);
#pragma omp parallel for schedule(dynamic)
    for (int64_t p = 0; p < count; ++p) one(p);
    Eigen::setNbThreads(saved);
  } else {
    for (int64_t p = 0; p < count; ++p) one(p);
  }
}

// Eigen's own float -> bfloat16 conversion (RTNE) and widening.
__attribute__((visibility("default"))) uint16_t ref_bf16_rtne(float f) {
  return Eigen::numext::bit_cast<uint16_t>(Eigen::bfloat16(f));
}
__attribute__((visibility("default"))) void ref_bf16_rtne_n(const float* x, int64_t n, uint16_t* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = Eigen::numext::bit_cast<uint16_t>(Eigen::bfloat16(x[i]));
}

// Eigen's own float -> half conversion (Eigen::half(float), float_to_half_rtne).
__attribute__((visibility("default"))) void ref_half_rtne_n(const float* x, int64_t n, uint16_t* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = Eigen::numext::bit_cast<uint16_t>(Eigen::half(x[i]));
}

// Eigen's own half product C = A*B (Matrix<half>, GEBP on half scalars) on
// host column-major arrays of half bits: the reference for what Eigen's CPU
// product of half matrices returns.
__attribute__((visibility("default"))) void ref_gemm_half(long m, long n, long k, const uint16_t* A, const uint16_t* B,
                                                          uint16_t* C) {
  using HM = Eigen::Matrix<Eigen::half, Eigen::Dynamic, Eigen::Dynamic>;
  Eigen::Map<const HM> a(reinterpret_cast<const Eigen::half*>(A), m, k);
  Eigen::Map<const HM> b(reinterpret_cast<const Eigen::half*>(B), k, n);
  Eigen::Map<HM> c(reinterpret_cast<Eigen::half*>(C), m, n);
  c.noalias() = a * b;
}

// Eigen's blocking choice for (m,n,k) on this host (informational).
__attribute__((visibility("default"))) void ref_blocking_f32(long m, long n, long k, long* kc, long* mc, long* nc) {
  Eigen::Index K = k, M = m, N = n;
  Eigen::internal::computeProductBlockingSizes<float, float, 1, Eigen::Index>(K, M, N, 1);
  *kc = K;
  *mc = M;
  *nc = N;
}
}
