// level3.cpp -- GEMMTR, SYR2K, SYMM, TRMM and TRSM on the device product.
//
// Eigen builds these on its GEBP engine with routine-specific packing
// (GeneralMatrixMatrixTriangular.h, SelfadjointMatrixMatrix.h,
// TriangularMatrixMatrix.h, TriangularSolverMatrix.h).  Here every routine
// is expressed as products on the MFMA kernels plus two small kernels
// (level3_kernels.hip):
//   GEMMTR, SYR2K  one or two products restricted to a triangle of C;
//   SYMM           the mirrored dense matrix (one HBM pass), then one product;
//   TRMM, TRSM     recursive 2x2 blocking of the triangle: at each level the
//                  off-diagonal block is a plain product on the rest of B;
//                  the leaves are products too (TRMM: the dense diagonal
//                  block; TRSM: the inverse of the diagonal block, all of
//                  them computed by one batched launch), so every flop but
//                  the O(order * 64^2) of the inversions runs on MFMA.
// Arguments are checked with the reference's rules (blas/level3_impl.h); the
// BLAS wrappers in blas.cpp add the xerbla_ codes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/eigen_amd.h"
#include "internal.hpp"
#include "kernels.hpp"

namespace eh {
namespace {

bool valid_op(int op) { return op == EH_OP_N || op == EH_OP_T || op == EH_OP_C; }
bool valid_dtype(int d) { return d == EH_F32 || d == EH_F64; }


#define EH_TRY(x)                         \
  do {                                    \
    eh_status st_ = (x);                  \
    if (st_ != EH_SUCCESS) return st_;    \
  } while (0)

// The triangular operator E = op(A) of order ka: E(i, k) = ta ? A[k + i*lda]
// : A[i + k*lda]; `lower` is the triangle of E (the stored triangle of A
// when not transposed, the other one when transposed).
template <typename T>
struct Tri {
  const T* A;
  int64_t lda;
  bool ta, lower, unit;
  // E[r0.., c0..] as a product operand: pointer and op code
  const T* block(int64_t r0, int64_t c0) const { return ta ? A + c0 + r0 * lda : A + r0 + c0 * lda; }
  int op() const { return ta ? EH_OP_T : EH_OP_N; }
  const T* diag(int64_t d) const { return A + d + d * lda; }
  bool stored_lower() const { return ta ? !lower : lower; }
};

// split point of a diagonal range: about half, on a multiple of the leaf
int64_t half(int64_t sz, int64_t leaf) {
  int64_t h = (sz / 2 + leaf - 1) / leaf * leaf;
  return std::min(h, sz - 1);
}

// ---- TRSM --------------------------------------------------------------------
// The diagonal blocks of order kTrsmLeaf are inverted up front (one launch);
// a leaf of the recursion multiplies its right-hand sides by the inverse, on
// the MFMA kernels, from a copy of them.
template <typename T>
struct TrsmWork {
  const T* inv;  // inverse of diagonal block b at inv + b*kTrsmLeaf^2 (ld kTrsmLeaf)
  T* copy;       // kTrsmLeaf x n (left) or m x kTrsmLeaf (right)
};

// E[d:d+sz, d:d+sz] X = alpha*B, B = rows d.. of the right-hand sides (sz x n).
template <typename T>
eh_status trsm_left(const Tri<T>& E, int64_t d, int64_t sz, T* B, int64_t ldb, int64_t n, T alpha,
                    const TrsmWork<T>& w, hipStream_t s) {
  if (sz <= kTrsmLeaf) {  // X = alpha * inv(E_dd) * B
    hipError_t e = hipMemcpy2DAsync(w.copy, sz * sizeof(T), B, ldb * sizeof(T), sz * sizeof(T), n,
                                    hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "trsm block copy");
    return device_product<T>(EH_OP_N, EH_OP_N, sz, n, sz, alpha, w.inv + (d / kTrsmLeaf) * kTrsmLeaf * kTrsmLeaf,
                             kTrsmLeaf, w.copy, sz, T(0), B, ldb, s);
  }
  const int64_t h = half(sz, kTrsmLeaf);
  T* B1 = B;
  T* B2 = B + h;
  if (E.lower) {  // X1 = E11 \ alpha*B1; B2 = alpha*B2 - E21*X1; X2 = E22 \ B2
    EH_TRY(trsm_left(E, d, h, B1, ldb, n, alpha, w, s));
    EH_TRY(device_product<T>(E.op(), EH_OP_N, sz - h, n, h, T(-1), E.block(d + h, d), E.lda, B1, ldb, alpha, B2, ldb,
                             s));
    return trsm_left(E, d + h, sz - h, B2, ldb, n, T(1), w, s);
  }
  // X2 = E22 \ alpha*B2; B1 = alpha*B1 - E12*X2; X1 = E11 \ B1
  EH_TRY(trsm_left(E, d + h, sz - h, B2, ldb, n, alpha, w, s));
  EH_TRY(device_product<T>(E.op(), EH_OP_N, h, n, sz - h, T(-1), E.block(d, d + h), E.lda, B2, ldb, alpha, B1, ldb,
                           s));
  return trsm_left(E, d, h, B1, ldb, n, T(1), w, s);
}

// X E[d:d+sz, d:d+sz] = alpha*B, B = columns d.. of the right-hand sides (m x sz).
template <typename T>
eh_status trsm_right(const Tri<T>& E, int64_t d, int64_t sz, T* B, int64_t ldb, int64_t m, T alpha,
                     const TrsmWork<T>& w, hipStream_t s) {
  if (sz <= kTrsmLeaf) {  // X = alpha * B * inv(E_dd)
    hipError_t e = hipMemcpy2DAsync(w.copy, m * sizeof(T), B, ldb * sizeof(T), m * sizeof(T), sz,
                                    hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "trsm block copy");
    return device_product<T>(EH_OP_N, EH_OP_N, m, sz, sz, alpha, w.copy, m,
                             w.inv + (d / kTrsmLeaf) * kTrsmLeaf * kTrsmLeaf, kTrsmLeaf, T(0), B, ldb, s);
  }
  const int64_t h = half(sz, kTrsmLeaf);
  T* B1 = B;
  T* B2 = B + h * ldb;
  if (E.lower) {  // X2 E22 = alpha*B2; X1 E11 = alpha*B1 - X2 E21
    EH_TRY(trsm_right(E, d + h, sz - h, B2, ldb, m, alpha, w, s));
    EH_TRY(device_product<T>(EH_OP_N, E.op(), m, h, sz - h, T(-1), B2, ldb, E.block(d + h, d), E.lda, alpha, B1, ldb,
                             s));
    return trsm_right(E, d, h, B1, ldb, m, T(1), w, s);
  }
  // X1 E11 = alpha*B1; X2 E22 = alpha*B2 - X1 E12
  EH_TRY(trsm_right(E, d, h, B1, ldb, m, alpha, w, s));
  EH_TRY(device_product<T>(EH_OP_N, E.op(), m, sz - h, h, T(-1), B1, ldb, E.block(d, d + h), E.lda, alpha, B2, ldb,
                           s));
  return trsm_right(E, d + h, sz - h, B2, ldb, m, T(1), w, s);
}

// ---- TRMM --------------------------------------------------------------------
constexpr int64_t kTrmmLeaf = 256;

// Workspace of the TRMM leaves: the dense diagonal block and a copy of the
// part of B it multiplies.
template <typename T>
struct TrmmWork {
  T* dense;  // kTrmmLeaf x kTrmmLeaf
  T* copy;   // leaf rows x n (left) or m x leaf columns (right)
};

template <typename T>
eh_status dense_diag(const Tri<T>& E, int64_t d, int64_t sz, T* D, hipStream_t s) {
  hipError_t e = launch_tri_dense<T>(E.diag(d), E.lda, (int)sz, D, sz, E.stored_lower(), false, E.unit, s);
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "triangle expansion launch");
}

// B = alpha * E[d:d+sz, d:d+sz] * B, B = rows d.. (sz x n)
template <typename T>
eh_status trmm_left(const Tri<T>& E, int64_t d, int64_t sz, T* B, int64_t ldb, int64_t n, T alpha,
                    const TrmmWork<T>& w, hipStream_t s) {
  if (sz <= kTrmmLeaf) {
    EH_TRY(dense_diag(E, d, sz, w.dense, s));
    hipError_t e = hipMemcpy2DAsync(w.copy, sz * sizeof(T), B, ldb * sizeof(T), sz * sizeof(T), n,
                                    hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "trmm block copy");
    return device_product<T>(E.op(), EH_OP_N, sz, n, sz, alpha, w.dense, sz, w.copy, sz, T(0), B, ldb, s);
  }
  const int64_t h = half(sz, kTrmmLeaf);
  T* B1 = B;
  T* B2 = B + h;
  if (E.lower) {  // B2 = alpha*(E22*B2 + E21*B1) before B1 changes
    EH_TRY(trmm_left(E, d + h, sz - h, B2, ldb, n, alpha, w, s));
    EH_TRY(device_product<T>(E.op(), EH_OP_N, sz - h, n, h, alpha, E.block(d + h, d), E.lda, B1, ldb, T(1), B2, ldb,
                             s));
    return trmm_left(E, d, h, B1, ldb, n, alpha, w, s);
  }
  // B1 = alpha*(E11*B1 + E12*B2) before B2 changes
  EH_TRY(trmm_left(E, d, h, B1, ldb, n, alpha, w, s));
  EH_TRY(device_product<T>(E.op(), EH_OP_N, h, n, sz - h, alpha, E.block(d, d + h), E.lda, B2, ldb, T(1), B1, ldb,
                           s));
  return trmm_left(E, d + h, sz - h, B2, ldb, n, alpha, w, s);
}

// B = alpha * B * E[d:d+sz, d:d+sz], B = columns d.. (m x sz)
template <typename T>
eh_status trmm_right(const Tri<T>& E, int64_t d, int64_t sz, T* B, int64_t ldb, int64_t m, T alpha,
                     const TrmmWork<T>& w, hipStream_t s) {
  if (sz <= kTrmmLeaf) {
    EH_TRY(dense_diag(E, d, sz, w.dense, s));
    hipError_t e = hipMemcpy2DAsync(w.copy, m * sizeof(T), B, ldb * sizeof(T), m * sizeof(T), sz,
                                    hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "trmm block copy");
    return device_product<T>(EH_OP_N, E.op(), m, sz, sz, alpha, w.copy, m, w.dense, sz, T(0), B, ldb, s);
  }
  const int64_t h = half(sz, kTrmmLeaf);
  T* B1 = B;
  T* B2 = B + h * ldb;
  if (E.lower) {  // B1 = alpha*(B1*E11 + B2*E21) before B2 changes
    EH_TRY(trmm_right(E, d, h, B1, ldb, m, alpha, w, s));
    EH_TRY(device_product<T>(EH_OP_N, E.op(), m, h, sz - h, alpha, B2, ldb, E.block(d + h, d), E.lda, T(1), B1, ldb,
                             s));
    return trmm_right(E, d + h, sz - h, B2, ldb, m, alpha, w, s);
  }
  // B2 = alpha*(B1*E12 + B2*E22) before B1 changes
  EH_TRY(trmm_right(E, d + h, sz - h, B2, ldb, m, alpha, w, s));
  EH_TRY(device_product<T>(EH_OP_N, E.op(), m, sz - h, h, alpha, B1, ldb, E.block(d, d + h), E.lda, T(1), B2, ldb,
                           s));
  return trmm_right(E, d, h, B1, ldb, m, alpha, w, s);
}

eh_status check_tri_args(int dtype, int side, int uplo, int transa, int diag, int64_t m, int64_t n, const void* alpha,
                         const void* A, int64_t lda, const void* B, int64_t ldb) {
  if (!valid_dtype(dtype)) return fail(EH_ERR_INVALID, "dtype must be EH_F32 or EH_F64");
  if (side != EH_LEFT && side != EH_RIGHT) return fail(EH_ERR_INVALID, "side must be EH_LEFT or EH_RIGHT");
  if (uplo != EH_UPPER && uplo != EH_LOWER) return fail(EH_ERR_INVALID, "uplo must be EH_UPPER or EH_LOWER");
  if (!valid_op(transa)) return fail(EH_ERR_INVALID, "transa must be N, T or C");
  if (diag != EH_NON_UNIT && diag != EH_UNIT) return fail(EH_ERR_INVALID, "diag must be EH_NON_UNIT or EH_UNIT");
  if (m < 0 || n < 0) return fail(EH_ERR_INVALID, "m, n must be >= 0");
  if (m > INT32_MAX || n > INT32_MAX) return fail(EH_ERR_UNSUPPORTED, "m, n must fit int32");
  if (lda < std::max<int64_t>(1, side == EH_LEFT ? m : n)) return fail(EH_ERR_INVALID, "lda too small");
  if (ldb < std::max<int64_t>(1, m)) return fail(EH_ERR_INVALID, "ldb too small");
  if (!alpha) return fail(EH_ERR_INVALID, "alpha must point to a host scalar");
  if (m > 0 && n > 0 && !B) return fail(EH_ERR_INVALID, "B is NULL");
  (void)A;  // checked once alpha is known: alpha == 0 does not read A
  return EH_SUCCESS;
}

template <typename T>
eh_status zero_b(T* B, int64_t m, int64_t n, int64_t ldb, hipStream_t s) {
  hipError_t e = launch_scale_c<T>(B, (int)m, (int)n, ldb, T(0), s);
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "zero fill launch");
}

template <typename T>
eh_status trsm_t(int side, int uplo, int transa, int diag, int64_t m, int64_t n, T alpha, const T* A, int64_t lda,
                 T* B, int64_t ldb, hipStream_t s) {
  if (alpha == T(0)) return zero_b(B, m, n, ldb, s);
  if (!A) return fail(EH_ERR_INVALID, "A is NULL");
  const bool ta = transa != EH_OP_N;
  const Tri<T> E{A, lda, ta, (uplo == EH_LOWER) != ta, diag == EH_UNIT};
  const int64_t ka = side == EH_LEFT ? m : n;
  const int64_t blocks = (ka + kTrsmLeaf - 1) / kTrsmLeaf;
  const int64_t inv = blocks * kTrsmLeaf * kTrsmLeaf;
  const int64_t copy = side == EH_LEFT ? kTrsmLeaf * n : m * kTrsmLeaf;
  StreamTemp t;
  EH_TRY(t.get((size_t)(inv + copy) * sizeof(T), s, "level-3 temporary"));
  const TrsmWork<T> w{static_cast<T*>(t.ptr), static_cast<T*>(t.ptr) + inv};
  const TrtriParams<T> tp{A, lda, (int)E.ta, (int)E.lower, (int)E.unit, (int)ka, static_cast<T*>(t.ptr)};
  hipError_t e = launch_trtri<T>(tp, s);
  if (e != hipSuccess) return hip_fail(e, "diagonal block inversion launch");
  return side == EH_LEFT ? trsm_left<T>(E, 0, m, B, ldb, n, alpha, w, s)
                         : trsm_right<T>(E, 0, n, B, ldb, m, alpha, w, s);
}

template <typename T>
eh_status trmm_t(int side, int uplo, int transa, int diag, int64_t m, int64_t n, T alpha, const T* A, int64_t lda,
                 T* B, int64_t ldb, hipStream_t s) {
  if (alpha == T(0)) return zero_b(B, m, n, ldb, s);
  if (!A) return fail(EH_ERR_INVALID, "A is NULL");
  const bool ta = transa != EH_OP_N;
  const Tri<T> E{A, lda, ta, (uplo == EH_LOWER) != ta, diag == EH_UNIT};
  const int64_t ka = side == EH_LEFT ? m : n;
  const int64_t leaf = std::min(ka, kTrmmLeaf);
  const int64_t copy = side == EH_LEFT ? leaf * n : m * leaf;
  StreamTemp t;
  EH_TRY(t.get((size_t)(leaf * leaf + copy) * sizeof(T), s, "level-3 temporary"));
  const TrmmWork<T> w{static_cast<T*>(t.ptr), static_cast<T*>(t.ptr) + leaf * leaf};
  return side == EH_LEFT ? trmm_left<T>(E, 0, m, B, ldb, n, alpha, w, s)
                         : trmm_right<T>(E, 0, n, B, ldb, m, alpha, w, s);
}

template <typename T>
eh_status symm_t(int side, int uplo, int64_t m, int64_t n, T alpha, const T* A, int64_t lda, const T* B, int64_t ldb,
                 T beta, T* C, int64_t ldc, hipStream_t s) {
  const int64_t ka = side == EH_LEFT ? m : n;
  if (alpha == T(0))  // only beta*C (A and B are not read)
    return device_product<T>(EH_OP_N, EH_OP_N, m, n, 0, alpha, A, lda, B, ldb, beta, C, ldc, s);
  if (!A || !B) return fail(EH_ERR_INVALID, "A or B is NULL");
  StreamTemp t;
  EH_TRY(t.get((size_t)ka * ka * sizeof(T), s, "level-3 temporary"));
  T* D = static_cast<T*>(t.ptr);
  hipError_t e = launch_tri_dense<T>(A, lda, (int)ka, D, ka, uplo == EH_LOWER, true, false, s);
  if (e != hipSuccess) return hip_fail(e, "symmetric expansion launch");
  if (side == EH_LEFT) return device_product<T>(EH_OP_N, EH_OP_N, m, n, m, alpha, D, ka, B, ldb, beta, C, ldc, s);
  return device_product<T>(EH_OP_N, EH_OP_N, m, n, n, alpha, B, ldb, D, ka, beta, C, ldc, s);
}

template <typename T>
T scalar(const void* p) {
  return *static_cast<const T*>(p);
}

}  // namespace
}  // namespace eh

using namespace eh;

extern "C" {

eh_status eh_gemmtr(eh_dtype dtype, eh_uplo uplo, eh_op opa, eh_op opb, int64_t n, int64_t k, const void* alpha,
                    const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc,
                    void* stream) {
  if (uplo != EH_UPPER && uplo != EH_LOWER) return fail(EH_ERR_INVALID, "uplo must be EH_UPPER or EH_LOWER");
  EH_TRY(check_gemm_args(dtype, opa, opb, n, n, k, alpha, A, lda, B, ldb, beta, C, ldc));
  const int tri = uplo == EH_LOWER ? 1 : 2;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == EH_F64)
    return device_product<double>(opa, opb, n, n, k, scalar<double>(alpha), static_cast<const double*>(A), lda,
                                  static_cast<const double*>(B), ldb, scalar<double>(beta), static_cast<double*>(C),
                                  ldc, s, tri);
  return device_product<float>(opa, opb, n, n, k, scalar<float>(alpha), static_cast<const float*>(A), lda,
                               static_cast<const float*>(B), ldb, scalar<float>(beta), static_cast<float*>(C), ldc, s,
                               tri);
}

eh_status eh_syr2k(eh_dtype dtype, eh_uplo uplo, eh_op trans, int64_t n, int64_t k, const void* alpha, const void* A,
                   int64_t lda, const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc, void* stream) {
  if (!valid_dtype(dtype)) return fail(EH_ERR_INVALID, "dtype must be EH_F32 or EH_F64");
  if (uplo != EH_UPPER && uplo != EH_LOWER) return fail(EH_ERR_INVALID, "uplo must be EH_UPPER or EH_LOWER");
  if (!valid_op(trans)) return fail(EH_ERR_INVALID, "trans must be N, T or C");
  if (n < 0 || k < 0) return fail(EH_ERR_INVALID, "n, k must be >= 0");
  if (n > INT32_MAX || k > INT32_MAX) return fail(EH_ERR_UNSUPPORTED, "n, k must fit int32");
  const int64_t rows = trans == EH_OP_N ? n : k;
  if (lda < std::max<int64_t>(1, rows)) return fail(EH_ERR_INVALID, "lda too small");
  if (ldb < std::max<int64_t>(1, rows)) return fail(EH_ERR_INVALID, "ldb too small");
  if (ldc < std::max<int64_t>(1, n)) return fail(EH_ERR_INVALID, "ldc too small");
  if (!alpha || !beta) return fail(EH_ERR_INVALID, "alpha and beta must point to host scalars");
  if (n == 0) return EH_SUCCESS;
  if (!C || (k > 0 && (!A || !B))) return fail(EH_ERR_INVALID, "A, B or C is NULL");
  const int tri = uplo == EH_LOWER ? 1 : 2;
  const bool tr = trans != EH_OP_N;
  const int o1 = tr ? EH_OP_T : EH_OP_N, o2 = tr ? EH_OP_N : EH_OP_T;
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto run = [&](auto zero) -> eh_status {
    using T = decltype(zero);
    const T al = scalar<T>(alpha), be = scalar<T>(beta);
    const T* Ap = static_cast<const T*>(A);
    const T* Bp = static_cast<const T*>(B);
    T* Cp = static_cast<T*>(C);
    // C = alpha*op(A)*op(B)^T + beta*C, then C += alpha*op(B)*op(A)^T
    EH_TRY(device_product<T>(o1, o2, n, n, k, al, Ap, lda, Bp, ldb, be, Cp, ldc, s, tri));
    return device_product<T>(o1, o2, n, n, k, al, Bp, ldb, Ap, lda, T(1), Cp, ldc, s, tri);
  };
  return dtype == EH_F64 ? run(0.0) : run(0.0f);
}

eh_status eh_symm(eh_dtype dtype, eh_side side, eh_uplo uplo, int64_t m, int64_t n, const void* alpha, const void* A,
                  int64_t lda, const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc, void* stream) {
  if (!valid_dtype(dtype)) return fail(EH_ERR_INVALID, "dtype must be EH_F32 or EH_F64");
  if (side != EH_LEFT && side != EH_RIGHT) return fail(EH_ERR_INVALID, "side must be EH_LEFT or EH_RIGHT");
  if (uplo != EH_UPPER && uplo != EH_LOWER) return fail(EH_ERR_INVALID, "uplo must be EH_UPPER or EH_LOWER");
  if (m < 0 || n < 0) return fail(EH_ERR_INVALID, "m, n must be >= 0");
  if (m > INT32_MAX || n > INT32_MAX) return fail(EH_ERR_UNSUPPORTED, "m, n must fit int32");
  if (lda < std::max<int64_t>(1, side == EH_LEFT ? m : n)) return fail(EH_ERR_INVALID, "lda too small");
  if (ldb < std::max<int64_t>(1, m)) return fail(EH_ERR_INVALID, "ldb too small");
  if (ldc < std::max<int64_t>(1, m)) return fail(EH_ERR_INVALID, "ldc too small");
  if (!alpha || !beta) return fail(EH_ERR_INVALID, "alpha and beta must point to host scalars");
  if (m == 0 || n == 0) return EH_SUCCESS;
  if (!C) return fail(EH_ERR_INVALID, "C is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == EH_F64)
    return symm_t<double>(side, uplo, m, n, scalar<double>(alpha), static_cast<const double*>(A), lda,
                          static_cast<const double*>(B), ldb, scalar<double>(beta), static_cast<double*>(C), ldc, s);
  return symm_t<float>(side, uplo, m, n, scalar<float>(alpha), static_cast<const float*>(A), lda,
                       static_cast<const float*>(B), ldb, scalar<float>(beta), static_cast<float*>(C), ldc, s);
}

eh_status eh_trmm(eh_dtype dtype, eh_side side, eh_uplo uplo, eh_op transa, eh_diag diag, int64_t m, int64_t n,
                  const void* alpha, const void* A, int64_t lda, void* B, int64_t ldb, void* stream) {
  EH_TRY(check_tri_args(dtype, side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb));
  if (m == 0 || n == 0) return EH_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == EH_F64)
    return trmm_t<double>(side, uplo, transa, diag, m, n, scalar<double>(alpha), static_cast<const double*>(A), lda,
                          static_cast<double*>(B), ldb, s);
  return trmm_t<float>(side, uplo, transa, diag, m, n, scalar<float>(alpha), static_cast<const float*>(A), lda,
                       static_cast<float*>(B), ldb, s);
}

eh_status eh_trsm(eh_dtype dtype, eh_side side, eh_uplo uplo, eh_op transa, eh_diag diag, int64_t m, int64_t n,
                  const void* alpha, const void* A, int64_t lda, void* B, int64_t ldb, void* stream) {
  EH_TRY(check_tri_args(dtype, side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb));
  if (m == 0 || n == 0) return EH_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == EH_F64)
    return trsm_t<double>(side, uplo, transa, diag, m, n, scalar<double>(alpha), static_cast<const double*>(A), lda,
                          static_cast<double*>(B), ldb, s);
  return trsm_t<float>(side, uplo, transa, diag, m, n, scalar<float>(alpha), static_cast<const float*>(A), lda,
                       static_cast<float*>(B), ldb, s);
}

}  // extern "C"
