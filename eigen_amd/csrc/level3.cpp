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
//
// The recursions are templated on the scalar: float, double, or cplx<R> for
// the complex routines (complex products run as one real product over 2x2
// real blocks, complex_gemm.cpp; op C conjugates).  The complex rank
// updates (SYRK, HERK, SYR2K, HER2K, GEMMTR) form the full product W in a
// temporary and combine its triangle into C (tri_combine, one HBM pass).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <type_traits>

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

// Product on the device, real or complex.
template <typename T>
eh_status prod(int opa, int opb, int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, const T* B,
               int64_t ldb, T beta, T* C, int64_t ldc, hipStream_t s) {
  return device_product<T>(opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
}
template <typename R>
eh_status prod(int opa, int opb, int64_t m, int64_t n, int64_t k, cplx<R> alpha, const cplx<R>* A, int64_t lda,
               const cplx<R>* B, int64_t ldb, cplx<R> beta, cplx<R>* C, int64_t ldc, hipStream_t s) {
  return complex_product<R>(opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

// B = 0 (m x n)
template <typename T>
eh_status zero_b(T* B, int64_t m, int64_t n, int64_t ldb, hipStream_t s) {
  hipError_t e = launch_scale_c<T>(B, (int)m, (int)n, ldb, T(0), s);
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "zero fill launch");
}
template <typename R>
eh_status zero_b(cplx<R>* B, int64_t m, int64_t n, int64_t ldb, hipStream_t s) {
  hipError_t e = launch_complex_scale<R>(reinterpret_cast<R*>(B), (int)m, (int)n, ldb, R(0), R(0), s);
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "zero fill launch");
}

// order of the inverted diagonal blocks of TRSM
template <typename T>
constexpr int64_t trsm_leaf() {
  return kTrsmLeaf;
}
template <>
constexpr int64_t trsm_leaf<cplx<float>>() {
  return kTrsmLeafComplex;
}
template <>
constexpr int64_t trsm_leaf<cplx<double>>() {
  return kTrsmLeafComplex;
}

// The triangular operator E = op(A) of order ka: E(i, k) = ta ? A[k + i*lda]
// : A[i + k*lda]; `lower` is the triangle of E (the stored triangle of A
// when not transposed, the other one when transposed).
template <typename T>
struct Tri {
  const T* A;
  int64_t lda;
  bool ta, lower, unit;
  bool conj;  // E = A^H (complex, op C)
  // E[r0.., c0..] as a product operand: pointer and op code
  const T* block(int64_t r0, int64_t c0) const { return ta ? A + c0 + r0 * lda : A + r0 + c0 * lda; }
  int op() const { return ta ? (conj ? EH_OP_C : EH_OP_T) : EH_OP_N; }
  const T* diag(int64_t d) const { return A + d + d * lda; }
  bool stored_lower() const { return ta ? !lower : lower; }
};

// split point of a diagonal range: about half, on a multiple of the leaf
int64_t half(int64_t sz, int64_t leaf) {
  int64_t h = (sz / 2 + leaf - 1) / leaf * leaf;
  return std::min(h, sz - 1);
}

// ---- TRSM --------------------------------------------------------------------
// The diagonal blocks of order trsm_leaf<T>() are inverted up front (one
// launch), and for real types their order is then doubled up to
// kTrsmLeafMax while it divides the order of the triangle: for every pair of
// consecutive blocks at once (two batched products of the current order),
//   lower:  inv([E11 0; E21 E22]) = [I1 0; -I2 E21 I1, I2]
//   upper:  inv([E11 E12; 0 E22]) = [I1 -I1 E12 I2; 0 I2].
// A leaf of the recursion multiplies its right-hand sides by the inverse of
// its block, on the MFMA kernels, from a copy of them.  Larger leaves halve
// the recursion's smallest and least efficient launches at each doubling
// (a 64-row leaf multiplies 64 x n by 64 x 64).
constexpr int64_t kTrsmLeafMax = 256;

template <typename T>
struct IsCplx : std::false_type {};
template <typename R>
struct IsCplx<cplx<R>> : std::true_type {};

template <typename T>
struct TrsmWork {
  const T* inv;  // inverse of diagonal block b at inv + b*leaf^2 (ld leaf)
  T* copy;       // leaf x n (left) or m x leaf (right)
  int64_t leaf;  // order of the inverted blocks
};

template <typename T>
constexpr eh_dtype real_dtype() {
  return std::is_same<T, double>::value ? EH_F64 : EH_F32;
}

// Inverses of order 2L in D (count = ka / 2L blocks) from those of order L in
// S; tmp holds count blocks of order L.
template <typename T>
eh_status double_inverses(const Tri<T>& E, int64_t ka, int64_t L, const T* S, T* D, T* tmp, hipStream_t s) {
  const int64_t count = ka / (2 * L), L2 = 2 * L;
  hipError_t e = launch_inverse_place<T>(S, D, (int)L, count, s);
  if (e != hipSuccess) return hip_fail(e, "inverse placement launch");
  const int64_t sd = L2 * L2, st = L * L, se = L2 * (1 + E.lda);  // strides between pairs
  const T one(1), zero(0), minus(-1);
  constexpr eh_dtype dt = real_dtype<T>();
  if (E.lower) {  // tmp = E21 I1; D(L:, :L) = -I2 tmp
    EH_TRY(eh_gemm_batched_strided(dt, (eh_op)E.op(), EH_OP_N, L, L, L, &one, E.block(L, 0), E.lda, se, D, L2, sd,
                                   &zero, tmp, L, st, count, s));
    return eh_gemm_batched_strided(dt, EH_OP_N, EH_OP_N, L, L, L, &minus, D + L + L * L2, L2, sd, tmp, L, st, &zero,
                                   D + L, L2, sd, count, s);
  }
  // tmp = E12 I2; D(:L, L:) = -I1 tmp
  EH_TRY(eh_gemm_batched_strided(dt, (eh_op)E.op(), EH_OP_N, L, L, L, &one, E.block(0, L), E.lda, se, D + L + L * L2,
                                 L2, sd, &zero, tmp, L, st, count, s));
  return eh_gemm_batched_strided(dt, EH_OP_N, EH_OP_N, L, L, L, &minus, D, L2, sd, tmp, L, st, &zero, D + L * L2, L2,
                                 sd, count, s);
}

// E[d:d+sz, d:d+sz] X = alpha*B, B = rows d.. of the right-hand sides (sz x n).
template <typename T>
eh_status trsm_left(const Tri<T>& E, int64_t d, int64_t sz, T* B, int64_t ldb, int64_t n, T alpha,
                    const TrsmWork<T>& w, hipStream_t s) {
  const int64_t leaf = w.leaf;
  if (sz <= leaf) {  // X = alpha * inv(E_dd) * B
    hipError_t e = hipMemcpy2DAsync(w.copy, sz * sizeof(T), B, ldb * sizeof(T), sz * sizeof(T), n,
                                    hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "trsm block copy");
    return prod(EH_OP_N, EH_OP_N, sz, n, sz, alpha, w.inv + (d / leaf) * leaf * leaf, leaf, w.copy, sz, T(0), B, ldb,
                s);
  }
  const int64_t h = half(sz, leaf);
  T* B1 = B;
  T* B2 = B + h;
  if (E.lower) {  // X1 = E11 \ alpha*B1; B2 = alpha*B2 - E21*X1; X2 = E22 \ B2
    EH_TRY(trsm_left(E, d, h, B1, ldb, n, alpha, w, s));
    EH_TRY(prod(E.op(), EH_OP_N, sz - h, n, h, T(-1), E.block(d + h, d), E.lda, B1, ldb, alpha, B2, ldb,
                             s));
    return trsm_left(E, d + h, sz - h, B2, ldb, n, T(1), w, s);
  }
  // X2 = E22 \ alpha*B2; B1 = alpha*B1 - E12*X2; X1 = E11 \ B1
  EH_TRY(trsm_left(E, d + h, sz - h, B2, ldb, n, alpha, w, s));
  EH_TRY(prod(E.op(), EH_OP_N, h, n, sz - h, T(-1), E.block(d, d + h), E.lda, B2, ldb, alpha, B1, ldb,
                           s));
  return trsm_left(E, d, h, B1, ldb, n, T(1), w, s);
}

// X E[d:d+sz, d:d+sz] = alpha*B, B = columns d.. of the right-hand sides (m x sz).
template <typename T>
eh_status trsm_right(const Tri<T>& E, int64_t d, int64_t sz, T* B, int64_t ldb, int64_t m, T alpha,
                     const TrsmWork<T>& w, hipStream_t s) {
  const int64_t leaf = w.leaf;
  if (sz <= leaf) {  // X = alpha * B * inv(E_dd)
    hipError_t e = hipMemcpy2DAsync(w.copy, m * sizeof(T), B, ldb * sizeof(T), m * sizeof(T), sz,
                                    hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "trsm block copy");
    return prod(EH_OP_N, EH_OP_N, m, sz, sz, alpha, w.copy, m, w.inv + (d / leaf) * leaf * leaf, leaf, T(0), B, ldb, s);
  }
  const int64_t h = half(sz, leaf);
  T* B1 = B;
  T* B2 = B + h * ldb;
  if (E.lower) {  // X2 E22 = alpha*B2; X1 E11 = alpha*B1 - X2 E21
    EH_TRY(trsm_right(E, d + h, sz - h, B2, ldb, m, alpha, w, s));
    EH_TRY(prod(EH_OP_N, E.op(), m, h, sz - h, T(-1), B2, ldb, E.block(d + h, d), E.lda, alpha, B1, ldb,
                             s));
    return trsm_right(E, d, h, B1, ldb, m, T(1), w, s);
  }
  // X1 E11 = alpha*B1; X2 E22 = alpha*B2 - X1 E12
  EH_TRY(trsm_right(E, d, h, B1, ldb, m, alpha, w, s));
  EH_TRY(prod(EH_OP_N, E.op(), m, sz - h, h, T(-1), B1, ldb, E.block(d, d + h), E.lda, alpha, B2, ldb,
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
    return prod(E.op(), EH_OP_N, sz, n, sz, alpha, w.dense, sz, w.copy, sz, T(0), B, ldb, s);
  }
  const int64_t h = half(sz, kTrmmLeaf);
  T* B1 = B;
  T* B2 = B + h;
  if (E.lower) {  // B2 = alpha*(E22*B2 + E21*B1) before B1 changes
    EH_TRY(trmm_left(E, d + h, sz - h, B2, ldb, n, alpha, w, s));
    EH_TRY(prod(E.op(), EH_OP_N, sz - h, n, h, alpha, E.block(d + h, d), E.lda, B1, ldb, T(1), B2, ldb,
                             s));
    return trmm_left(E, d, h, B1, ldb, n, alpha, w, s);
  }
  // B1 = alpha*(E11*B1 + E12*B2) before B2 changes
  EH_TRY(trmm_left(E, d, h, B1, ldb, n, alpha, w, s));
  EH_TRY(prod(E.op(), EH_OP_N, h, n, sz - h, alpha, E.block(d, d + h), E.lda, B2, ldb, T(1), B1, ldb,
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
    return prod(EH_OP_N, E.op(), m, sz, sz, alpha, w.copy, m, w.dense, sz, T(0), B, ldb, s);
  }
  const int64_t h = half(sz, kTrmmLeaf);
  T* B1 = B;
  T* B2 = B + h * ldb;
  if (E.lower) {  // B1 = alpha*(B1*E11 + B2*E21) before B2 changes
    EH_TRY(trmm_right(E, d, h, B1, ldb, m, alpha, w, s));
    EH_TRY(prod(EH_OP_N, E.op(), m, h, sz - h, alpha, B2, ldb, E.block(d + h, d), E.lda, T(1), B1, ldb,
                             s));
    return trmm_right(E, d + h, sz - h, B2, ldb, m, alpha, w, s);
  }
  // B2 = alpha*(B1*E12 + B2*E22) before B1 changes
  EH_TRY(trmm_right(E, d + h, sz - h, B2, ldb, m, alpha, w, s));
  EH_TRY(prod(EH_OP_N, E.op(), m, sz - h, h, alpha, B1, ldb, E.block(d, d + h), E.lda, T(1), B2, ldb,
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
eh_status trsm_t(int side, int uplo, int transa, int diag, int64_t m, int64_t n, T alpha, const T* A, int64_t lda,
                 T* B, int64_t ldb, hipStream_t s) {
  if (alpha == T(0)) return zero_b(B, m, n, ldb, s);
  if (!A) return fail(EH_ERR_INVALID, "A is NULL");
  const bool ta = transa != EH_OP_N;
  const Tri<T> E{A, lda, ta, (uplo == EH_LOWER) != ta, diag == EH_UNIT, transa == EH_OP_C};
  const int64_t ka = side == EH_LEFT ? m : n;
  constexpr int64_t leaf0 = trsm_leaf<T>();
  // doubled leaf order (real types: the doubling runs on the batched
  // product, which takes real operands)
  int64_t leaf = leaf0;
  if constexpr (!IsCplx<T>::value)
    while (leaf < kTrsmLeafMax && ka % (2 * leaf) == 0) leaf *= 2;
  const int64_t blocks0 = (ka + leaf0 - 1) / leaf0;
  const int64_t inv0 = blocks0 * leaf0 * leaf0;                 // the inversion's blocks
  const int64_t invd = leaf > leaf0 ? ka * leaf : 0;           // each doubled set (ka / L blocks of L^2)
  const int64_t tmp = leaf > leaf0 ? ka * leaf / 4 : 0;        // (ka / 2L) blocks of L^2, L <= leaf / 2
  const int64_t copy = side == EH_LEFT ? leaf * n : m * leaf;
  StreamTemp t;
  EH_TRY(t.get((size_t)(inv0 + 2 * invd + tmp + copy) * sizeof(T), s, "level-3 temporary"));
  T* base = static_cast<T*>(t.ptr);
  T* bufs[2] = {base + inv0, base + inv0 + invd};
  T* tbuf = base + inv0 + 2 * invd;
  TrtriParams<T> tp{A, lda, (int)E.ta, (int)E.lower, (int)E.unit, (int)ka, base};
  tp.conj = E.conj;
  hipError_t e = launch_trtri<T>(tp, s);
  if (e != hipSuccess) return hip_fail(e, "diagonal block inversion launch");
  const T* inv = base;
  int which = 0;
  if constexpr (!IsCplx<T>::value) {
    for (int64_t L = leaf0; L < leaf; L *= 2, which ^= 1) {
      EH_TRY(double_inverses(E, ka, L, inv, bufs[which], tbuf, s));
      inv = bufs[which];
    }
  }
  const TrsmWork<T> w{inv, tbuf + tmp, leaf};
  return side == EH_LEFT ? trsm_left<T>(E, 0, m, B, ldb, n, alpha, w, s)
                         : trsm_right<T>(E, 0, n, B, ldb, m, alpha, w, s);
}

template <typename T>
eh_status trmm_t(int side, int uplo, int transa, int diag, int64_t m, int64_t n, T alpha, const T* A, int64_t lda,
                 T* B, int64_t ldb, hipStream_t s) {
  if (alpha == T(0)) return zero_b(B, m, n, ldb, s);
  if (!A) return fail(EH_ERR_INVALID, "A is NULL");
  const bool ta = transa != EH_OP_N;
  const Tri<T> E{A, lda, ta, (uplo == EH_LOWER) != ta, diag == EH_UNIT, transa == EH_OP_C};
  const int64_t ka = side == EH_LEFT ? m : n;
  const int64_t leaf = std::min(ka, kTrmmLeaf);
  const int64_t copy = side == EH_LEFT ? leaf * n : m * leaf;
  StreamTemp t;
  EH_TRY(t.get((size_t)(leaf * leaf + copy) * sizeof(T), s, "level-3 temporary"));
  const TrmmWork<T> w{static_cast<T*>(t.ptr), static_cast<T*>(t.ptr) + leaf * leaf};
  return side == EH_LEFT ? trmm_left<T>(E, 0, m, B, ldb, n, alpha, w, s)
                         : trmm_right<T>(E, 0, n, B, ldb, m, alpha, w, s);
}

// herm (complex): A is Hermitian (the mirror is conjugated, the diagonal's
// imaginary part is not read), else symmetric.
template <typename T>
eh_status symm_t(int side, int uplo, int64_t m, int64_t n, T alpha, const T* A, int64_t lda, const T* B, int64_t ldb,
                 T beta, T* C, int64_t ldc, hipStream_t s, bool herm = false) {
  const int64_t ka = side == EH_LEFT ? m : n;
  if (alpha == T(0))  // only beta*C (A and B are not read)
    return prod(EH_OP_N, EH_OP_N, m, n, 0, alpha, A, lda, B, ldb, beta, C, ldc, s);
  if (!A || !B) return fail(EH_ERR_INVALID, "A or B is NULL");
  StreamTemp t;
  EH_TRY(t.get((size_t)ka * ka * sizeof(T), s, "level-3 temporary"));
  T* D = static_cast<T*>(t.ptr);
  hipError_t e = launch_tri_dense<T>(A, lda, (int)ka, D, ka, uplo == EH_LOWER, true, false, s, herm);
  if (e != hipSuccess) return hip_fail(e, "symmetric expansion launch");
  if (side == EH_LEFT) return prod(EH_OP_N, EH_OP_N, m, n, m, alpha, D, ka, B, ldb, beta, C, ldc, s);
  return prod(EH_OP_N, EH_OP_N, m, n, n, alpha, B, ldb, D, ka, beta, C, ldc, s);
}

template <typename T>
T scalar(const void* p) {
  return *static_cast<const T*>(p);
}

template <typename R>
cplx<R> cscalar(const void* p) {
  const R* q = static_cast<const R*>(p);
  return {q[0], q[1]};
}

// Complex update of one triangle of C (n x n): W = op1(X)*op2(Y) over the
// whole square in a temporary, then C = alpha*W + alpha2*f(W^T) + beta*C on
// the triangle (tri_combine).  k == 0 or alpha == 0 reads neither X nor Y.
template <typename R>
eh_status rank_update(int uplo, int o1, int o2, int64_t n, int64_t k, cplx<R> alpha, const cplx<R>* X, int64_t ldx,
                      const cplx<R>* Y, int64_t ldy, cplx<R> alpha2, bool use2, bool herm, cplx<R> beta, cplx<R>* C,
                      int64_t ldc, hipStream_t s) {
  const bool lower = uplo == EH_LOWER;
  if (k == 0 || alpha == cplx<R>(0)) {
    if (beta == cplx<R>(1)) return EH_SUCCESS;  // C untouched (as ?herk / ?her2k leave it)
    hipError_t e = launch_tri_combine<R>(nullptr, 0, (int)n, C, ldc, lower, cplx<R>(0), cplx<R>(0), false, beta, herm,
                                         s);
    return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "triangle scale launch");
  }
  if (!X || !Y) return fail(EH_ERR_INVALID, "A or B is NULL");
  StreamTemp t;
  EH_TRY(t.get((size_t)n * n * sizeof(cplx<R>), s, "rank update temporary"));
  cplx<R>* W = static_cast<cplx<R>*>(t.ptr);
  EH_TRY(prod(o1, o2, n, n, k, cplx<R>(1), X, ldx, Y, ldy, cplx<R>(0), W, n, s));
  hipError_t e = launch_tri_combine<R>(W, n, (int)n, C, ldc, lower, alpha, alpha2, use2, beta, herm, s);
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "triangle combine launch");
}

eh_status check_rank_args(int dtype, int uplo, int trans, int64_t n, int64_t k, const void* alpha, const void* A,
                          int64_t lda, const void* B, int64_t ldb, bool needs_b, const void* beta, const void* C,
                          int64_t ldc, int bad_trans) {
  if (!valid_dtype(dtype)) return fail(EH_ERR_INVALID, "dtype must be EH_F32 or EH_F64");
  if (uplo != EH_UPPER && uplo != EH_LOWER) return fail(EH_ERR_INVALID, "uplo must be EH_UPPER or EH_LOWER");
  if (!valid_op(trans) || trans == bad_trans) return fail(EH_ERR_INVALID, "trans not allowed for this routine");
  if (n < 0 || k < 0) return fail(EH_ERR_INVALID, "n, k must be >= 0");
  if (n > INT32_MAX / 2 || k > INT32_MAX / 2) return fail(EH_ERR_UNSUPPORTED, "complex n, k must be below 2^30");
  const int64_t rows = trans == EH_OP_N ? n : k;
  if (lda < std::max<int64_t>(1, rows)) return fail(EH_ERR_INVALID, "lda too small");
  if (needs_b && ldb < std::max<int64_t>(1, rows)) return fail(EH_ERR_INVALID, "ldb too small");
  if (ldc < std::max<int64_t>(1, n)) return fail(EH_ERR_INVALID, "ldc too small");
  if (!alpha || !beta) return fail(EH_ERR_INVALID, "alpha and beta must point to host scalars");
  if (n > 0 && !C) return fail(EH_ERR_INVALID, "C is NULL");
  (void)A;
  (void)B;
  return EH_SUCCESS;
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

// ---- complex routines -----------------------------------------------------

eh_status eh_trsm_complex(eh_dtype dtype, eh_side side, eh_uplo uplo, eh_op transa, eh_diag diag, int64_t m,
                          int64_t n, const void* alpha, const void* A, int64_t lda, void* B, int64_t ldb,
                          void* stream) {
  EH_TRY(check_tri_args(dtype, side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb));
  if (m == 0 || n == 0) return EH_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == EH_F64)
    return trsm_t<cplx<double>>(side, uplo, transa, diag, m, n, cscalar<double>(alpha),
                                static_cast<const cplx<double>*>(A), lda, static_cast<cplx<double>*>(B), ldb, s);
  return trsm_t<cplx<float>>(side, uplo, transa, diag, m, n, cscalar<float>(alpha), static_cast<const cplx<float>*>(A),
                             lda, static_cast<cplx<float>*>(B), ldb, s);
}

eh_status eh_trmm_complex(eh_dtype dtype, eh_side side, eh_uplo uplo, eh_op transa, eh_diag diag, int64_t m,
                          int64_t n, const void* alpha, const void* A, int64_t lda, void* B, int64_t ldb,
                          void* stream) {
  EH_TRY(check_tri_args(dtype, side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb));
  if (m == 0 || n == 0) return EH_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == EH_F64)
    return trmm_t<cplx<double>>(side, uplo, transa, diag, m, n, cscalar<double>(alpha),
                                static_cast<const cplx<double>*>(A), lda, static_cast<cplx<double>*>(B), ldb, s);
  return trmm_t<cplx<float>>(side, uplo, transa, diag, m, n, cscalar<float>(alpha), static_cast<const cplx<float>*>(A),
                             lda, static_cast<cplx<float>*>(B), ldb, s);
}

static eh_status symm_complex(eh_dtype dtype, eh_side side, eh_uplo uplo, int64_t m, int64_t n, const void* alpha,
                              const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C,
                              int64_t ldc, void* stream, bool herm) {
  if (!valid_dtype(dtype)) return fail(EH_ERR_INVALID, "dtype must be EH_F32 or EH_F64");
  if (side != EH_LEFT && side != EH_RIGHT) return fail(EH_ERR_INVALID, "side must be EH_LEFT or EH_RIGHT");
  if (uplo != EH_UPPER && uplo != EH_LOWER) return fail(EH_ERR_INVALID, "uplo must be EH_UPPER or EH_LOWER");
  if (m < 0 || n < 0) return fail(EH_ERR_INVALID, "m, n must be >= 0");
  if (m > INT32_MAX / 2 || n > INT32_MAX / 2) return fail(EH_ERR_UNSUPPORTED, "complex m, n must be below 2^30");
  if (lda < std::max<int64_t>(1, side == EH_LEFT ? m : n)) return fail(EH_ERR_INVALID, "lda too small");
  if (ldb < std::max<int64_t>(1, m)) return fail(EH_ERR_INVALID, "ldb too small");
  if (ldc < std::max<int64_t>(1, m)) return fail(EH_ERR_INVALID, "ldc too small");
  if (!alpha || !beta) return fail(EH_ERR_INVALID, "alpha and beta must point to host scalars");
  if (m == 0 || n == 0) return EH_SUCCESS;
  if (!C) return fail(EH_ERR_INVALID, "C is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == EH_F64)
    return symm_t<cplx<double>>(side, uplo, m, n, cscalar<double>(alpha), static_cast<const cplx<double>*>(A), lda,
                                static_cast<const cplx<double>*>(B), ldb, cscalar<double>(beta),
                                static_cast<cplx<double>*>(C), ldc, s, herm);
  return symm_t<cplx<float>>(side, uplo, m, n, cscalar<float>(alpha), static_cast<const cplx<float>*>(A), lda,
                             static_cast<const cplx<float>*>(B), ldb, cscalar<float>(beta),
                             static_cast<cplx<float>*>(C), ldc, s, herm);
}

eh_status eh_symm_complex(eh_dtype dtype, eh_side side, eh_uplo uplo, int64_t m, int64_t n, const void* alpha,
                          const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C,
                          int64_t ldc, void* stream) {
  return symm_complex(dtype, side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc, stream, false);
}

eh_status eh_hemm(eh_dtype dtype, eh_side side, eh_uplo uplo, int64_t m, int64_t n, const void* alpha, const void* A,
                  int64_t lda, const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc, void* stream) {
  return symm_complex(dtype, side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc, stream, true);
}

// rank updates: run the complex update with the scalar type of `dtype`
#define EH_RANK_DISPATCH(call)                                       \
  do {                                                               \
    hipStream_t s = static_cast<hipStream_t>(stream);                \
    if (dtype == EH_F64) {                                           \
      using R = double;                                              \
      return call;                                                   \
    }                                                                \
    using R = float;                                                 \
    return call;                                                     \
  } while (0)

eh_status eh_syrk_complex(eh_dtype dtype, eh_uplo uplo, eh_op trans, int64_t n, int64_t k, const void* alpha,
                          const void* A, int64_t lda, const void* beta, void* C, int64_t ldc, void* stream) {
  EH_TRY(check_rank_args(dtype, uplo, trans, n, k, alpha, A, lda, nullptr, 0, false, beta, C, ldc, EH_OP_C));
  if (n == 0) return EH_SUCCESS;
  const int o1 = trans == EH_OP_N ? EH_OP_N : EH_OP_T, o2 = trans == EH_OP_N ? EH_OP_T : EH_OP_N;
  EH_RANK_DISPATCH((rank_update<R>(uplo, o1, o2, n, k, cscalar<R>(alpha), static_cast<const cplx<R>*>(A), lda,
                                   static_cast<const cplx<R>*>(A), lda, cplx<R>(0), false, false, cscalar<R>(beta),
                                   static_cast<cplx<R>*>(C), ldc, s)));
}

eh_status eh_herk(eh_dtype dtype, eh_uplo uplo, eh_op trans, int64_t n, int64_t k, const void* alpha, const void* A,
                  int64_t lda, const void* beta, void* C, int64_t ldc, void* stream) {
  EH_TRY(check_rank_args(dtype, uplo, trans, n, k, alpha, A, lda, nullptr, 0, false, beta, C, ldc, EH_OP_T));
  if (n == 0) return EH_SUCCESS;
  const int o1 = trans == EH_OP_N ? EH_OP_N : EH_OP_C, o2 = trans == EH_OP_N ? EH_OP_C : EH_OP_N;
  EH_RANK_DISPATCH((rank_update<R>(uplo, o1, o2, n, k, cplx<R>(scalar<R>(alpha)), static_cast<const cplx<R>*>(A), lda,
                                   static_cast<const cplx<R>*>(A), lda, cplx<R>(0), false, true,
                                   cplx<R>(scalar<R>(beta)), static_cast<cplx<R>*>(C), ldc, s)));
}

eh_status eh_syr2k_complex(eh_dtype dtype, eh_uplo uplo, eh_op trans, int64_t n, int64_t k, const void* alpha,
                           const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C,
                           int64_t ldc, void* stream) {
  EH_TRY(check_rank_args(dtype, uplo, trans, n, k, alpha, A, lda, B, ldb, true, beta, C, ldc, EH_OP_C));
  if (n == 0) return EH_SUCCESS;
  const int o1 = trans == EH_OP_N ? EH_OP_N : EH_OP_T, o2 = trans == EH_OP_N ? EH_OP_T : EH_OP_N;
  // C = alpha*W + alpha*W^T + beta*C with W = op(A)*op(B)^T
  EH_RANK_DISPATCH((rank_update<R>(uplo, o1, o2, n, k, cscalar<R>(alpha), static_cast<const cplx<R>*>(A), lda,
                                   static_cast<const cplx<R>*>(B), ldb, cscalar<R>(alpha), true, false,
                                   cscalar<R>(beta), static_cast<cplx<R>*>(C), ldc, s)));
}

eh_status eh_her2k(eh_dtype dtype, eh_uplo uplo, eh_op trans, int64_t n, int64_t k, const void* alpha, const void* A,
                   int64_t lda, const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc, void* stream) {
  EH_TRY(check_rank_args(dtype, uplo, trans, n, k, alpha, A, lda, B, ldb, true, beta, C, ldc, EH_OP_T));
  if (n == 0) return EH_SUCCESS;
  const int o1 = trans == EH_OP_N ? EH_OP_N : EH_OP_C, o2 = trans == EH_OP_N ? EH_OP_C : EH_OP_N;
  // C = alpha*W + conj(alpha)*W^H + beta*C with W = op(A)*op(B)^H, beta real
  EH_RANK_DISPATCH((rank_update<R>(uplo, o1, o2, n, k, cscalar<R>(alpha), static_cast<const cplx<R>*>(A), lda,
                                   static_cast<const cplx<R>*>(B), ldb, conj_if(cscalar<R>(alpha), true), true, true,
                                   cplx<R>(scalar<R>(beta)), static_cast<cplx<R>*>(C), ldc, s)));
}

eh_status eh_gemmtr_complex(eh_dtype dtype, eh_uplo uplo, eh_op opa, eh_op opb, int64_t n, int64_t k,
                            const void* alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
                            const void* beta, void* C, int64_t ldc, void* stream) {
  if (uplo != EH_UPPER && uplo != EH_LOWER) return fail(EH_ERR_INVALID, "uplo must be EH_UPPER or EH_LOWER");
  EH_TRY(check_gemm_args(dtype, opa, opb, n, n, k, alpha, A, lda, B, ldb, beta, C, ldc));
  if (n > INT32_MAX / 2 || k > INT32_MAX / 2) return fail(EH_ERR_UNSUPPORTED, "complex n, k must be below 2^30");
  if (n == 0) return EH_SUCCESS;
  EH_RANK_DISPATCH((rank_update<R>(uplo, opa, opb, n, k, cscalar<R>(alpha), static_cast<const cplx<R>*>(A), lda,
                                   static_cast<const cplx<R>*>(B), ldb, cplx<R>(0), false, false, cscalar<R>(beta),
                                   static_cast<cplx<R>*>(C), ldc, s)));
}

#undef EH_RANK_DISPATCH

}  // extern "C"
