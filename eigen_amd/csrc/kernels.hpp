// kernels.hpp -- launch interfaces of the device kernels (host side).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cplx.hpp"

namespace eh {

// Parameters of one product C = alpha*op(A)*op(B) + beta*C (column-major).
// W is the split-K partial buffer (splits x M x N, column-major with ld = M),
// used only when the grid has more than one split.
template <typename T>
struct GemmParams {
  const T* A;
  const T* B;
  T* C;
  T* W;
  int64_t lda, ldb, ldc;
  int M, N, K;
  int k_per_split;  // multiple of the kernel's k slice
  T alpha, beta;
  // Byte extent of A and B from their base pointers (last element + 1); the
  // kernels read them through raw buffer resources of exactly this size, so
  // both must stay below 2^32 with a margin of one tile (the host splits
  // larger products).
  uint32_t bytesA, bytesB;
  // Triangle of C the product updates: 0 all of it, 1 the lower (row >= col),
  // 2 the upper (row <= col); the rest of C is neither read nor written.
  int tri;
};

// Plain MFMA product: tile 0..5 = 64x64, 64x128, 128x64, 128x128, 128x256, 256x128.
template <typename T>
hipError_t launch_gemm_mfma(const GemmParams<T>& p, int tile, bool ta, bool tb, int splits, hipStream_t s);
template <typename T>
hipError_t launch_splitk_reduce(const GemmParams<T>& p, int splits, hipStream_t s);
template <typename T>
hipError_t launch_scale_c(T* C, int M, int N, int64_t ldc, T beta, hipStream_t s, int tri = 0);
template <typename T>
hipError_t launch_strided_combine(const T* W, int M, int N, T* C, int64_t incc, int64_t ldc, T alpha, T beta,
                                  hipStream_t s);

// Matrix-vector product y = alpha*op(A)*x + beta*y, op(A) m x k: trans
// means op(A)(i,k) = A[k + i*lda], otherwise A[i + k*lda].  The plan fixes
// the grid (row blocks x K splits); with more than one split the launcher
// needs gemv_workspace_bytes of workspace for the partial sums.
struct GemvPlan {
  int blocks, splits, per;
};
GemvPlan gemv_plan(bool trans, int m, int k, int es);
size_t gemv_workspace_bytes(const GemvPlan& g, int m, int es);
template <typename T>
hipError_t launch_gemv(const GemvPlan& g, bool trans, int m, int k, T alpha, const T* A, int64_t lda,
                       uint32_t bytesA, const T* x, int64_t incx, T beta, T* y, int64_t incy, void* workspace,
                       hipStream_t s);

// Dense form of the n x n matrix stored in one triangle of A (lower: i >= j):
// sym mirrors the stored triangle, otherwise the other triangle is zero and
// unit puts ones on the diagonal.  D has leading dimension ldd.  herm (complex
// types): the mirror is conjugated and the diagonal's imaginary part taken as
// zero (a Hermitian matrix); conj: every element is conjugated.
template <typename T>
hipError_t launch_tri_dense(const T* A, int64_t lda, int n, T* D, int64_t ldd, bool lower, bool sym, bool unit,
                            hipStream_t s, bool herm = false, bool conj = false);

// Inverses of the diagonal blocks of order `leaf` of the triangular matrix
// M(i, k) = mt ? A[k + i*lda] : A[i + k*lda] (conjugated when conj) of the
// given order (lower or upper, unit diagonal or not): block b goes to
// W + b*leaf^2 as a dense leaf x leaf column-major matrix (zeros outside the
// triangle, and past the order in the last block).  leaf = kTrsmLeaf for
// real types, kTrsmLeafComplex for complex ones (the column of the inverse a
// thread keeps in registers is twice as wide).
constexpr int kTrsmLeaf = 64;
constexpr int kTrsmLeafComplex = 32;
template <typename T>
struct TrtriParams {
  const T* A;
  int64_t lda;
  int mt, lower, unit, order;
  T* W;
  int conj = 0;
};
template <typename T>
hipError_t launch_trtri(const TrtriParams<T>& p, hipStream_t s);
// D_p = [S_2p 0; 0 S_2p+1] (blocks of order L into blocks of order 2L,
// off-diagonal blocks zero), p < count: the first step of doubling the
// order of the inverted diagonal blocks.
template <typename T>
hipError_t launch_inverse_place(const T* S, T* D, int L, int64_t count, hipStream_t s);

// Triangle of a complex rank update (level3_kernels.hip): on the lower
// (upper) triangle, C = alpha*W + alpha2*f(W^T) + beta*C with f = conj when
// herm (W^H), the alpha2 term only when use2, C not read when beta == 0, the
// diagonal's imaginary part zeroed when herm; W == nullptr: C = beta*C.
template <typename R>
hipError_t launch_tri_combine(const cplx<R>* W, int64_t ldw, int n, cplx<R>* C, int64_t ldc, bool lower,
                              cplx<R> alpha, cplx<R> alpha2, bool use2, cplx<R> beta, bool herm, hipStream_t s);

// Complex product as one real product (complex_kernels.hip): A3 (2m x 2k
// real, ld3) = the 2x2 real blocks of alpha*op(A); B3 (2k x n real, ld3) =
// op(B) with real and imaginary parts in consecutive rows; C = beta*C for a
// complex beta.  T is the real type; A, B, C hold interleaved (re, im) pairs
// with leading dimensions in complex elements.
template <typename T>
hipError_t launch_complex_pack_a(const T* A, int64_t lda, int m, int k, bool ta, bool conj, T ar, T ai, T* A3,
                                 int64_t ld3, hipStream_t s);
template <typename T>
hipError_t launch_complex_pack_b(const T* B, int64_t ldb, int k, int n, bool tb, bool conj, T* B3, int64_t ld3,
                                 hipStream_t s);
template <typename T>
hipError_t launch_complex_scale(T* C, int m, int n, int64_t ldc, T br, T bi, hipStream_t s);
// Complex matrix-vector product on the real view of A: the real right-hand
// sides from x (op 0 = N, 1 = T, 2 = C) and the combination of the real
// product P into y (complex_kernels.hip documents the layouts).
template <typename T>
hipError_t launch_complex_gemv_gather(const T* x, int64_t incx, int len, int op, T* X, int64_t ld, hipStream_t s);
template <typename T>
hipError_t launch_complex_gemv_combine(const T* P, int64_t ldp, int len, int op, T ar, T ai, T br, T bi, T* y,
                                       int64_t incy, hipStream_t s);

// Split-BF16 FP32 product over packed operands.  Each operand is first packed
// into bf16 planes P[ks][term][row][32] (rows padded to split_rows_padded,
// ks = k / 32 slices) by launch_pack_split, then the product kernel streams
// the planes into LDS.
inline int split_rows_padded(int rows) { return (rows + 255) / 256 * 256; }
inline int split_kslices(int K) { return (K + 31) / 32; }
inline size_t split_packed_bytes(int rows, int K, int terms) {
  return (size_t)split_kslices(K) * terms * split_rows_padded(rows) * 64;
}
struct SplitParams {
  const void* Ap;  // packed op(A): M rows
  const void* Bp;  // packed op(B)^T: N rows
  float* C;
  float* W;
  int64_t ldc;
  int M, N, KS, Mp, Np;
  int ks_per_split;
  float alpha, beta;
  // One-term products only: a nonzero value is the leading dimension (in
  // elements) of an operand read in place instead of from packed planes --
  // its op(X) rows k-contiguous, K a multiple of 32, 16-byte aligned rows.
  int64_t ldAd, ldBd;
};
hipError_t launch_pack_split(const float* X, int64_t ld, bool kcont, int rows, int K, int terms, void* P,
                             hipStream_t s);
// f16: IEEE half planes (terms = 1) on the FP16 MFMA instruction.
hipError_t launch_gemm_split_packed(const SplitParams& p, int terms, int tile, int splits, hipStream_t s,
                                   bool f16 = false);
// 16-bit operand (bf16 or half bits) -> one plane in the same packed layout
// (terms = 1); a bit copy, so it serves both formats.
hipError_t launch_pack_bf16(const unsigned short* X, int64_t ld, bool kcont, int rows, int K, void* P,
                            hipStream_t s);
hipError_t launch_split_bf16(const float* x, int64_t n, int terms, unsigned short* t0, unsigned short* t1,
                             unsigned short* t2, hipStream_t s);

// Seeded device generators (same streams as the documented splitmix64 scheme).
hipError_t launch_fill(int dtype, void* x, int64_t n, uint64_t seed, int dist, double p0, double p1,
                       hipStream_t s);

}  // namespace eh
