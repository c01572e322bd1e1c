// level3_kernels.hip -- device kernels of the level-3 routines built on the product.
//
// Eigen runs TRMM, TRSM, SYMM and the rank updates on the same pack + GEBP
// engine as GEMM (Eigen/src/Core/products/TriangularMatrixMatrix.h,
// TriangularSolverMatrix.h, SelfadjointMatrixMatrix.h,
// GeneralMatrixMatrixTriangular.h).  Here the bulk of their flops goes to the
// MFMA product kernels through a recursive 2x2 blocking on the host; the two
// kernels below do the rest:
//
//   tri_dense_kernel   the dense form of a matrix stored in one triangle:
//                      symmetric (mirror the stored triangle) or triangular
//                      (zeros outside it, optional unit diagonal).  64x64
//                      tiles; the mirrored half goes through LDS so that both
//                      the reads and the writes are coalesced.  HBM-bound.
//   trsm_leaf_kernel   a diagonal block of at most 32 rows of a triangular
//                      solve: one thread per right-hand side, the solution
//                      vector in registers, the block (and the reciprocals of
//                      its diagonal, as TriangularSolverMatrix.h:67 uses them)
//                      in LDS, read as wave-wide broadcasts.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace eh {

template <typename T>
__global__ void __launch_bounds__(256) tri_dense_kernel(const T* A, int64_t lda, int n, T* D, int64_t ldd, int lower,
                                                        int sym, int unit) {
  __shared__ T mir[64][65];
  const int bi = blockIdx.x * 64, bj = blockIdx.y * 64;  // D rows bi.., columns bj..
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  // (i, j) is stored when i >= j (lower) or i <= j (upper)
  auto stored = [&](int i, int j) { return lower ? i >= j : i <= j; };
  // does the tile hold elements outside the stored triangle?
  const bool mixed = lower ? bi < bj + 63 : bj < bi + 63;
  if (sym && mixed) {
    // mir[c][r] = A(bj + c, bi + r) = the mirror of D(bi + r, bj + c); lanes
    // walk rows of A (coalesced), only stored elements are read
    const int ar = bj + tx;
    for (int r = ty; r < 64; r += 4) {
      const int ac = bi + r;
      mir[tx][r] = (ar < n && ac < n && stored(ar, ac)) ? A[ar + (int64_t)ac * lda] : T(0);
    }
    __syncthreads();
  }
  const int i = bi + tx;
  if (i >= n) return;
  for (int r = ty; r < 64; r += 4) {
    const int j = bj + r;
    if (j >= n) break;
    T v;
    if (unit && i == j)
      v = T(1);
    else if (stored(i, j))
      v = A[i + (int64_t)j * lda];
    else
      v = sym ? mir[r][tx] : T(0);
    D[i + (int64_t)j * ldd] = v;
  }
}

template <typename T>
hipError_t launch_tri_dense(const T* A, int64_t lda, int n, T* D, int64_t ldd, bool lower, bool sym, bool unit,
                            hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned nb = (unsigned)((n + 63) / 64);
  hipLaunchKernelGGL(tri_dense_kernel<T>, dim3(nb, nb), dim3(256), 0, s, A, lda, n, D, ldd, (int)lower, (int)sym,
                     (int)unit);
  return hipGetLastError();
}

// Solve M x = alpha b for every right-hand side v < nvec, in place in B:
// element i of vector v is B[i*es + v*vs]; M(i, k) = mt ? A[k + i*lda]
// : A[i + k*lda] is lower or upper triangular of order nb <= NB.  Entries of
// M outside its triangle (and past nb) are zero in LDS and the reciprocal
// diagonal past nb is one, so the unrolled substitution needs no guards:
// padded components stay zero.
template <typename T, int NB>
__global__ void __launch_bounds__(256) trsm_leaf_kernel(TrsmLeaf<T> p) {
  __shared__ T M[NB][NB];
  __shared__ T dinv[NB];
  for (int idx = threadIdx.x; idx < NB * NB; idx += 256) {
    const int i = idx % NB, k = idx / NB;
    const bool in = i < p.nb && k < p.nb && (p.lower ? k < i : k > i);
    M[i][k] = in ? (p.mt ? p.A[k + (int64_t)i * p.lda] : p.A[i + (int64_t)k * p.lda]) : T(0);
  }
  for (int i = threadIdx.x; i < NB; i += 256)
    dinv[i] = (i >= p.nb || p.unit) ? T(1) : T(1) / p.A[i + (int64_t)i * p.lda];
  __syncthreads();
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= p.nvec) return;
  T* b = p.B + (int64_t)v * p.vs;
  T x[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) x[i] = i < p.nb ? p.alpha * b[(int64_t)i * p.es] : T(0);
  if (p.lower) {
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      T s = x[i];
#pragma unroll
      for (int k = 0; k < i; ++k) s = fma_t(-M[i][k], x[k], s);
      x[i] = s * dinv[i];
    }
  } else {
#pragma unroll
    for (int i = NB - 1; i >= 0; --i) {
      T s = x[i];
#pragma unroll
      for (int k = i + 1; k < NB; ++k) s = fma_t(-M[i][k], x[k], s);
      x[i] = s * dinv[i];
    }
  }
#pragma unroll
  for (int i = 0; i < NB; ++i)
    if (i < p.nb) b[(int64_t)i * p.es] = x[i];
}

template <typename T>
hipError_t launch_trsm_leaf(const TrsmLeaf<T>& p, hipStream_t s) {
  if (p.nb < 1 || p.nb > kTrsmLeaf) return hipErrorInvalidValue;
  if (p.nvec <= 0) return hipSuccess;
  hipLaunchKernelGGL((trsm_leaf_kernel<T, kTrsmLeaf>), dim3((unsigned)((p.nvec + 255) / 256)), dim3(256), 0, s, p);
  return hipGetLastError();
}

template hipError_t launch_tri_dense<float>(const float*, int64_t, int, float*, int64_t, bool, bool, bool,
                                            hipStream_t);
template hipError_t launch_tri_dense<double>(const double*, int64_t, int, double*, int64_t, bool, bool, bool,
                                             hipStream_t);
template hipError_t launch_trsm_leaf<float>(const TrsmLeaf<float>&, hipStream_t);
template hipError_t launch_trsm_leaf<double>(const TrsmLeaf<double>&, hipStream_t);

}  // namespace eh
