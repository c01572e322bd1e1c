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
//   trtri_kernel       the inverses of all 64 x 64 diagonal blocks of a
//                      triangular matrix in one launch (one workgroup per
//                      block, one thread per column of the inverse, the
//                      column in registers, the block and the reciprocals of
//                      its diagonal in LDS); the solve then multiplies each
//                      diagonal block of right-hand sides by its inverse on
//                      the MFMA kernels.
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

// Inverses of the diagonal blocks of a triangular matrix, one workgroup per
// block: block b is M_b(i, k) = mt ? A[k + i*lda] : A[i + k*lda] at A + b*NB
// (1 + lda), of order nb = min(NB, order - b*NB), lower or upper, unit
// diagonal or not.  Thread v solves M_b x = e_v by substitution, multiplying
// by the reciprocal of the diagonal as TriangularSolverMatrix.h:67 does, and
// writes column v of inv(M_b) to W + b*NB*NB (NB x NB, column-major, zero
// outside the triangle and past nb).  Entries of M_b outside its triangle
// and past nb are zero in LDS and the reciprocal diagonal past nb is one, so
// the unrolled substitution needs no guards.  Each dot product runs as two
// interleaved partial sums (half the dependent chain).
template <typename T, int NB>
__global__ void __launch_bounds__(NB) trtri_kernel(TrtriParams<T> p) {
  __shared__ T M[NB][NB];
  __shared__ T dinv[NB];
  const int b = blockIdx.x;
  const int nb = min(NB, p.order - b * NB);
  const T* A = p.A + (int64_t)b * NB * (1 + p.lda);
  for (int idx = threadIdx.x; idx < NB * NB; idx += NB) {
    const int i = idx % NB, k = idx / NB;
    const bool in = i < nb && k < nb && (p.lower ? k < i : k > i);
    M[i][k] = in ? (p.mt ? A[k + (int64_t)i * p.lda] : A[i + (int64_t)k * p.lda]) : T(0);
  }
  if (threadIdx.x < NB) {
    const int i = threadIdx.x;
    dinv[i] = (i >= nb || p.unit) ? T(1) : T(1) / A[i + (int64_t)i * p.lda];
  }
  __syncthreads();
  const int v = threadIdx.x;
  T x[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) x[i] = (i == v && i < nb) ? T(1) : T(0);
  if (p.lower) {
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      T s0 = x[i], s1 = T(0);
#pragma unroll
      for (int k = 0; k + 1 < i; k += 2) {
        s0 = fma_t(-M[i][k], x[k], s0);
        s1 = fma_t(-M[i][k + 1], x[k + 1], s1);
      }
      if (i & 1) s0 = fma_t(-M[i][i - 1], x[i - 1], s0);
      x[i] = (s0 + s1) * dinv[i];
    }
  } else {
#pragma unroll
    for (int i = NB - 1; i >= 0; --i) {
      T s0 = x[i], s1 = T(0);
#pragma unroll
      for (int k = i + 1; k + 1 < NB; k += 2) {
        s0 = fma_t(-M[i][k], x[k], s0);
        s1 = fma_t(-M[i][k + 1], x[k + 1], s1);
      }
      if ((NB - 1 - i) & 1) s0 = fma_t(-M[i][NB - 1], x[NB - 1], s0);
      x[i] = (s0 + s1) * dinv[i];
    }
  }
  T* w = p.W + (int64_t)b * NB * NB + (int64_t)v * NB;
#pragma unroll
  for (int i = 0; i < NB; ++i) w[i] = x[i];
}

template <typename T>
hipError_t launch_trtri(const TrtriParams<T>& p, hipStream_t s) {
  if (p.order <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((p.order + kTrsmLeaf - 1) / kTrsmLeaf);
  hipLaunchKernelGGL((trtri_kernel<T, kTrsmLeaf>), dim3(blocks), dim3(kTrsmLeaf), 0, s, p);
  return hipGetLastError();
}

template hipError_t launch_tri_dense<float>(const float*, int64_t, int, float*, int64_t, bool, bool, bool,
                                            hipStream_t);
template hipError_t launch_tri_dense<double>(const double*, int64_t, int, double*, int64_t, bool, bool, bool,
                                             hipStream_t);
template hipError_t launch_trtri<float>(const TrtriParams<float>&, hipStream_t);
template hipError_t launch_trtri<double>(const TrtriParams<double>&, hipStream_t);

}  // namespace eh
