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

// T is a real type or cplx<R>; herm (complex only): the mirror is conjugated
// and the diagonal's imaginary part is dropped (BLAS ?hemm reads only the
// real part of a Hermitian diagonal); conj: every element is conjugated.
template <typename T>
__global__ void __launch_bounds__(256) tri_dense_kernel(const T* A, int64_t lda, int n, T* D, int64_t ldd, int lower,
                                                        int sym, int unit, int herm, int conj) {
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
      mir[tx][r] = (ar < n && ac < n && stored(ar, ac)) ? conj_if(A[ar + (int64_t)ac * lda], herm) : T(0);
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
    else if (herm && i == j)
      v = real_part(A[i + (int64_t)j * lda]);
    else if (stored(i, j))
      v = A[i + (int64_t)j * lda];
    else
      v = sym ? mir[r][tx] : T(0);
    D[i + (int64_t)j * ldd] = conj_if(v, conj);
  }
}

template <typename T>
hipError_t launch_tri_dense(const T* A, int64_t lda, int n, T* D, int64_t ldd, bool lower, bool sym, bool unit,
                            hipStream_t s, bool herm, bool conj) {
  if (n <= 0) return hipSuccess;
  const unsigned nb = (unsigned)((n + 63) / 64);
  hipLaunchKernelGGL(tri_dense_kernel<T>, dim3(nb, nb), dim3(256), 0, s, A, lda, n, D, ldd, (int)lower, (int)sym,
                     (int)unit, (int)herm, (int)conj);
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
// interleaved partial sums (half the dependent chain).  T is a real type or
// cplx<R> (then p.conj conjugates M, for an inverse of A^H's blocks).
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
    M[i][k] = in ? conj_if(p.mt ? A[k + (int64_t)i * p.lda] : A[i + (int64_t)k * p.lda], p.conj != 0) : T(0);
  }
  if (threadIdx.x < NB) {
    const int i = threadIdx.x;
    dinv[i] = (i >= nb || p.unit) ? T(1) : recip(conj_if(A[i + (int64_t)i * p.lda], p.conj != 0));
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

// Complex rank updates: on the lower (upper) triangle of the n x n result,
//   C(i, j) = alpha*W(i, j) + alpha2*f(W(j, i)) + beta*C(i, j)
// with f = conj when herm (the W^H term of HER2K) else identity (W^T of
// SYR2K), the alpha2 term only when use2, C not read when beta == 0, and
// (herm) the diagonal's imaginary part set to zero as ?herk / ?her2k do.
// W == nullptr: the W terms are zero (C = beta*C on the triangle).  64 x 64
// tiles, tiles off the triangle return at once; the transposed term goes
// through LDS so both its reads and the writes of C are coalesced.
template <typename R>
__global__ void __launch_bounds__(256) tri_combine_kernel(const cplx<R>* W, int64_t ldw, int n, cplx<R>* C,
                                                          int64_t ldc, int lower, cplx<R> alpha, cplx<R> alpha2,
                                                          int use2, cplx<R> beta, int herm) {
  __shared__ cplx<R> tr[64][65];
  const int bi = blockIdx.x * 64, bj = blockIdx.y * 64;  // C rows bi.., columns bj..
  if (lower ? bi + 63 < bj : bj + 63 < bi) return;      // the whole tile is off the triangle
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  if (W && use2) {
    // tr[c][r] = W(bj + c, bi + r): lanes walk rows of W (coalesced)
    const int wr = bj + tx;
    for (int r = ty; r < 64; r += 4) {
      const int wc = bi + r;
      tr[tx][r] = (wr < n && wc < n) ? W[wr + (int64_t)wc * ldw] : cplx<R>(0);
    }
    __syncthreads();
  }
  const int i = bi + tx;
  if (i >= n) return;
  const bool read_c = !(beta == cplx<R>(0));
  for (int r = ty; r < 64; r += 4) {
    const int j = bj + r;
    if (j >= n) break;
    if (lower ? i < j : i > j) continue;
    cplx<R> v(0);
    if (W) {
      v = alpha * W[i + (int64_t)j * ldw];
      if (use2) v = v + alpha2 * conj_if(tr[r][tx], herm != 0);
    }
    if (read_c) v = v + beta * C[i + (int64_t)j * ldc];
    if (herm && i == j) v.im = R(0);
    C[i + (int64_t)j * ldc] = v;
  }
}

template <typename R>
hipError_t launch_tri_combine(const cplx<R>* W, int64_t ldw, int n, cplx<R>* C, int64_t ldc, bool lower,
                              cplx<R> alpha, cplx<R> alpha2, bool use2, cplx<R> beta, bool herm, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned nb = (unsigned)((n + 63) / 64);
  hipLaunchKernelGGL(tri_combine_kernel<R>, dim3(nb, nb), dim3(256), 0, s, W, ldw, n, C, ldc, (int)lower, alpha, alpha2,
                     (int)use2, beta, (int)herm);
  return hipGetLastError();
}
template hipError_t launch_tri_combine<float>(const cplx<float>*, int64_t, int, cplx<float>*, int64_t, bool,
                                              cplx<float>, cplx<float>, bool, cplx<float>, bool, hipStream_t);
template hipError_t launch_tri_combine<double>(const cplx<double>*, int64_t, int, cplx<double>*, int64_t, bool,
                                               cplx<double>, cplx<double>, bool, cplx<double>, bool, hipStream_t);

template <typename T>
struct TrsmLeafOf {
  static constexpr int value = kTrsmLeaf;
};
template <typename R>
struct TrsmLeafOf<cplx<R>> {
  static constexpr int value = kTrsmLeafComplex;
};

template <typename T>
hipError_t launch_trtri(const TrtriParams<T>& p, hipStream_t s) {
  if (p.order <= 0) return hipSuccess;
  constexpr int NB = TrsmLeafOf<T>::value;
  const unsigned blocks = (unsigned)((p.order + NB - 1) / NB);
  hipLaunchKernelGGL((trtri_kernel<T, NB>), dim3(blocks), dim3(NB), 0, s, p);
  return hipGetLastError();
}

// D_p = [S_2p 0; 0 S_2p+1] for p < count: S holds 2*count blocks of order
// L (L x L, column-major, consecutive), D count blocks of order 2L; the
// off-diagonal blocks are zeroed (one of them is then overwritten by the
// doubling product).  One thread per element of D.
template <typename T>
__global__ void __launch_bounds__(256) inverse_place_kernel(const T* __restrict__ S, T* __restrict__ D, int L,
                                                            int64_t total) {
  const int64_t idx = blockIdx.x * 256ll + threadIdx.x;
  if (idx >= total) return;
  const int64_t blk = 4ll * L * L;
  const int64_t p = idx / blk, r = idx % blk;
  const int j = (int)(r / (2 * L)), i = (int)(r % (2 * L));
  T v = T(0);
  if (i < L && j < L) v = S[(2 * p) * L * L + i + (int64_t)j * L];
  if (i >= L && j >= L) v = S[(2 * p + 1) * L * L + (i - L) + (int64_t)(j - L) * L];
  D[idx] = v;
}

template <typename T>
hipError_t launch_inverse_place(const T* S, T* D, int L, int64_t count, hipStream_t s) {
  const int64_t total = count * 4ll * L * L;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(inverse_place_kernel<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, S, D, L, total);
  return hipGetLastError();
}

#define EH_L3_INST(T)                                                                                       \
  template hipError_t launch_tri_dense<T>(const T*, int64_t, int, T*, int64_t, bool, bool, bool, hipStream_t, \
                                          bool, bool);                                                      \
  template hipError_t launch_trtri<T>(const TrtriParams<T>&, hipStream_t);                                  \
  template hipError_t launch_inverse_place<T>(const T*, T*, int, int64_t, hipStream_t);
EH_L3_INST(float)
EH_L3_INST(double)
EH_L3_INST(cplx<float>)
EH_L3_INST(cplx<double>)
#undef EH_L3_INST

}  // namespace eh
