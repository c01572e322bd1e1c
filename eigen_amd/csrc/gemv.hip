// gemv.hip -- matrix-vector products (C with one column or one row).
//
// Eigen sends a product whose result is a vector to its GEMV kernels
// instead of GEBP (GeneralMatrixMatrix.h:480-490 ->
// Eigen/src/Core/products/GeneralMatrixVector.h:71-289, the column-major
// "y += alpha*A*x" and row-major "y += alpha*A^T*x" kernels).  Such a
// product reads every element of the matrix once and does two flops per
// element, so it is bound by HBM bandwidth; MFMA tiles would waste all but
// one column.  Two streaming kernels:
//
//   gemv_n_kernel  op(A)(i,k) = A[i + k*lda] (rows contiguous): each thread
//                  owns 16 contiguous bytes of rows and walks k, the vector
//                  element of each k being wave-uniform (scalar loads);
//   gemv_t_kernel  op(A)(i,k) = A[k + i*lda] (k contiguous): each wave owns
//                  eight rows, lanes read 16-byte chunks along k and the
//                  vector chunk once for all eight rows, then reduce across
//                  the wave.
//
// Both split K over a second grid dimension when the rows alone cannot fill
// the chip; partial sums go to the workspace and a reduction kernel applies
// alpha/beta in a fixed order (deterministic).  Operands are read through
// raw buffer resources: reads past the end of the matrix return zero and
// rows past m only feed outputs that are never stored.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace eh {

template <typename T>
struct GemvParams {
  const T* A;
  const T* x;
  T* y;
  T* W;            // split-K partials (splits x m) or nullptr
  int64_t lda, incx, incy;
  int m, k, k_per_split;
  uint32_t bytesA;  // extent of A from its base pointer (< 2^32, checked on the host)
  T alpha, beta;
};

template <typename T>
__device__ inline void gemv_store(const GemvParams<T>& p, int i, T acc, int z) {
  if (p.W) {
    p.W[(int64_t)z * p.m + i] = acc;
  } else {
    T* dst = p.y + (int64_t)i * p.incy;
    *dst = (p.beta == T(0)) ? p.alpha * acc : fma_t(p.alpha, acc, p.beta * *dst);
  }
}

// Rows per wave of gemv_t_kernel: eight (against four) read the vector
// chunk once per eight rows, 16384^2 FP64 A^T x 387 -> 352 us (FP32 equal).
constexpr int kGemvRowsPerWave = 8;
// k unroll of gemv_n_kernel (loads in flight per thread): 16 for FP64
// (16384^2: 370 -> 362 us), 8 for FP32 (16 measured 2% slower).
__host__ __device__ constexpr int gemv_n_unroll(int es) { return es == 8 ? 16 : 8; }

// 256 threads x EPT rows; k range of split blockIdx.y, unrolled by U.
template <typename T>
__global__ void __launch_bounds__(256) gemv_n_kernel(GemvParams<T> p) {
  constexpr int EPT = 16 / sizeof(T);
  constexpr int U = gemv_n_unroll((int)sizeof(T));
  const int i0 = (blockIdx.x * 256 + threadIdx.x) * EPT;
  const int z = blockIdx.y;
  const int kb = z * p.k_per_split, ke = min(p.k, kb + p.k_per_split);
  const __amdgpu_buffer_rsrc_t rA = make_rsrc(p.A, p.bytesA);
  const uint32_t vo = (uint32_t)i0 * sizeof(T);
  T acc[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) acc[e] = T(0);
  int kk = kb;
  for (; kk + U <= ke; kk += U) {
    u32x4 v[U];
    T xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      v[u] = buf_load16(rA, vo, (uint32_t)((uint64_t)(kk + u) * p.lda * sizeof(T)));
      xv[u] = p.x[(int64_t)(kk + u) * p.incx];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const T* a = reinterpret_cast<const T*>(&v[u]);
#pragma unroll
      for (int e = 0; e < EPT; ++e) acc[e] = fma_t(a[e], xv[u], acc[e]);
    }
  }
  for (; kk < ke; ++kk) {
    const u32x4 v = buf_load16(rA, vo, (uint32_t)((uint64_t)kk * p.lda * sizeof(T)));
    const T xv = p.x[(int64_t)kk * p.incx];
    const T* a = reinterpret_cast<const T*>(&v);
#pragma unroll
    for (int e = 0; e < EPT; ++e) acc[e] = fma_t(a[e], xv, acc[e]);
  }
#pragma unroll
  for (int e = 0; e < EPT; ++e)
    if (i0 + e < p.m) gemv_store(p, i0 + e, acc[e], z);
}

// Sum over the 64 lanes of a wave (butterfly, every lane ends with the sum).
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// 4 waves x kGemvRowsPerWave rows per workgroup; each lane covers EPT
// consecutive k of a 64*EPT chunk; k range of split blockIdx.y.
template <typename T>
__global__ void __launch_bounds__(256) gemv_t_kernel(GemvParams<T> p) {
  constexpr int EPT = 16 / sizeof(T);
  constexpr int RW = kGemvRowsPerWave;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int r0 = (blockIdx.x * 4 + wave) * RW;
  const int z = blockIdx.y;
  const int kb = z * p.k_per_split, ke = min(p.k, kb + p.k_per_split);
  const __amdgpu_buffer_rsrc_t rA = make_rsrc(p.A, p.bytesA);
  uint32_t vo[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) vo[r] = (uint32_t)(((uint64_t)(r0 + r) * p.lda + lane * EPT) * sizeof(T));
  T acc[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) acc[r] = T(0);
  for (int k0 = kb; k0 < ke; k0 += 64 * EPT) {
    const int kl = k0 + lane * EPT;
    T xv[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) xv[e] = kl + e < ke ? p.x[(int64_t)(kl + e) * p.incx] : T(0);
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const u32x4 v = buf_load16(rA, vo[r], (uint32_t)((uint64_t)k0 * sizeof(T)));
      const T* a = reinterpret_cast<const T*>(&v);
      // elements past ke (next row's data or the next split's) meet x = 0
      // only if they are finite; select them away instead
#pragma unroll
      for (int e = 0; e < EPT; ++e) acc[r] = fma_t(kl + e < ke ? a[e] : T(0), xv[e], acc[r]);
    }
  }
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    const T s = wave_sum(acc[r]);
    if (lane == 0 && r0 + r < p.m) gemv_store(p, r0 + r, s, z);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) gemv_reduce_kernel(GemvParams<T> p, int splits) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < p.m; i += gridDim.x * 256) {
    // partials summed in split order; loads issued eight at a time so that a
    // long split list costs splits/8 memory latencies, not splits
    T s = T(0);
    int z = 0;
    for (; z + 8 <= splits; z += 8) {
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = p.W[(int64_t)(z + u) * p.m + i];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; z < splits; ++z) s += p.W[(int64_t)z * p.m + i];
    T* dst = p.y + (int64_t)i * p.incy;
    *dst = (p.beta == T(0)) ? p.alpha * s : fma_t(p.alpha, s, p.beta * *dst);
  }
}

// Grid of a matrix-vector product: row blocks, and splits of K so that the
// grid has at least ~2048 (Aᵀx) or ~256 (Ax) workgroups' worth of streaming,
// each split a whole number of chunks of at least `min_k` columns.  Ax reads
// fastest with one long-running workgroup per CU: each then streams long runs
// of its columns, where 2048 shorter workgroups measured 10-20% slower.
GemvPlan gemv_plan(bool trans, int m, int k, int es) {
  const int ept = 16 / es;
  GemvPlan g{};
  g.blocks = trans ? (m + 4 * kGemvRowsPerWave - 1) / (4 * kGemvRowsPerWave) : (m + 256 * ept - 1) / (256 * ept);
  const int chunk = trans ? 64 * ept : gemv_n_unroll(es);
  const int min_k = trans ? 4 * chunk : 64;
  const int64_t target = trans ? 2048 : 256;
  int s = 1;
  while ((int64_t)g.blocks * s < target && k / (2 * s) >= min_k && s < 256) s *= 2;
  int64_t per = ((int64_t)k + s - 1) / s;
  per = (per + chunk - 1) / chunk * chunk;
  g.per = (int)per;
  g.splits = (int)(((int64_t)k + per - 1) / per);
  if (g.splits < 1) g.splits = 1;
  return g;
}

size_t gemv_workspace_bytes(const GemvPlan& g, int m, int es) {
  return g.splits > 1 ? (size_t)g.splits * m * es : 0;
}

template <typename T>
hipError_t launch_gemv(const GemvPlan& g, bool trans, int m, int k, T alpha, const T* A, int64_t lda,
                       uint32_t bytesA, const T* x, int64_t incx, T beta, T* y, int64_t incy, void* workspace,
                       hipStream_t s) {
  GemvParams<T> p{};
  p.A = A;
  p.x = x;
  p.y = y;
  p.lda = lda;
  p.incx = incx;
  p.incy = incy;
  p.m = m;
  p.k = k;
  p.bytesA = bytesA;
  p.alpha = alpha;
  p.beta = beta;
  p.k_per_split = g.per;
  p.W = g.splits > 1 ? static_cast<T*>(workspace) : nullptr;
  dim3 grid((unsigned)g.blocks, (unsigned)g.splits);
  if (trans)
    hipLaunchKernelGGL(gemv_t_kernel<T>, grid, dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL(gemv_n_kernel<T>, grid, dim3(256), 0, s, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || g.splits == 1) return e;
  int64_t rb = (m + 255) / 256;
  if (rb > 65536) rb = 65536;
  hipLaunchKernelGGL(gemv_reduce_kernel<T>, dim3((unsigned)rb), dim3(256), 0, s, p, g.splits);
  return hipGetLastError();
}

template hipError_t launch_gemv<float>(const GemvPlan&, bool, int, int, float, const float*, int64_t, uint32_t,
                                       const float*, int64_t, float, float*, int64_t, void*, hipStream_t);
template hipError_t launch_gemv<double>(const GemvPlan&, bool, int, int, double, const double*, int64_t, uint32_t,
                                        const double*, int64_t, double, double*, int64_t, void*, hipStream_t);

}  // namespace eh
