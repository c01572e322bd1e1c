// grouped.hip -- many independent products of uneven shapes in one launch.
//
// Eigen evaluates small and fixed-size products one by one on the host
// (Eigen/src/Core/GeneralProduct.h:26-150 chooses a coefficient-based or GEMM
// path per product; Eigen/src/Core/ProductEvaluators.h:606-692 is the lazy
// small-product kernel).  Here a whole batch of FP32 and FP64 products runs
// in one persistent kernel that walks a device-side work list:
//
//   1. eh_grouped_create uploads one descriptor per product (ordered by
//      decreasing depth, longest work first);
//   2. a build kernel computes each product's 32x32 output-tile count and an
//      exclusive prefix sum over the products (the work list);
//   3. the persistent kernel starts 8 waves per SIMD; wave w takes tiles
//      w, w + W, w + 2W, ... of the list (W waves in all: the list is sorted
//      deepest first, so the round-robin deal balances like longest-first
//      scheduling, without a shared counter to contend on), finds each tile's
//      product by binary search in the prefix sums and computes the 32x32 tile
//      over the full depth with MFMA micro-tiles (v_mfma_f32_32x32x2_f32 for
//      FP32, a 2x2 block of v_mfma_f64_16x16x4_f64 for FP64), reading its
//      operands straight from HBM (no LDS: one wave owns one tile), and fuses
//      alpha/beta into the store.  A wave exits after its last tile.
//
// The two launches are captured once in a hipGraph and replayed
// (eh_grouped_graph_launch); the graph holds kernel nodes only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <numeric>
#include <vector>

#include "../../include/eigen_amd.h"
#include "common.hpp"
#include "internal.hpp"

namespace eh {

struct GroupedDesc {
  const void* A;
  const void* B;
  void* C;
  int64_t lda, ldb, ldc;
  double alpha, beta;
  int32_t dtype, ta, tb;
  int32_t m, n, k;
  int32_t tiles_m, tiles_n;
};

// Work list: tile_begin[p] = first tile index of product p (exclusive prefix
// sum of tiles_m * tiles_n), tile_begin[count] = total tiles.
__global__ void __launch_bounds__(1024) grouped_build_kernel(GroupedDesc* d, int64_t count, int64_t* tile_begin) {
  __shared__ int64_t partial[1024];
  const int tid = threadIdx.x;
  // each thread owns a contiguous run of products
  const int64_t per = (count + 1023) / 1024;
  const int64_t lo = std::min<int64_t>(count, tid * per), hi = std::min<int64_t>(count, lo + per);
  int64_t sum = 0;
  for (int64_t p = lo; p < hi; ++p) {
    const int tm = (d[p].m + 31) / 32, tn = (d[p].n + 31) / 32;
    d[p].tiles_m = tm;
    d[p].tiles_n = tn;
    sum += (int64_t)tm * tn;
  }
  partial[tid] = sum;
  __syncthreads();
  // inclusive scan of the 1024 partial sums (Hillis-Steele)
  for (int off = 1; off < 1024; off <<= 1) {
    const int64_t v = tid >= off ? partial[tid - off] : 0;
    __syncthreads();
    partial[tid] += v;
    __syncthreads();
  }
  int64_t run = tid > 0 ? partial[tid - 1] : 0;
  for (int64_t p = lo; p < hi; ++p) {
    tile_begin[p] = run;
    run += (int64_t)d[p].tiles_m * d[p].tiles_n;
  }
  if (tid == 1023) tile_begin[count] = partial[1023];
}

template <typename T>
__device__ inline T load_or_zero(const T* X, int64_t ld, bool trans, int r, int c, int rows, int cols) {
  // element (r, c) of op(X) (rows x cols); op(X) = X or X^T as stored column-major
  if (r >= rows || c >= cols) return T(0);
  return trans ? X[c + (int64_t)r * ld] : X[r + (int64_t)c * ld];
}

// Element (r, k) of a 32-row slab of an operand, zero outside rows x K.
// KC: the operand is k-contiguous (X[k + r*ld]); otherwise rows are
// contiguous (X[r + k*ld]).
template <typename T, bool KC>
__device__ inline T slab_elem(const T* X, int64_t ld, int r, int k, int rows, int K) {
  if (r >= rows || k >= K) return T(0);
  return KC ? X[k + (int64_t)r * ld] : X[r + (int64_t)k * ld];
}

// Per-wave LDS scratch for one 32 x 16 slab (row stride padded to keep the
// 16-byte reads aligned): 32 * 20 floats or 32 * 18 doubles.
constexpr int kSlabF32 = 32 * 20, kSlabF64 = 32 * 18;

// FP32: lane (i, g) needs v[s] = op(X)(r0 + i, k0 + g*8 + s), s < 8.  Rows-
// contiguous operands are read directly (32 consecutive rows per load);
// k-contiguous ones are loaded 16 k x 4 rows per instruction (64-byte runs),
// transposed through the wave's LDS slab and read back.
template <bool KC>
__device__ inline void slab_f32(const float* X, int64_t ld, int r0, int k0, int rows, int K, int lane,
                                float* slab, float (&v)[8]) {
  const int i = lane & 31, g = lane >> 5;
  if constexpr (!KC) {
#pragma unroll
    for (int s = 0; s < 8; ++s) v[s] = slab_elem<float, false>(X, ld, r0 + i, k0 + g * 8 + s, rows, K);
  } else {
    const int kk = lane & 15, rr = lane >> 4;
#pragma unroll
    for (int t = 0; t < 8; ++t) slab[(rr + 4 * t) * 20 + kk] = slab_elem<float, true>(X, ld, r0 + rr + 4 * t, k0 + kk, rows, K);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int s = 0; s < 8; ++s) v[s] = slab[i * 20 + g * 8 + s];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();  // the slab is rewritten by the next chunk
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// FP64: lane (i, g) needs v[x][s] = op(X)(r0 + 16x + i, k0 + g*4 + s), s < 4.
template <bool KC>
__device__ inline void slab_f64(const double* X, int64_t ld, int r0, int k0, int rows, int K, int lane,
                                double* slab, double (&v)[2][4]) {
  const int i = lane & 15, g = lane >> 4;
  if constexpr (!KC) {
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int s = 0; s < 4; ++s) v[x][s] = slab_elem<double, false>(X, ld, r0 + 16 * x + i, k0 + g * 4 + s, rows, K);
  } else {
    const int kk = lane & 15, rr = lane >> 4;
#pragma unroll
    for (int t = 0; t < 8; ++t)
      slab[(rr + 4 * t) * 18 + kk] = slab_elem<double, true>(X, ld, r0 + rr + 4 * t, k0 + kk, rows, K);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int s = 0; s < 4; ++s) v[x][s] = slab[(16 * x + i) * 18 + g * 4 + s];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// One 32x32 FP32 tile of C = alpha*op(A)*op(B) + beta*C.  op(A) is read as a
// 32-row slab over k (k-contiguous when A is transposed); op(B) as the
// 32-row slab of op(B)^T (k-contiguous when B is not transposed).
template <bool TA, bool TB>
__device__ void grouped_tile_f32(const GroupedDesc& g, int tm, int tn, int lane, float* slab) {
  const float* A = static_cast<const float*>(g.A);
  const float* B = static_cast<const float*>(g.B);
  float* C = static_cast<float*>(g.C);
  floatx16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  // lane (i, grp) feeds k = k0 + grp*8 + s to instruction s of each 16-deep step
  for (int k0 = 0; k0 < g.k; k0 += 16) {
    float a[8], b[8];
    slab_f32<TA>(A, g.lda, tm * 32, k0, g.m, g.k, lane, slab, a);
    slab_f32<!TB>(B, g.ldb, tn * 32, k0, g.n, g.k, lane, slab, b);
#pragma unroll
    for (int s = 0; s < 8; ++s) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[s], b[s], acc, 0, 0, 0);
  }
  const float alpha = (float)g.alpha, beta = (float)g.beta;
  const int c = tn * 32 + Mfma<float>::col(lane);
  if (c >= g.n) return;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int rr = tm * 32 + Mfma<float>::row(r, lane);
    if (rr >= g.m) continue;
    float* dst = C + rr + (int64_t)c * g.ldc;
    *dst = beta == 0.0f ? alpha * acc[r] : __builtin_fmaf(alpha, acc[r], beta * *dst);
  }
}

// One 32x32 FP64 tile as a 2x2 block of 16x16 MFMA tiles.
template <bool TA, bool TB>
__device__ void grouped_tile_f64(const GroupedDesc& g, int tm, int tn, int lane, double* slab) {
  const double* A = static_cast<const double*>(g.A);
  const double* B = static_cast<const double*>(g.B);
  double* C = static_cast<double*>(g.C);
  doublex4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[x][y][r] = 0.0;
  for (int k0 = 0; k0 < g.k; k0 += 16) {
    double a[2][4], b[2][4];
    slab_f64<TA>(A, g.lda, tm * 32, k0, g.m, g.k, lane, slab, a);
    slab_f64<!TB>(B, g.ldb, tn * 32, k0, g.n, g.k, lane, slab, b);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x][s], b[y][s], acc[x][y], 0, 0, 0);
  }
#pragma unroll
  for (int y = 0; y < 2; ++y) {
    const int c = tn * 32 + y * 16 + Mfma<double>::col(lane);
    if (c >= g.n) continue;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = tm * 32 + x * 16 + Mfma<double>::row(r, lane);
        if (rr >= g.m) continue;
        double* dst = C + rr + (int64_t)c * g.ldc;
        *dst = g.beta == 0.0 ? g.alpha * acc[x][y][r] : __builtin_fma(g.alpha, acc[x][y][r], g.beta * *dst);
      }
  }
}

template <typename F>
__device__ inline void dispatch_trans(int ta, int tb, F&& f) {
  if (!ta && !tb)
    f(std::integral_constant<bool, false>(), std::integral_constant<bool, false>());
  else if (ta && !tb)
    f(std::integral_constant<bool, true>(), std::integral_constant<bool, false>());
  else if (!ta && tb)
    f(std::integral_constant<bool, false>(), std::integral_constant<bool, true>());
  else
    f(std::integral_constant<bool, true>(), std::integral_constant<bool, true>());
}

// Persistent kernel: wave w computes tiles w, w + W, w + 2W, ...
__global__ void __launch_bounds__(256) grouped_gemm_kernel(const GroupedDesc* __restrict__ d, int64_t count,
                                                           const int64_t* __restrict__ tile_begin) {
  __shared__ __attribute__((aligned(16))) double lds[4 * kSlabF64];  // one slab per wave
  const int lane = threadIdx.x & 63;
  double* slab64 = lds + (threadIdx.x >> 6) * kSlabF64;
  float* slab32 = reinterpret_cast<float*>(slab64);
  const int64_t total = tile_begin[count];
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
  const int64_t first = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  for (int64_t t = first; t < total; t += waves) {
    // product p with tile_begin[p] <= t < tile_begin[p+1]
    int64_t lo = 0, hi = count - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (tile_begin[mid] <= t)
        lo = mid;
      else
        hi = mid - 1;
    }
    const GroupedDesc g = d[lo];
    const int local = (int)(t - tile_begin[lo]);
    const int tm = local % g.tiles_m, tn = local / g.tiles_m;
    if (g.dtype == EH_F32)
      dispatch_trans(g.ta, g.tb, [&](auto ta, auto tb) { grouped_tile_f32<ta, tb>(g, tm, tn, lane, slab32); });
    else
      dispatch_trans(g.ta, g.tb, [&](auto ta, auto tb) { grouped_tile_f64<ta, tb>(g, tm, tn, lane, slab64); });
  }
}

}  // namespace eh

using namespace eh;

struct eh_grouped {
  std::vector<eh_gemm_problem> host;  // as given (for the per-product path)
  GroupedDesc* d_desc = nullptr;
  int64_t* d_begin = nullptr;
  int64_t count = 0;
  int64_t tiles = 0;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  std::mutex mu;
};

static eh_status grouped_enqueue(eh_grouped* g, hipStream_t s) {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
      cus = prop.multiProcessorCount;
  }
  hipLaunchKernelGGL(grouped_build_kernel, dim3(1), dim3(1024), 0, s, g->d_desc, g->count, g->d_begin);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "grouped build launch");
  // 8 waves per SIMD: every wave is independent (no LDS, no barrier)
  const int64_t want = std::max<int64_t>(1, std::min<int64_t>((int64_t)cus * 8, (g->tiles + 3) / 4));
  hipLaunchKernelGGL(grouped_gemm_kernel, dim3((unsigned)want), dim3(256), 0, s, g->d_desc, g->count, g->d_begin);
  e = hipGetLastError();
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "grouped kernel launch");
}

extern "C" {

eh_status eh_grouped_create(const eh_gemm_problem* problems, int64_t count, eh_grouped** out) {
  if (!out) return fail(EH_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (count <= 0 || !problems) return fail(EH_ERR_INVALID, "need at least one product");
  std::vector<GroupedDesc> desc;
  desc.reserve(count);
  int64_t tiles = 0;
  for (int64_t p = 0; p < count; ++p) {
    const eh_gemm_problem& q = problems[p];
    const double one = 1.0;
    eh_status st = check_gemm_args(q.dtype, q.opa, q.opb, q.m, q.n, q.k, &one, q.A, q.lda, q.B, q.ldb, &one, q.C,
                                   q.ldc);
    if (st != EH_SUCCESS) return fail(st, "product " + std::to_string(p) + ": " + eh_last_error());
    if (q.m == 0 || q.n == 0) continue;  // nothing to do
    GroupedDesc d{};
    d.A = q.A;
    d.B = q.B;
    d.C = q.C;
    d.lda = q.lda;
    d.ldb = q.ldb;
    d.ldc = q.ldc;
    d.alpha = (q.k == 0) ? 0.0 : q.alpha;
    d.beta = q.beta;
    d.dtype = q.dtype;
    d.ta = q.opa != EH_OP_N;
    d.tb = q.opb != EH_OP_N;
    d.m = (int32_t)q.m;
    d.n = (int32_t)q.n;
    d.k = d.alpha == 0.0 ? 0 : (int32_t)q.k;  // alpha == 0: A and B are never read
    desc.push_back(d);
    tiles += ((q.m + 31) / 32) * ((q.n + 31) / 32);
  }
  // longest products first: the round-robin deal then balances like longest-first scheduling
  std::stable_sort(desc.begin(), desc.end(), [](const GroupedDesc& x, const GroupedDesc& y) {
    return (int64_t)x.k * (x.dtype == EH_F64 ? 2 : 1) > (int64_t)y.k * (y.dtype == EH_F64 ? 2 : 1);
  });
  auto* g = new eh_grouped();
  g->host.assign(problems, problems + count);
  g->count = (int64_t)desc.size();
  g->tiles = tiles;
  if (g->count == 0) {
    *out = g;
    return EH_SUCCESS;
  }
  hipError_t e = hipMalloc(&g->d_desc, sizeof(GroupedDesc) * g->count);
  if (e == hipSuccess) e = hipMalloc(&g->d_begin, sizeof(int64_t) * (g->count + 1));
  if (e == hipSuccess) e = hipMemcpy(g->d_desc, desc.data(), sizeof(GroupedDesc) * g->count, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    eh_grouped_destroy(g);
    return fail(EH_ERR_ALLOC, std::string("grouped setup: ") + hipGetErrorString(e));
  }
  *out = g;
  return EH_SUCCESS;
}

eh_status eh_grouped_run(eh_grouped* g, void* stream) {
  if (!g) return fail(EH_ERR_INVALID, "NULL handle");
  if (g->count == 0) return EH_SUCCESS;
  return grouped_enqueue(g, static_cast<hipStream_t>(stream));
}

eh_status eh_grouped_graph_launch(eh_grouped* g, void* stream) {
  if (!g) return fail(EH_ERR_INVALID, "NULL handle");
  if (g->count == 0) return EH_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(g->mu);
  if (!g->exec) {
    // capture on a private stream so the caller's stream is never in capture mode
    hipStream_t cs;
    hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      hipStreamDestroy(cs);
      return hip_fail(e, "hipStreamBeginCapture");
    }
    eh_status st = grouped_enqueue(g, cs);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(cs, &graph);
    hipStreamDestroy(cs);
    if (st != EH_SUCCESS) return st;
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      hipGraphDestroy(graph);
      g->exec = nullptr;
      return hip_fail(e, "hipGraphInstantiate");
    }
    g->graph = graph;
  }
  hipError_t e = hipGraphLaunch(g->exec, s);
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "hipGraphLaunch");
}

eh_status eh_grouped_run_per_product(eh_grouped* g, void* stream) {
  if (!g) return fail(EH_ERR_INVALID, "NULL handle");
  for (const eh_gemm_problem& q : g->host) {
    eh_status st;
    if (q.dtype == EH_F32) {
      const float a = (float)q.alpha, b = (float)q.beta;
      st = eh_gemm(EH_F32, (eh_op)q.opa, (eh_op)q.opb, q.m, q.n, q.k, &a, q.A, q.lda, q.B, q.ldb, &b, q.C, q.ldc,
                   EH_PREC_PLAIN, stream);
    } else {
      st = eh_gemm(EH_F64, (eh_op)q.opa, (eh_op)q.opb, q.m, q.n, q.k, &q.alpha, q.A, q.lda, q.B, q.ldb, &q.beta, q.C,
                   q.ldc, EH_PREC_PLAIN, stream);
    }
    if (st != EH_SUCCESS) return st;
  }
  return EH_SUCCESS;
}

eh_status eh_grouped_info(eh_grouped* g, int64_t* tiles, int64_t* products) {
  if (!g) return fail(EH_ERR_INVALID, "NULL handle");
  if (tiles) *tiles = g->tiles;
  if (products) *products = g->count;
  return EH_SUCCESS;
}

eh_status eh_grouped_destroy(eh_grouped* g) {
  if (!g) return EH_SUCCESS;
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  if (g->d_desc) hipFree(g->d_desc);
  if (g->d_begin) hipFree(g->d_begin);
  delete g;
  return EH_SUCCESS;
}

}  // extern "C"
