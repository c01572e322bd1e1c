// grouped.hip -- many independent products of uneven shapes in one launch.
//
// Eigen evaluates small and fixed-size products one by one on the host
// (Eigen/src/Core/GeneralProduct.h:26-150 chooses a coefficient-based or GEMM
// path per product; Eigen/src/Core/ProductEvaluators.h:606-692 is the lazy
// small-product kernel).  Here a whole batch of FP32 and FP64 products runs
// in one persistent kernel that walks a work list:
//
//   1. eh_grouped_create builds the work list once, on the host: one
//      descriptor per product, grouped by operand type (dtype, op(A), op(B))
//      and deepest first within a type, the first 32x32 output tile of each
//      product (prefix sums of tile counts), and a partition of the tile
//      sequence into contiguous runs, one per resident workgroup.  Each
//      type gets runs in proportion to its estimated cost (k steps per tile
//      plus a per-tile overhead) and its runs carry equal cost, so every run
//      holds one operand type.  Within a product the tiles are numbered in
//      bands of Band tile columns, row by row inside a band, so any G
//      consecutive tiles form a block of about G / Band rows by Band columns
//      of tiles (schedule (G, Band): 16 x 4 by default, 8 x 2 or 4 x 2).
//   2. the kernel starts exactly the workgroups that fit on the chip at once;
//      each picks the instantiation for its run's operand type once, and
//      wave w takes tiles w, w + G, w + 2G, ... of the run, so the G waves of
//      a workgroup compute one block of a product at the same time: each
//      op(A) row panel and op(B) column panel of the block is read from HBM
//      once and shared through the CU's and the XCD's caches (a wave walking
//      a product alone finds its panels evicted by the time it comes back to
//      them: this layout halved the kernel's HBM reads and cut its time by a
//      fifth).  A barrier after every block keeps the waves in step (without
//      it they drift apart and lose the sharing).  Each wave advances a
//      product cursor (wave-uniform, scalar loads) and computes each 32x32
//      tile over the full depth with MFMA (v_mfma_f32_32x32x2_f32 for FP32, a
//      2x2 block of v_mfma_f64_16x16x4_f64 for FP64).  Operands are read with
//      raw buffer loads (reads past the end of an operand return zero, so no
//      load can fault), double-buffered in registers so the next k step is in
//      flight while the current one is multiplied; the final step skips the
//      MFMAs whose k values all lie past K.  The output tile is produced
//      transposed (C^T = op(B)^T op(A)^T) so that consecutive lanes store
//      consecutive rows of the column-major C, with alpha/beta fused into
//      the store.
//   3. products whose operands span 2^31 bytes or more (32-bit buffer
//      offsets) run after the kernel through the single-product path.
//
// eh_grouped_graph_launch captures these launches once in a hipGraph and
// replays it; the graph holds kernel nodes only.
//
// eh_gemm_batched_strided (arrays of one shape) writes the same work list on
// the device, except for small shapes, which take kernels of their own: up
// to 4 x 4 x 4 one thread per product (whole-column accesses for dense
// 2 x 2 and 4 x 4 matrices, nine-element runs for dense 3 x 3 ones), up to
// 16 x 16 with depth up to 64 (FP64) or 32 (FP32) one wave per product.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
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
  int64_t ldc;
  double alpha, beta;
  uint32_t bytesA, bytesB;  // extents of the stored operands (buffer ranges)
  uint32_t lda, ldb;        // clamped to 2^31 - 1: a leading dimension is only used when it is below that
  int32_t dtype, ta, tb;
  int32_t m, n, k;
  int32_t tiles_m, tiles_n;
};

// Schedules: waves per workgroup G and band width (tile columns of a block).
// Schedule 0 (16 waves, 4x4-tile blocks) is the default; EH_TUNE_TABLE
// times all of them on each new batch (eh_grouped_create).
constexpr int kNumSchedules = 3;
constexpr int kSchedWaves[kNumSchedules] = {16, 8, 4};
constexpr int kSchedBand[kNumSchedules] = {4, 2, 2};
// k depth of one step: 16 for FP32 (8 MFMAs), 8 for FP64 (8 MFMAs of a 2x2
// block); two steps of operands per wave stay within 128 registers (four
// waves per SIMD).
constexpr int kStepF32 = 16, kStepF64 = 8;
constexpr int kLaneF32 = kStepF32 / 2, kLaneF64 = kStepF64 / 4;  // consecutive k per lane

// ---- slab loads -------------------------------------------------------------
// A slab is 32 rows (of op(A), or of op(B)^T) by one k step.  KC: the operand
// is k-contiguous in memory (X[k + r*ld]); otherwise its rows are contiguous
// (X[r + k*ld]).  `vo` is the lane's byte offset at k0 = 0; elements with
// r >= rows only reach outputs that are never stored, elements with k >= K
// are zeroed by tail_* on the last step.

// FP32: lane (i = lane%32, g = lane/32) holds v[s] = X(r0 + i, k0 + kLaneF32*g + s).
template <bool KC>
__device__ inline uint32_t lane_offset_f32(int r0, int lane, uint32_t ld) {
  const uint32_t r = (uint32_t)(r0 + (lane & 31)), g = (uint32_t)(lane >> 5);
  return KC ? (r * ld + kLaneF32 * g) * 4u : (r + kLaneF32 * g * ld) * 4u;
}

template <bool KC>
__device__ inline void load_f32(__amdgpu_buffer_rsrc_t rs, uint32_t vo, int k0, uint32_t ld, float (&v)[kLaneF32]) {
  if constexpr (KC) {
    const uint32_t so = (uint32_t)k0 * 4u;
#pragma unroll
    for (int q = 0; q < kLaneF32 / 4; ++q) {
      const u32x4 x = buf_load16(rs, vo + 16u * q, so);
#pragma unroll
      for (int s = 0; s < 4; ++s) v[4 * q + s] = __uint_as_float(x[s]);
    }
  } else {
    // one load per k, the k stride added to a running scalar offset; the
    // empty asm keeps the compiler from precomputing all kLaneF32 offsets
    // into scalar registers ahead of the loop
    uint32_t so = (uint32_t)k0 * ld * 4u;
#pragma unroll
    for (int s = 0; s < kLaneF32; ++s) {
      v[s] = buf_load(rs, vo, so, (float*)nullptr);
      so += ld * 4u;
      asm volatile("" : "+s"(so));
    }
  }
}

__device__ inline void tail_f32(float (&v)[kLaneF32], int k0, int K, int lane) {
  const int kb = k0 + kLaneF32 * (lane >> 5);
#pragma unroll
  for (int s = 0; s < kLaneF32; ++s)
    if (kb + s >= K) v[s] = 0.0f;
}

// FP64: lane (i = lane%16, g = lane/16) holds v[x][s] = X(r0 + 16x + i, k0 + kLaneF64*g + s).
template <bool KC>
__device__ inline uint32_t lane_offset_f64(int r0, int lane, uint32_t ld) {
  const uint32_t r = (uint32_t)(r0 + (lane & 15)), g = (uint32_t)(lane >> 4);
  return KC ? (r * ld + kLaneF64 * g) * 8u : (r + kLaneF64 * g * ld) * 8u;
}

template <bool KC>
__device__ inline void load_f64(__amdgpu_buffer_rsrc_t rs, uint32_t vo, int k0, uint32_t ld,
                                double (&v)[2][kLaneF64]) {
#pragma unroll
  for (int x = 0; x < 2; ++x) {
    if constexpr (KC) {
      const uint32_t so = (uint32_t)k0 * 8u + (uint32_t)x * 16u * ld * 8u;
#pragma unroll
      for (int q = 0; q < kLaneF64 / 2; ++q) {
        const u32x4 w = buf_load16(rs, vo + 16u * q, so);
        v[x][2 * q] = __hiloint2double((int)w[1], (int)w[0]);
        v[x][2 * q + 1] = __hiloint2double((int)w[3], (int)w[2]);
      }
    } else {
#pragma unroll
      for (int s = 0; s < kLaneF64; ++s)
        v[x][s] = buf_load(rs, vo + (uint32_t)x * 128u, (uint32_t)(k0 + s) * ld * 8u, (double*)nullptr);
    }
  }
}

__device__ inline void tail_f64(double (&v)[2][kLaneF64], int k0, int K, int lane) {
  const int kb = k0 + kLaneF64 * (lane >> 4);
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int s = 0; s < kLaneF64; ++s)
      if (kb + s >= K) v[x][s] = 0.0;
}

// Tile `l` of a product in band order: bands of Band tile columns (the last
// band holds the remaining columns), each band walked row by row.
template <int Band>
__device__ inline void band_tile(const GroupedDesc& g, int l, int& tm, int& tn) {
  const int band_tiles = g.tiles_m * Band, full = g.tiles_n / Band;
  const int b = l / band_tiles;
  if (b < full) {
    const int r = l - b * band_tiles;
    tm = r / Band;
    tn = b * Band + r % Band;
  } else {
    const int w = g.tiles_n - full * Band, r = l - full * band_tiles;
    tm = r / w;
    tn = full * Band + r % w;
  }
}

// ---- tiles ------------------------------------------------------------------
// Tile<T, TA, TB> holds one 32x32 output tile of a product (its descriptor,
// tile coordinates, buffer resources and lane offsets, all wave-uniform
// except the offsets) and the operations on it: load k step `st` into a
// fragment, multiply a fragment into the accumulator, and store the tile.
// The tile is computed as C^T = op(B)^T op(A)^T: the MFMA's first operand is
// the op(B)^T slab (32 columns of C), its second the op(A) slab (32 rows of
// C), so consecutive lanes store consecutive rows of the column-major C.
template <typename T, bool TA, bool TB>
struct Tile;

// FP32: v_mfma_f32_32x32x2_f32; accumulator r of lane l holds
// C(m0 + l%32, n0 + row(r, l)).
template <bool TA, bool TB>
struct Tile<float, TA, TB> {
  static constexpr int kStep = kStepF32;
  struct Frag {
    float a[kLaneF32], b[kLaneF32];
  };
  using Acc = floatx16;
  GroupedDesc g;
  int tm, tn, steps;
  __amdgpu_buffer_rsrc_t rA, rB;
  uint32_t vA, vB;

  __device__ void init(const GroupedDesc& d, int tile_m, int tile_n, int lane) {
    g = d;
    tm = tile_m;
    tn = tile_n;
    steps = (g.k + kStep - 1) / kStep;
    rA = make_rsrc(g.A, g.bytesA);
    rB = make_rsrc(g.B, g.bytesB);
    vA = lane_offset_f32<TA>(tm * 32, lane, g.lda);
    vB = lane_offset_f32<!TB>(tn * 32, lane, g.ldb);
  }
  __device__ void load(int st, Frag& f) const {
    load_f32<TA>(rA, vA, st * kStep, g.lda, f.a);
    load_f32<!TB>(rB, vB, st * kStep, g.ldb, f.b);
  }
  __device__ static void zero(Acc& acc) {
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  }
  // The final step (possibly partial) masks the k values at or past K and
  // skips the MFMAs whose k values all lie there: MFMA s multiplies k0 + s
  // and k0 + kLaneF32 + s, so with r = K - k0 <= kLaneF32 only s < r remain.
  __device__ void mma(int st, Frag& f, Acc& acc, int lane) const {
    if (st < steps - 1) {
#pragma unroll
      for (int s = 0; s < kLaneF32; ++s) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(f.b[s], f.a[s], acc, 0, 0, 0);
      return;
    }
    tail_f32(f.a, st * kStep, g.k, lane);
    tail_f32(f.b, st * kStep, g.k, lane);
    const int live = min(kLaneF32, g.k - st * kStep);
#pragma unroll
    for (int s = 0; s < kLaneF32; ++s)
      if (s < live) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(f.b[s], f.a[s], acc, 0, 0, 0);
  }
  __device__ void store(Acc& acc, int lane) const {
    const int m = tm * 32 + (lane & 31);
    if (m >= g.m) return;
    const float alpha = (float)g.alpha, beta = (float)g.beta;
    float* C = static_cast<float*>(g.C) + m;
    if (beta != 0.0f) {  // read every C value of the lane before the first store
      float c[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int n = tn * 32 + Mfma<float>::row(r, lane);
        c[r] = n < g.n ? C[(int64_t)n * g.ldc] : 0.0f;
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[r] = __builtin_fmaf(alpha, acc[r], beta * c[r]);
    } else {
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[r] *= alpha;
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int n = tn * 32 + Mfma<float>::row(r, lane);
      if (n < g.n) C[(int64_t)n * g.ldc] = acc[r];
    }
  }
};

// FP64: a 2x2 block of v_mfma_f64_16x16x4_f64 tiles; acc.v[y][x] of lane l,
// entry r, holds C(m0 + 16x + l%16, n0 + 16y + l/16 + 4r).
template <bool TA, bool TB>
struct Tile<double, TA, TB> {
  static constexpr int kStep = kStepF64;
  struct Frag {
    double a[2][kLaneF64], b[2][kLaneF64];
  };
  struct Acc {
    doublex4 v[2][2];
  };
  GroupedDesc g;
  int tm, tn, steps;
  __amdgpu_buffer_rsrc_t rA, rB;
  uint32_t vA, vB;

  __device__ void init(const GroupedDesc& d, int tile_m, int tile_n, int lane) {
    g = d;
    tm = tile_m;
    tn = tile_n;
    steps = (g.k + kStep - 1) / kStep;
    rA = make_rsrc(g.A, g.bytesA);
    rB = make_rsrc(g.B, g.bytesB);
    vA = lane_offset_f64<TA>(tm * 32, lane, g.lda);
    vB = lane_offset_f64<!TB>(tn * 32, lane, g.ldb);
  }
  __device__ void load(int st, Frag& f) const {
    load_f64<TA>(rA, vA, st * kStep, g.lda, f.a);
    load_f64<!TB>(rB, vB, st * kStep, g.ldb, f.b);
  }
  __device__ static void zero(Acc& acc) {
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc.v[y][x][r] = 0.0;
  }
  // as for FP32: MFMA s multiplies k0 + kLaneF64*g + s (g = 0..3), so with
  // r = K - k0 <= kLaneF64 only s < r remain on the final step
  __device__ void mma(int st, Frag& f, Acc& acc, int lane) const {
    auto k_slice = [&](int s) {
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int x = 0; x < 2; ++x)
          acc.v[y][x] = __builtin_amdgcn_mfma_f64_16x16x4f64(f.b[y][s], f.a[x][s], acc.v[y][x], 0, 0, 0);
    };
    if (st < steps - 1) {
#pragma unroll
      for (int s = 0; s < kLaneF64; ++s) k_slice(s);
      return;
    }
    tail_f64(f.a, st * kStep, g.k, lane);
    tail_f64(f.b, st * kStep, g.k, lane);
    const int live = min(kLaneF64, g.k - st * kStep);
#pragma unroll
    for (int s = 0; s < kLaneF64; ++s)
      if (s < live) k_slice(s);
  }
  __device__ void store(Acc& acc, int lane) const {
    double* C = static_cast<double*>(g.C);
    // one 16-row half of the tile at a time (x), its C values read first
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      const int m = tm * 32 + 16 * x + (lane & 15);
      if (m >= g.m) continue;
      if (g.beta != 0.0) {
        double c[2][4];
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int n = tn * 32 + 16 * y + (lane >> 4) + 4 * r;
            c[y][r] = n < g.n ? C[m + (int64_t)n * g.ldc] : 0.0;
          }
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc.v[y][x][r] = __builtin_fma(g.alpha, acc.v[y][x][r], g.beta * c[y][r]);
      } else {
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc.v[y][x][r] *= g.alpha;
      }
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = tn * 32 + 16 * y + (lane >> 4) + 4 * r;
          if (n < g.n) C[m + (int64_t)n * g.ldc] = acc.v[y][x][r];
        }
    }
  }
};

// The tiles of one workgroup run, all of one operand type <T, TA, TB>.  Wave
// w computes tiles tbeg + w, tbeg + w + G, ...; per tile the k steps go
// through two register fragments, one step in flight while the other is
// multiplied.
template <typename T, bool TA, bool TB, int G, int Band>
__device__ void run_tiles(const GroupedDesc* __restrict__ d, const int64_t* __restrict__ tile_begin, int64_t tbeg,
                          int64_t tend, int p, int w, int lane) {
  using TL = Tile<T, TA, TB>;
  for (int64_t t0 = tbeg, t = tbeg + w; t0 < tend; t0 += G, t += G) {
    if (t < tend) {
      while (tile_begin[p + 1] <= t) ++p;
      const GroupedDesc& g = d[p];
      int tm, tn;
      band_tile<Band>(g, (int)(t - tile_begin[p]), tm, tn);
      TL cur;
      cur.init(g, tm, tn, lane);
      typename TL::Frag f0, f1;
      typename TL::Acc acc;
      TL::zero(acc);
      const int n = cur.steps;
      if (n > 0) cur.load(0, f0);
      if (n > 1) cur.load(1, f1);
      for (int i = 0; i < n; i += 2) {
        cur.mma(i, f0, acc, lane);
        if (i + 2 < n) cur.load(i + 2, f0);
        if (i + 1 < n) {
          cur.mma(i + 1, f1, acc, lane);
          if (i + 3 < n) cur.load(i + 3, f1);
        }
      }
      cur.store(acc, lane);
    }
    __syncthreads();  // every wave runs the same number of blocks
  }
}

// Persistent kernel of schedule (G, Band), four waves per SIMD (at most 128
// registers) for every G: workgroup v (XCD-major numbering
// of the resident workgroups) computes tiles run[v] .. run[v+1]-1 of the
// work list; the host cuts runs where the operand type changes, so one
// instantiation of run_tiles serves the whole run.
template <int G, int Band>
__global__ void __launch_bounds__(64 * G, 4) grouped_gemm_kernel(const GroupedDesc* __restrict__ d,
                                                              const int64_t* __restrict__ tile_begin,
                                                              const int64_t* __restrict__ run,
                                                              const int32_t* __restrict__ run_prod) {
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int v = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t tbeg = run[v], tend = run[v + 1];
  if (tbeg >= tend) return;
  const int p = run_prod[v];
  const GroupedDesc& g = d[p];
  if (g.dtype == EH_F32) {
    if (!g.ta && !g.tb)
      run_tiles<float, false, false, G, Band>(d, tile_begin, tbeg, tend, p, w, lane);
    else if (g.ta && !g.tb)
      run_tiles<float, true, false, G, Band>(d, tile_begin, tbeg, tend, p, w, lane);
    else if (!g.ta && g.tb)
      run_tiles<float, false, true, G, Band>(d, tile_begin, tbeg, tend, p, w, lane);
    else
      run_tiles<float, true, true, G, Band>(d, tile_begin, tbeg, tend, p, w, lane);
  } else {
    if (!g.ta && !g.tb)
      run_tiles<double, false, false, G, Band>(d, tile_begin, tbeg, tend, p, w, lane);
    else if (g.ta && !g.tb)
      run_tiles<double, true, false, G, Band>(d, tile_begin, tbeg, tend, p, w, lane);
    else if (!g.ta && g.tb)
      run_tiles<double, false, true, G, Band>(d, tile_begin, tbeg, tend, p, w, lane);
    else
      run_tiles<double, true, true, G, Band>(d, tile_begin, tbeg, tend, p, w, lane);
  }
}

using GroupedKernel = void (*)(const GroupedDesc*, const int64_t*, const int64_t*, const int32_t*);
static const GroupedKernel kSchedKernel[kNumSchedules] = {grouped_gemm_kernel<16, 4>, grouped_gemm_kernel<8, 2>,
                                                          grouped_gemm_kernel<4, 2>};
static_assert(kSchedWaves[0] == 16 && kSchedBand[0] == 4 && kSchedWaves[1] == 8 && kSchedBand[1] == 2 &&
                  kSchedWaves[2] == 4 && kSchedBand[2] == 2,
              "kernel table and schedule constants agree");

// Resident workgroups of schedule s on the current device (queried once per
// process and schedule; function-local statics initialise thread-safely).
static int64_t resident_groups(int s) {
  static int64_t cache[kNumSchedules] = {};
  static std::once_flag once[kNumSchedules];
  std::call_once(once[s], [s] {
    int dev = 0, cus = 256, per_cu = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
        prop.multiProcessorCount > 0)
      cus = prop.multiProcessorCount;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kSchedKernel[s], 64 * kSchedWaves[s], 0) !=
            hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    cache[s] = (int64_t)cus * per_cu;
  });
  return cache[s];
}

static hipError_t launch_grouped(int s, int64_t groups, const GroupedDesc* d, const int64_t* begin,
                                 const int64_t* run, const int32_t* run_prod, hipStream_t stream) {
  hipLaunchKernelGGL(kSchedKernel[s], dim3((unsigned)groups), dim3(64 * kSchedWaves[s]), 0, stream, d, begin, run,
                     run_prod);
  return hipGetLastError();
}

// Bytes spanned by a column-major operand of `rows` x `cols` (leading
// dimension ld); 0 for an empty one.
static uint64_t extent_bytes(int64_t rows, int64_t cols, int64_t ld, int es) {
  if (rows <= 0 || cols <= 0) return 0;
  return ((uint64_t)(cols - 1) * (uint64_t)ld + (uint64_t)rows) * (uint64_t)es;
}

// Estimated cost of one tile: its k steps (8 MFMAs of equal duration each,
// for both dtypes) plus the store and the first load.
static int64_t tile_cost(const GroupedDesc& d) {
  const int step = d.dtype == EH_F64 ? kStepF64 : kStepF32;
  return 2 + (d.k + step - 1) / step;
}

// Work list of `count` products of one shape at fixed byte strides, built on
// the device: descriptor p is `base` with its operands advanced by p strides,
// product p owns tiles [p*T, (p+1)*T), and with equal tile costs the W
// workgroups' equal-cost runs start at tile floor(v * count * T / W).
__global__ void __launch_bounds__(256) strided_build_kernel(GroupedDesc base, int64_t sa, int64_t sb, int64_t sc,
                                                            int64_t count, int64_t T, int64_t W, GroupedDesc* d,
                                                            int64_t* begin, int64_t* run, int32_t* run_prod) {
  const int64_t total = count * T;
  const int64_t n = max(count + 1, W + 1);
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    if (i < count) {
      GroupedDesc g = base;
      g.A = static_cast<const char*>(base.A) + i * sa;
      g.B = static_cast<const char*>(base.B) + i * sb;
      g.C = static_cast<char*>(base.C) + i * sc;
      d[i] = g;
    }
    if (i <= count) begin[i] = i * T;
    if (i <= W) run[i] = i == W ? total : (int64_t)((__int128)total * i / W);
    if (i < W) {
      const int64_t t = (int64_t)((__int128)total * i / W);
      run_prod[i] = (int32_t)min(t / T, count - 1);
    }
  }
}

// Products of at most 4 x 4 x 4 (Matrix2/3/4-sized), one thread each: the
// operands of product p are loaded whole into registers (unconditional loads
// clamped into the product, zeroed by selects, so no load waits on another),
// C(i, j) = alpha * sum_k op(A)(i, k) op(B)(k, j) + beta * C(i, j) with the
// k sum in order.  A 32 x 32 tile of the grouped kernel would spend 1/64 of
// its work on such a product.
template <typename T>
__global__ void __launch_bounds__(256) tiny_batched_kernel(const T* __restrict__ A, int64_t lda, int64_t sa,
                                                           const T* __restrict__ B, int64_t ldb, int64_t sb,
                                                           T* __restrict__ C, int64_t ldc, int64_t sc, int m, int n,
                                                           int k, int ta, int tb, T alpha, T beta, int64_t count) {
  const int64_t p = blockIdx.x * 256ll + threadIdx.x;
  if (p >= count) return;
  T a[4][4], b[4][4];
  const bool reads = k > 0 && alpha != T(0);
  const T* Ap = A + p * sa;
  const T* Bp = B + p * sb;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int ic = min(i, m - 1), kc = min(kk, max(k - 1, 0));
      const T va = reads ? Ap[ta ? kc + (int64_t)ic * lda : ic + (int64_t)kc * lda] : T(0);
      a[i][kk] = (i < m && kk < k) ? va : T(0);
    }
#pragma unroll
  for (int kk = 0; kk < 4; ++kk)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int jc = min(j, n - 1), kc = min(kk, max(k - 1, 0));
      const T vb = reads ? Bp[tb ? jc + (int64_t)kc * ldb : kc + (int64_t)jc * ldb] : T(0);
      b[kk][j] = (j < n && kk < k) ? vb : T(0);
    }
  T* Cp = C + p * sc;
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (i >= m || j >= n) continue;
      T acc = T(0);
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) acc = fma_t(a[i][kk], b[kk][j], acc);
      T* dst = Cp + i + (int64_t)j * ldc;
      *dst = beta == T(0) ? alpha * acc : fma_t(alpha, acc, beta * *dst);
    }
}

// Dense N x N x N products, N = 2 or 4 (Matrix2f/2d/4f/4d arrays: op N,
// leading dimensions N, operands aligned to a column): one thread each,
// every column of A, B and C moved as one vector access (two 16-byte ones
// for FP64 N = 4) instead of one access per element.  100,000 FP32 4x4
// products: 7.1 against 18.6 us (kernel events); FP64: 12.1 against
// 21.9 us.  The k sum runs in the same order as in tiny_batched_kernel, so
// the results are identical.
template <typename T, int N>
__global__ void __launch_bounds__(256) dense_batched_kernel(const T* __restrict__ A, int64_t sa,
                                                            const T* __restrict__ B, int64_t sb, T* __restrict__ C,
                                                            int64_t sc, T alpha, T beta, int64_t count) {
  using V = T __attribute__((ext_vector_type(N)));
  const int64_t p = blockIdx.x * 256ll + threadIdx.x;
  if (p >= count) return;
  const V* Ap = reinterpret_cast<const V*>(A + p * sa);
  const V* Bp = reinterpret_cast<const V*>(B + p * sb);
  V* Cp = reinterpret_cast<V*>(C + p * sc);
  V a[N], b[N];
#pragma unroll
  for (int j = 0; j < N; ++j) a[j] = Ap[j];
#pragma unroll
  for (int j = 0; j < N; ++j) b[j] = Bp[j];
  V c[N];
  if (beta != T(0)) {
#pragma unroll
    for (int j = 0; j < N; ++j) c[j] = Cp[j];
  }
#pragma unroll
  for (int j = 0; j < N; ++j) {
    V acc;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T s = T(0);
#pragma unroll
      for (int kk = 0; kk < N; ++kk) s = fma_t(a[kk][i], b[j][kk], s);
      acc[i] = beta == T(0) ? alpha * s : fma_t(alpha, s, beta * c[j][i]);
    }
    Cp[j] = acc;
  }
}

// Dense 3 x 3 x 3 products (Matrix3f / Matrix3d arrays: op N, leading
// dimensions 3, any alignment): one thread each, the nine elements of each
// operand loaded as one contiguous run (the compiler merges them into wide
// loads) instead of sixteen clamped loads (100,000 FP32 products: 6.5
// against 9.4 us, FP64 7.7 against 11.5 us); same k order as
// tiny_batched_kernel.
template <typename T, int N>
__global__ void __launch_bounds__(256) dense_run_batched_kernel(const T* __restrict__ A, int64_t sa,
                                                                const T* __restrict__ B, int64_t sb,
                                                                T* __restrict__ C, int64_t sc, T alpha, T beta,
                                                                int64_t count) {
  const int64_t p = blockIdx.x * 256ll + threadIdx.x;
  if (p >= count) return;
  const T* Ap = A + p * sa;
  const T* Bp = B + p * sb;
  T* Cp = C + p * sc;
  T a[N * N], b[N * N];
#pragma unroll
  for (int e = 0; e < N * N; ++e) a[e] = Ap[e];
#pragma unroll
  for (int e = 0; e < N * N; ++e) b[e] = Bp[e];
  T c[N * N];
  if (beta != T(0)) {
#pragma unroll
    for (int e = 0; e < N * N; ++e) c[e] = Cp[e];
  }
#pragma unroll
  for (int j = 0; j < N; ++j)
#pragma unroll
    for (int i = 0; i < N; ++i) {
      T s = T(0);
#pragma unroll
      for (int kk = 0; kk < N; ++kk) s = fma_t(a[kk * N + i], b[j * N + kk], s);
      Cp[j * N + i] = beta == T(0) ? alpha * s : fma_t(alpha, s, beta * c[j * N + i]);
    }
}

template <int N>
static void launch_dense_t(eh_dtype dtype, const void* A, int64_t sa, const void* B, int64_t sb, void* C, int64_t sc,
                           const void* alpha, const void* beta, int64_t count, unsigned blocks, hipStream_t s) {
  if (dtype == EH_F64)
    hipLaunchKernelGGL((dense_batched_kernel<double, N>), dim3(blocks), dim3(256), 0, s, static_cast<const double*>(A),
                       sa, static_cast<const double*>(B), sb, static_cast<double*>(C), sc,
                       *static_cast<const double*>(alpha), *static_cast<const double*>(beta), count);
  else
    hipLaunchKernelGGL((dense_batched_kernel<float, N>), dim3(blocks), dim3(256), 0, s, static_cast<const float*>(A),
                       sa, static_cast<const float*>(B), sb, static_cast<float*>(C), sc,
                       *static_cast<const float*>(alpha), *static_cast<const float*>(beta), count);
}

// Products of at most 16 x 16 with depth at most 4 * KS (KS = 4, 8 or 16),
// one wave each: a single 16 x 16 MFMA block (v_mfma_f32_16x16x4_f32 /
// v_mfma_f64_16x16x4_f64) holds C^T = op(B)^T op(A)^T, so consecutive lanes
// store consecutive rows of C.  Lane l feeds row / column l % 16 at depth
// k0 + l / 16 of each 4-deep step; all KS steps are loaded up front from
// clamped addresses inside the product and zeroed by selects past m, n or k,
// so no load waits on another and none leaves the product.  A 32 x 32 tile
// of the grouped kernel would spend 3/4 of its MFMAs and half its loads on
// zeros for a 16 x 16 product: 20,000 FP64 16^3 products take 23.6 us here
// against 30.8 us there (kernel events; 29.0 against 40.6 us per call, which
// also saves the work-list kernel), FP32 18.7 against 24.9 us, 50,000 FP64
// 8^3 30.2 against 47.1 us.  Deeper products stay on the grouped kernel,
// whose 16-byte loads of k-contiguous slabs win there (FP32 beyond depth
// 32, FP64 beyond 64): a loop over passes of 64 here measured 48% slower on
// FP32 16x16x256 (225 against 152 us), 4% faster on FP64.
template <typename T>
struct Small16;
template <>
struct Small16<float> {
  using acc_t = floatx4;
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int col(int r, int q) { return 4 * q + r; }  // D row (a column of C) of accumulator r
};
template <>
struct Small16<double> {
  using acc_t = doublex4;
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int col(int r, int q) { return q + 4 * r; }
};

template <typename T, bool TA, bool TB, int KS>
__global__ void __launch_bounds__(256) small_batched_kernel(const T* __restrict__ A, int64_t lda, int64_t sa,
                                                            const T* __restrict__ B, int64_t ldb, int64_t sb,
                                                            T* __restrict__ C, int64_t ldc, int64_t sc, int m, int n,
                                                            int k, T alpha, T beta, int64_t count) {
  using S = Small16<T>;
  const int64_t p = blockIdx.x * 4ll + (threadIdx.x >> 6);
  if (p >= count) return;
  const int lane = threadIdx.x & 63, r16 = lane & 15, q = lane >> 4;
  const T* Ap = A + p * sa;
  const T* Bp = B + p * sb;
  const int ic = min(r16, m - 1), jc = min(r16, n - 1);
  T a[KS], b[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    const int kk = 4 * s + q, kc = min(kk, k - 1);
    const T va = TA ? Ap[kc + (int64_t)ic * lda] : Ap[ic + (int64_t)kc * lda];
    const T vb = TB ? Bp[jc + (int64_t)kc * ldb] : Bp[kc + (int64_t)jc * ldb];
    a[s] = (r16 < m && kk < k) ? va : T(0);
    b[s] = (r16 < n && kk < k) ? vb : T(0);
  }
  typename S::acc_t acc;
#pragma unroll
  for (int r = 0; r < 4; ++r) acc[r] = T(0);
#pragma unroll
  for (int s = 0; s < KS; ++s) acc = S::mma(b[s], a[s], acc);
  const bool row_in = r16 < m;
  T* Cp = C + p * sc + r16;
  if (beta != T(0)) {  // every C value of the lane read before the first store
    T c[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int j = S::col(r, q);
      c[r] = row_in && j < n ? Cp[(int64_t)j * ldc] : T(0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] = fma_t(alpha, acc[r], beta * c[r]);
  } else {
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] *= alpha;
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int j = S::col(r, q);
    if (row_in && j < n) Cp[(int64_t)j * ldc] = acc[r];
  }
}

template <typename T, int KS>
static void launch_small_t(bool ta, bool tb, const void* A, int64_t lda, int64_t sa, const void* B, int64_t ldb,
                           int64_t sb, void* C, int64_t ldc, int64_t sc, int m, int n, int k, T alpha, T beta,
                           int64_t count, hipStream_t s) {
  const dim3 grid((unsigned)((count + 3) / 4)), block(256);
  const T* a = static_cast<const T*>(A);
  const T* b = static_cast<const T*>(B);
  T* c = static_cast<T*>(C);
  if (!ta && !tb)
    hipLaunchKernelGGL((small_batched_kernel<T, false, false, KS>), grid, block, 0, s, a, lda, sa, b, ldb, sb, c, ldc,
                       sc, m, n, k, alpha, beta, count);
  else if (ta && !tb)
    hipLaunchKernelGGL((small_batched_kernel<T, true, false, KS>), grid, block, 0, s, a, lda, sa, b, ldb, sb, c, ldc,
                       sc, m, n, k, alpha, beta, count);
  else if (!ta && tb)
    hipLaunchKernelGGL((small_batched_kernel<T, false, true, KS>), grid, block, 0, s, a, lda, sa, b, ldb, sb, c, ldc,
                       sc, m, n, k, alpha, beta, count);
  else
    hipLaunchKernelGGL((small_batched_kernel<T, true, true, KS>), grid, block, 0, s, a, lda, sa, b, ldb, sb, c, ldc,
                       sc, m, n, k, alpha, beta, count);
}

}  // namespace eh

using namespace eh;

extern "C" eh_status eh_gemm_batched_strided(eh_dtype dtype, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k,
                                             const void* alpha, const void* A, int64_t lda, int64_t stride_a,
                                             const void* B, int64_t ldb, int64_t stride_b, const void* beta, void* C,
                                             int64_t ldc, int64_t stride_c, int64_t count, void* stream) {
  eh_status st = check_gemm_args(dtype, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  if (st != EH_SUCCESS) return st;
  if (count < 0 || stride_a < 0 || stride_b < 0 || stride_c < 0)
    return fail(EH_ERR_INVALID, "count and strides must be >= 0");
  if (count > 1 && m > 0 && n > 0 && stride_c < (n - 1) * ldc + m)
    return fail(EH_ERR_INVALID, "stride_c too small: results of different products would overlap");
  if (m == 0 || n == 0 || count == 0) return EH_SUCCESS;
  if (count > INT32_MAX) return fail(EH_ERR_UNSUPPORTED, "count must fit int32");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (m <= 4 && n <= 4 && k <= 4) {  // fixed-size products up to 4 x 4: one thread each
    const unsigned blocks = (unsigned)((count + 255) / 256);
    KernelTimer t(EH_KC_GROUPED, s);
    // dense N x N arrays (N = 2, 4) with every column aligned to its size
    const int es = dtype == EH_F64 ? 8 : 4;
    const int64_t col = m * es;
    const bool aligned = ((uintptr_t)A % col | (uintptr_t)B % col | (uintptr_t)C % col) == 0 &&
                         (stride_a * es) % col == 0 && (stride_b * es) % col == 0 && (stride_c * es) % col == 0;
    const double alv = dtype == EH_F64 ? *static_cast<const double*>(alpha) : *static_cast<const float*>(alpha);
    if ((m == 2 || m == 4) && n == m && k == m && opa == EH_OP_N && opb == EH_OP_N && lda == m && ldb == m &&
        ldc == m && aligned && alv != 0.0) {
      if (m == 4)
        launch_dense_t<4>(dtype, A, stride_a, B, stride_b, C, stride_c, alpha, beta, count, blocks, s);
      else
        launch_dense_t<2>(dtype, A, stride_a, B, stride_b, C, stride_c, alpha, beta, count, blocks, s);
      hipError_t e = hipGetLastError();
      return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "dense batched launch");
    }
    if (m == 3 && n == 3 && k == 3 && opa == EH_OP_N && opb == EH_OP_N && lda == 3 && ldb == 3 && ldc == 3 &&
        alv != 0.0) {
      if (dtype == EH_F64)
        hipLaunchKernelGGL((dense_run_batched_kernel<double, 3>), dim3(blocks), dim3(256), 0, s,
                           static_cast<const double*>(A), stride_a, static_cast<const double*>(B), stride_b,
                           static_cast<double*>(C), stride_c, *static_cast<const double*>(alpha),
                           *static_cast<const double*>(beta), count);
      else
        hipLaunchKernelGGL((dense_run_batched_kernel<float, 3>), dim3(blocks), dim3(256), 0, s,
                           static_cast<const float*>(A), stride_a, static_cast<const float*>(B), stride_b,
                           static_cast<float*>(C), stride_c, *static_cast<const float*>(alpha),
                           *static_cast<const float*>(beta), count);
      hipError_t e = hipGetLastError();
      return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "dense 3x3 batched launch");
    }
    if (dtype == EH_F64)
      hipLaunchKernelGGL(tiny_batched_kernel<double>, dim3(blocks), dim3(256), 0, s, static_cast<const double*>(A), lda,
                         stride_a, static_cast<const double*>(B), ldb, stride_b, static_cast<double*>(C), ldc, stride_c,
                         (int)m, (int)n, (int)k, (int)(opa != EH_OP_N), (int)(opb != EH_OP_N),
                         *static_cast<const double*>(alpha), *static_cast<const double*>(beta), count);
    else
      hipLaunchKernelGGL(tiny_batched_kernel<float>, dim3(blocks), dim3(256), 0, s, static_cast<const float*>(A), lda,
                         stride_a, static_cast<const float*>(B), ldb, stride_b, static_cast<float*>(C), ldc, stride_c,
                         (int)m, (int)n, (int)k, (int)(opa != EH_OP_N), (int)(opb != EH_OP_N),
                         *static_cast<const float*>(alpha), *static_cast<const float*>(beta), count);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "tiny batched launch");
  }
  const int es = dtype == EH_F64 ? 8 : 4;
  const double al = dtype == EH_F64 ? *static_cast<const double*>(alpha) : *static_cast<const float*>(alpha);
  const double be = dtype == EH_F64 ? *static_cast<const double*>(beta) : *static_cast<const float*>(beta);
  const bool ta = opa != EH_OP_N, tb = opb != EH_OP_N;
  const bool reads = k > 0 && al != 0.0;
  // one wave per product up to 16 x 16: FP64 to depth 64, FP32 to depth 32,
  // in one pass of 4, 8 or 16 four-deep steps (a pass with steps past the
  // depth still loads them: FP64 16x16x32 took 65.7 us with 16 steps, 37.5
  // with 8; FP32 16x16x32 29.5 us against 33.4 through the grouped kernel,
  // 16x16x64 with 16 steps 50 against 46 us there)
  if (reads && m <= 16 && n <= 16 && k <= (dtype == EH_F64 ? 64 : 32)) {
    KernelTimer t(EH_KC_GROUPED, s);
    if (dtype == EH_F64) {
      const double a = *static_cast<const double*>(alpha), b = *static_cast<const double*>(beta);
      if (k <= 16)
        launch_small_t<double, 4>(ta, tb, A, lda, stride_a, B, ldb, stride_b, C, ldc, stride_c, (int)m, (int)n, (int)k,
                                  a, b, count, s);
      else if (k <= 32)
        launch_small_t<double, 8>(ta, tb, A, lda, stride_a, B, ldb, stride_b, C, ldc, stride_c, (int)m, (int)n, (int)k,
                                  a, b, count, s);
      else
        launch_small_t<double, 16>(ta, tb, A, lda, stride_a, B, ldb, stride_b, C, ldc, stride_c, (int)m, (int)n,
                                   (int)k, a, b, count, s);
    } else {
      const float a = *static_cast<const float*>(alpha), b = *static_cast<const float*>(beta);
      if (k <= 16)
        launch_small_t<float, 4>(ta, tb, A, lda, stride_a, B, ldb, stride_b, C, ldc, stride_c, (int)m, (int)n, (int)k,
                                 a, b, count, s);
      else  // k <= 32
        launch_small_t<float, 8>(ta, tb, A, lda, stride_a, B, ldb, stride_b, C, ldc, stride_c, (int)m, (int)n, (int)k,
                                 a, b, count, s);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "small batched launch");
  }
  constexpr uint64_t kMaxBytes = 0x7FFFFFFFull;
  const uint64_t ea = reads ? (ta ? extent_bytes(k, m, lda, es) : extent_bytes(m, k, lda, es)) : 0;
  const uint64_t eb = reads ? (tb ? extent_bytes(n, k, ldb, es) : extent_bytes(k, n, ldb, es)) : 0;
  const int64_t tm = (m + 31) / 32, tn = (n + 31) / 32;
  if (ea > kMaxBytes || eb > kMaxBytes || m > (1 << 30) || n > (1 << 30) || tm * tn > (1ll << 30) ||
      count * tm * tn > (1ll << 40)) {
    // past the grouped kernel's 32-bit offsets: one product at a time
    for (int64_t p = 0; p < count; ++p) {
      st = gemm_dispatch(dtype, opa, opb, m, n, k, alpha, static_cast<const char*>(A) + p * stride_a * es, lda,
                         static_cast<const char*>(B) + p * stride_b * es, ldb, beta,
                         static_cast<char*>(C) + p * stride_c * es, ldc, EH_PREC_PLAIN, -1, stream);
      if (st != EH_SUCCESS) return st;
    }
    return EH_SUCCESS;
  }
  GroupedDesc base{};
  base.A = A;
  base.B = B;
  base.C = C;
  base.ldc = ldc;
  base.alpha = reads ? al : 0.0;
  base.beta = be;
  base.bytesA = (uint32_t)ea;
  base.bytesB = (uint32_t)eb;
  base.lda = (uint32_t)std::min<int64_t>(lda, kMaxBytes);
  base.ldb = (uint32_t)std::min<int64_t>(ldb, kMaxBytes);
  base.dtype = dtype;
  base.ta = ta;
  base.tb = tb;
  base.m = (int32_t)m;
  base.n = (int32_t)n;
  base.k = reads ? (int32_t)k : 0;
  base.tiles_m = (int32_t)tm;
  base.tiles_n = (int32_t)tn;
  const int64_t T = tm * tn, tiles = count * T;
  // Schedule by tiles per product: products of a few tiles share no panels
  // across a block, so smaller workgroups (less waiting at the block
  // barrier) win.  Measured on batches of one shape: 4 waves from 16^3 to
  // 64^3 (1 to 4 tiles, 36 vs 40 us for 20,000 FP64 16^3, a shape that now
  // takes the one-wave path above), 8 waves at 100^3 (16 tiles), 16 waves
  // at 200^3 (49 tiles).
  const int sched = T < 8 ? 2 : (T < 32 ? 1 : 0);
  int64_t groups = std::min<int64_t>(resident_groups(sched), (tiles + kSchedWaves[sched] - 1) / kSchedWaves[sched]);
  if (groups > 8) groups = (groups + 7) / 8 * 8;
  groups = std::max<int64_t>(groups, 1);
  const int64_t W = groups;
  // work list in the per-stream workspace: descriptors, tile starts, workgroup runs
  const size_t bd = (size_t)count * sizeof(GroupedDesc), bb = (size_t)(count + 1) * sizeof(int64_t),
               bw = (size_t)(W + 1) * sizeof(int64_t), bp = (size_t)W * sizeof(int32_t);
  void* ws = nullptr;
  WorkspaceLease lease;
  st = workspace_get(s, bd + bb + bw + bp, &lease, &ws);
  if (st != EH_SUCCESS) return st;
  char* p = static_cast<char*>(ws);
  GroupedDesc* d = reinterpret_cast<GroupedDesc*>(p);
  int64_t* begin = reinterpret_cast<int64_t*>(p + bd);
  int64_t* run = reinterpret_cast<int64_t*>(p + bd + bb);
  int32_t* run_prod = reinterpret_cast<int32_t*>(p + bd + bb + bw);
  const int64_t nb = std::min<int64_t>((std::max(count, W) + 256) / 256, 65536);
  hipLaunchKernelGGL(strided_build_kernel, dim3((unsigned)nb), dim3(256), 0, s, base, stride_a * es, stride_b * es,
                     stride_c * es, count, T, W, d, begin, run, run_prod);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "batched build launch");
  {
    KernelTimer t(EH_KC_GROUPED, s);
    e = launch_grouped(sched, groups, d, begin, run, run_prod, s);
  }
  return e == hipSuccess ? EH_SUCCESS : hip_fail(e, "grouped kernel launch");
}

struct eh_grouped {
  std::vector<eh_gemm_problem> host;   // as given (for the per-product path)
  std::vector<eh_gemm_problem> large;  // products past 32-bit buffer offsets
  std::vector<GroupedDesc> desc;       // the kernel's work list (host copy)
  std::vector<int64_t> begin, cost_begin;
  GroupedDesc* d_desc = nullptr;
  int64_t* d_begin = nullptr;
  int64_t* d_run = nullptr;
  int32_t* d_run_prod = nullptr;
  int64_t count = 0;   // products in the kernel's work list
  int64_t tiles = 0;
  int schedule = 0;
  int64_t groups = 0;  // workgroups launched (kSchedWaves[schedule] waves each)
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  std::mutex mu;
};

static eh_status grouped_enqueue(eh_grouped* g, hipStream_t s) {
  if (g->count > 0) {
    KernelTimer t(EH_KC_GROUPED, s);
    hipError_t e = launch_grouped(g->schedule, g->groups, g->d_desc, g->d_begin, g->d_run, g->d_run_prod, s);
    if (e != hipSuccess) return hip_fail(e, "grouped kernel launch");
  }
  for (const eh_gemm_problem& q : g->large) {
    // one 128x128 tile configuration without split-K: no workspace, so the
    // launch can be captured in a graph
    const float af = (float)q.alpha, bf = (float)q.beta;
    const void* al = q.dtype == EH_F32 ? (const void*)&af : (const void*)&q.alpha;
    const void* be = q.dtype == EH_F32 ? (const void*)&bf : (const void*)&q.beta;
    eh_status st = gemm_dispatch(q.dtype, q.opa, q.opb, q.m, q.n, q.k, al, q.A, q.lda, q.B, q.ldb, be, q.C, q.ldc,
                                 EH_PREC_PLAIN, 3 * kNumSplits, s);
    if (st != EH_SUCCESS) return st;
  }
  return EH_SUCCESS;
}

// Runs of schedule `sched` for g's work list, uploaded to the device.  Every
// run holds one operand type: each type's products get runs in proportion
// to their cost (at least one), cut so that the runs of a type carry equal
// cost (run v starts at the first tile whose starting cost reaches the
// run's share).  A captured graph of the previous runs is dropped.
static eh_status grouped_build_runs(eh_grouped* g, int sched) {
  const auto& desc = g->desc;
  const auto& begin = g->begin;
  const auto& cost_begin = g->cost_begin;
  const int G = kSchedWaves[sched];
  const int64_t groups = std::max<int64_t>(1, std::min<int64_t>(resident_groups(sched), (g->tiles + G - 1) / G));
  auto type_of = [](const GroupedDesc& x) { return x.dtype * 4 + x.ta * 2 + x.tb; };
  std::vector<int64_t> seg{0};
  for (int64_t p = 1; p < g->count; ++p)
    if (type_of(desc[p]) != type_of(desc[p - 1])) seg.push_back(p);
  seg.push_back(g->count);
  const int64_t total_cost = cost_begin[g->count];
  std::vector<int64_t> seg_runs(seg.size() - 1);
  int64_t W = 0;
  for (size_t i = 0; i + 1 < seg.size(); ++i) {
    const int64_t c = cost_begin[seg[i + 1]] - cost_begin[seg[i]];
    seg_runs[i] = std::max<int64_t>(1, (int64_t)((__int128)groups * c / std::max<int64_t>(total_cost, 1)));
    W += seg_runs[i];
  }
  // launched workgroups: whole XCD rows for the remap; the extra runs are empty
  const int64_t WL = W > 8 ? (W + 7) / 8 * 8 : W;
  std::vector<int64_t> run(WL + 1, g->tiles);
  std::vector<int32_t> run_prod(WL, (int32_t)(g->count - 1));
  int64_t v = 0;
  for (size_t i = 0; i + 1 < seg.size(); ++i) {
    const int64_t ps = seg[i], pe = seg[i + 1], cb = cost_begin[ps], ce = cost_begin[pe], R = seg_runs[i];
    int64_t p = ps;
    for (int64_t r = 0; r < R; ++r, ++v) {
      const int64_t target = cb + (int64_t)((__int128)(ce - cb) * r / R);
      while (p < pe - 1 && cost_begin[p + 1] <= target) ++p;
      const int64_t c = tile_cost(desc[p]);
      int64_t t = begin[p] + (target - cost_begin[p] + c - 1) / c;
      t = std::min(t, begin[p + 1]);
      if (r > 0) t = std::max(t, run[v - 1]);
      run[v] = t;
      // product holding tile t (the next product of the type when t ends
      // this one; a run starting at the type's end is empty)
      int64_t q = p;
      while (q < pe - 1 && begin[q + 1] <= t) ++q;
      run_prod[v] = (int32_t)q;
    }
  }
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  g->exec = nullptr;
  g->graph = nullptr;
  if (g->d_run) (void)hipFree(g->d_run);
  if (g->d_run_prod) (void)hipFree(g->d_run_prod);
  g->d_run = nullptr;
  g->d_run_prod = nullptr;
  hipError_t e = hipMalloc(&g->d_run, sizeof(int64_t) * (WL + 1));
  if (e == hipSuccess) e = hipMalloc(&g->d_run_prod, sizeof(int32_t) * WL);
  if (e == hipSuccess) e = hipMemcpy(g->d_run, run.data(), sizeof(int64_t) * (WL + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(g->d_run_prod, run_prod.data(), sizeof(int32_t) * WL, hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail(EH_ERR_ALLOC, std::string("grouped runs: ") + hipGetErrorString(e));
  g->schedule = sched;
  g->groups = WL;
  return EH_SUCCESS;
}

// Time every schedule on this batch and keep the fastest.  The timed runs
// write to a scratch copy of the results (each product's C as its own m x n
// block, beta = 0), so the caller's C is untouched; batches whose results
// exceed kTuneScratchMax keep the default schedule.
static eh_status grouped_tune(eh_grouped* g) {
  constexpr size_t kTuneScratchMax = size_t(1) << 30;
  std::vector<GroupedDesc> td = g->desc;
  size_t bytes = 0;
  for (auto& x : td) {
    const size_t es = x.dtype == EH_F64 ? 8 : 4;
    bytes = (bytes + 255) / 256 * 256;
    x.C = reinterpret_cast<void*>(bytes);  // offset, rebased below
    x.ldc = x.m;
    x.beta = 0.0;
    bytes += (size_t)x.m * x.n * es;
  }
  if (bytes > kTuneScratchMax) return EH_SUCCESS;
  char* scratch = nullptr;
  GroupedDesc* d_td = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&scratch, std::max<size_t>(bytes, 256));
  if (e == hipSuccess) e = hipMalloc(&d_td, sizeof(GroupedDesc) * td.size());
  if (e == hipSuccess) {
    for (auto& x : td) x.C = scratch + reinterpret_cast<size_t>(x.C);
    e = hipMemcpy(d_td, td.data(), sizeof(GroupedDesc) * td.size(), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  eh_status st = e == hipSuccess ? EH_SUCCESS : hip_fail(e, "grouped tuning setup");
  int best = 0;
  float best_ms = 0.0f;
  for (int sched = 0; sched < kNumSchedules && st == EH_SUCCESS; ++sched) {
    st = grouped_build_runs(g, sched);
    std::vector<float> t;
    for (int r = 0; r < 4 && st == EH_SUCCESS; ++r) {  // one warm-up, median of three
      (void)hipEventRecord(e0, s);
      e = launch_grouped(sched, g->groups, d_td, g->d_begin, g->d_run, g->d_run_prod, s);
      if (e == hipSuccess) e = hipEventRecord(e1, s);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) st = hip_fail(e, "grouped tuning run");
      if (r > 0) t.push_back(ms);
    }
    if (st != EH_SUCCESS) break;
    std::sort(t.begin(), t.end());
    if (sched == 0 || t[1] < best_ms) {
      best = sched;
      best_ms = t[1];
    }
  }
  if (e1) (void)hipEventDestroy(e1);
  if (e0) (void)hipEventDestroy(e0);
  if (s) (void)hipStreamDestroy(s);
  if (d_td) (void)hipFree(d_td);
  if (scratch) (void)hipFree(scratch);
  if (st != EH_SUCCESS) return st;
  return grouped_build_runs(g, best);
}

extern "C" {

eh_status eh_grouped_create(const eh_gemm_problem* problems, int64_t count, eh_grouped** out) {
  if (!out) return fail(EH_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (count <= 0 || !problems) return fail(EH_ERR_INVALID, "need at least one product");
  std::vector<GroupedDesc> desc;
  std::vector<eh_gemm_problem> large;
  desc.reserve(count);
  constexpr uint64_t kMaxBytes = 0x7FFFFFFFull;  // buffer offsets of valid elements stay below 2^31
  for (int64_t p = 0; p < count; ++p) {
    const eh_gemm_problem& q = problems[p];
    const double one = 1.0;
    eh_status st = check_gemm_args(q.dtype, q.opa, q.opb, q.m, q.n, q.k, &one, q.A, q.lda, q.B, q.ldb, &one, q.C,
                                   q.ldc);
    if (st != EH_SUCCESS) return fail(st, "product " + std::to_string(p) + ": " + eh_last_error());
    if (q.m == 0 || q.n == 0) continue;  // nothing to do
    const bool ta = q.opa != EH_OP_N, tb = q.opb != EH_OP_N;
    const int es = q.dtype == EH_F64 ? 8 : 4;
    const bool reads = q.k > 0 && q.alpha != 0.0;  // alpha == 0: A and B are never read
    const uint64_t ea = reads ? (ta ? extent_bytes(q.k, q.m, q.lda, es) : extent_bytes(q.m, q.k, q.lda, es)) : 0;
    const uint64_t eb = reads ? (tb ? extent_bytes(q.n, q.k, q.ldb, es) : extent_bytes(q.k, q.n, q.ldb, es)) : 0;
    const int64_t tm = (q.m + 31) / 32, tn = (q.n + 31) / 32;
    if (ea > kMaxBytes || eb > kMaxBytes || q.m > (1 << 30) || q.n > (1 << 30) || tm * tn > (1ll << 30)) {
      large.push_back(q);
      continue;
    }
    GroupedDesc d{};
    d.A = q.A;
    d.B = q.B;
    d.C = q.C;
    d.ldc = q.ldc;
    d.alpha = reads ? q.alpha : 0.0;
    d.beta = q.beta;
    d.bytesA = (uint32_t)ea;
    d.bytesB = (uint32_t)eb;
    d.lda = (uint32_t)std::min<int64_t>(q.lda, kMaxBytes);
    d.ldb = (uint32_t)std::min<int64_t>(q.ldb, kMaxBytes);
    d.dtype = q.dtype;
    d.ta = ta;
    d.tb = tb;
    d.m = (int32_t)q.m;
    d.n = (int32_t)q.n;
    d.k = reads ? (int32_t)q.k : 0;
    d.tiles_m = (int32_t)tm;
    d.tiles_n = (int32_t)tn;
    desc.push_back(d);
  }
  // grouped by operand type (FP32 before FP64), each type deepest first
  auto type_of = [](const GroupedDesc& x) { return x.dtype * 4 + x.ta * 2 + x.tb; };
  std::stable_sort(desc.begin(), desc.end(), [&](const GroupedDesc& x, const GroupedDesc& y) {
    if (type_of(x) != type_of(y)) return type_of(x) < type_of(y);
    return x.k > y.k;
  });
  auto* g = new eh_grouped();
  g->host.assign(problems, problems + count);
  g->large = std::move(large);
  g->count = (int64_t)desc.size();
  if (g->count == 0) {
    *out = g;
    return EH_SUCCESS;
  }

  // work list: first tile of each product and cumulative cost
  g->desc = std::move(desc);
  g->begin.assign(g->count + 1, 0);
  g->cost_begin.assign(g->count + 1, 0);
  for (int64_t p = 0; p < g->count; ++p) {
    const int64_t tiles = (int64_t)g->desc[p].tiles_m * g->desc[p].tiles_n;
    g->begin[p + 1] = g->begin[p] + tiles;
    g->cost_begin[p + 1] = g->cost_begin[p] + tiles * tile_cost(g->desc[p]);
  }
  g->tiles = g->begin[g->count];

  hipError_t e = hipMalloc(&g->d_desc, sizeof(GroupedDesc) * g->count);
  if (e == hipSuccess) e = hipMalloc(&g->d_begin, sizeof(int64_t) * (g->count + 1));
  if (e == hipSuccess)
    e = hipMemcpy(g->d_desc, g->desc.data(), sizeof(GroupedDesc) * g->count, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(g->d_begin, g->begin.data(), sizeof(int64_t) * (g->count + 1), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    eh_grouped_destroy(g);
    return fail(EH_ERR_ALLOC, std::string("grouped setup: ") + hipGetErrorString(e));
  }
  eh_status st = grouped_build_runs(g, 0);
  if (st == EH_SUCCESS && tune_mode() == EH_TUNE_TABLE && !fixed_scratch_active()) st = grouped_tune(g);
  if (st != EH_SUCCESS) {
    eh_grouped_destroy(g);
    return st;
  }
  *out = g;
  return EH_SUCCESS;
}

eh_status eh_grouped_run(eh_grouped* g, void* stream) {
  if (!g) return fail(EH_ERR_INVALID, "NULL handle");
  return grouped_enqueue(g, static_cast<hipStream_t>(stream));
}

eh_status eh_grouped_graph_launch(eh_grouped* g, void* stream) {
  if (!g) return fail(EH_ERR_INVALID, "NULL handle");
  if (g->count == 0 && g->large.empty()) return EH_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(g->mu);
  if (!g->exec) {
    // capture on a private stream so the caller's stream is never in capture mode
    hipStream_t cs;
    hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      (void)hipStreamDestroy(cs);
      return hip_fail(e, "hipStreamBeginCapture");
    }
    eh_status st = grouped_enqueue(g, cs);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(cs, &graph);
    (void)hipStreamDestroy(cs);
    if (st != EH_SUCCESS) {
      if (graph) (void)hipGraphDestroy(graph);
      return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(graph);
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
  if (products) *products = g->count + (int64_t)g->large.size();
  return EH_SUCCESS;
}

int eh_grouped_num_schedules(void) { return kNumSchedules; }

eh_status eh_grouped_schedule_info(int schedule, int* waves_per_group, int* band) {
  if (schedule < 0 || schedule >= kNumSchedules) return fail(EH_ERR_INVALID, "schedule out of range");
  if (waves_per_group) *waves_per_group = kSchedWaves[schedule];
  if (band) *band = kSchedBand[schedule];
  return EH_SUCCESS;
}

eh_status eh_grouped_set_schedule(eh_grouped* g, int schedule) {
  if (!g) return fail(EH_ERR_INVALID, "NULL handle");
  if (schedule < 0 || schedule >= kNumSchedules) return fail(EH_ERR_INVALID, "schedule out of range");
  std::lock_guard<std::mutex> lk(g->mu);
  if (g->count == 0) {
    g->schedule = schedule;
    return EH_SUCCESS;
  }
  return grouped_build_runs(g, schedule);
}

eh_status eh_grouped_get_schedule(eh_grouped* g, int* schedule) {
  if (!g || !schedule) return fail(EH_ERR_INVALID, "NULL argument");
  *schedule = g->schedule;
  return EH_SUCCESS;
}

eh_status eh_grouped_destroy(eh_grouped* g) {
  if (!g) return EH_SUCCESS;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->d_desc) (void)hipFree(g->d_desc);
  if (g->d_begin) (void)hipFree(g->d_begin);
  if (g->d_run) (void)hipFree(g->d_run);
  if (g->d_run_prod) (void)hipFree(g->d_run_prod);
  delete g;
  return EH_SUCCESS;
}

}  // extern "C"
