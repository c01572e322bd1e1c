// gemm_mfma.hip -- plain FP32 / FP64 MFMA products for gfx950.
//
// One workgroup computes a BM x BN tile of C = alpha*op(A)*op(B) + beta*C
// over a K range (the whole K, or one split of it).  This is the device form
// of Eigen's GEBP (Eigen/src/Core/products/GeneralBlockPanelKernel.h):
//   * gemm_pack_lhs / gemm_pack_rhs (:1893-2478) become the global->LDS
//     staging of one BK-deep slice of op(A) and op(B), double buffered, with
//     every LDS row holding 128 contiguous bytes of k for one row of op(A) /
//     one column of op(B) (an XOR-swizzled, bank-conflict-free packed layout);
//   * the register-blocked micro-kernel (gebp_micro_panel_impl, :1352-1456)
//     becomes a 2x2 grid of waves, each issuing RM x RN MFMA micro-tiles;
//   * the store `C += alpha*acc` (:1446-1455) becomes a fused alpha/beta
//     epilogue (beta == 0 never reads C, matching blas/level3_impl.h:82-87).
// Transposed operands only change which global index is contiguous; the LDS
// layout is the same for all four op combinations.
#include "common.hpp"
#include "kernels.hpp"

namespace eh {

template <typename T, int BM, int BN, int WGM, int WGN, bool TA, bool TB>
__global__ void __launch_bounds__(64 * WGM * WGN) gemm_mfma_kernel(GemmParams<T> p) {
  using F = Mfma<T>;
  constexpr int NT = 64 * WGM * WGN;
  constexpr int ROWB = 128;                // bytes of k per LDS row
  constexpr int BK = ROWB / sizeof(T);     // 32 (f32) or 16 (f64)
  constexpr int EPC = 16 / sizeof(T);      // elements per 16-byte chunk
  constexpr int CH = ROWB / 16;            // chunks per row = 8
  constexpr int WTM = BM / WGM, WTN = BN / WGN;
  constexpr int RM = WTM / F::MI, RN = WTN / F::NI;
  constexpr int KPL = BK / F::KG;          // k values one lane feeds per slice
  constexpr int CPL = KPL / EPC;           // chunks one lane reads per row
  constexpr int A_CH = BM * CH / NT, B_CH = BN * CH / NT;
  static_assert(RM >= 1 && RN >= 1, "wave tile smaller than one MFMA block");
  static_assert(BM * CH % NT == 0 && BN * CH % NT == 0, "tile not divisible by threads");
  static_assert(CPL * EPC == KPL, "lane k range must be whole chunks");

  __shared__ __attribute__((aligned(16))) unsigned char smem[2 * (BM + BN) * ROWB];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wid = tid >> 6;
  const int wm = wid % WGM, wn = wid / WGM;

  const int ntm = (p.M + BM - 1) / BM, ntn = (p.N + BN - 1) / BN;
  int tm, tn;
  tile_coords(xcd_remap(blockIdx.x, gridDim.x), ntm, ntn, 8, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int z = blockIdx.y;
  const int kbeg = z * p.k_per_split;
  const int kend = min(p.K, kbeg + p.k_per_split);
  const int nkt = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;

  u32x4 ra[A_CH], rb[B_CH];

  // ---- global -> registers for one BK slice starting at k0 ----------------
  auto load_a = [&](int k0) {
#pragma unroll
    for (int j = 0; j < A_CH; ++j) {
      const int q = tid + NT * j;
      u32x4 v = {0u, 0u, 0u, 0u};
      if constexpr (TA) {  // op(A)(i,k) = A[k + i*lda]: k contiguous
        const int r = q / CH, c = q % CH;
        const int gi = m0 + r, gk = k0 + c * EPC;
        if (gi < p.M) {
          const T* src = p.A + (int64_t)gi * p.lda + gk;
          if (p.vecA && gk + EPC <= kend) {
            v = *reinterpret_cast<const u32x4*>(src);
          } else {
            T t[EPC];
#pragma unroll
            for (int e = 0; e < EPC; ++e) t[e] = (gk + e < kend) ? src[e] : T(0);
            v = *reinterpret_cast<u32x4*>(t);
          }
        }
      } else {  // op(A)(i,k) = A[i + k*lda]: rows contiguous
        const int r = q % BM, c = q / BM;
        const int gi = m0 + r, gk = k0 + c * EPC;
        T t[EPC];
#pragma unroll
        for (int e = 0; e < EPC; ++e)
          t[e] = (gi < p.M && gk + e < kend) ? p.A[gi + (int64_t)(gk + e) * p.lda] : T(0);
        v = *reinterpret_cast<u32x4*>(t);
      }
      ra[j] = v;
    }
  };
  auto load_b = [&](int k0) {
#pragma unroll
    for (int j = 0; j < B_CH; ++j) {
      const int q = tid + NT * j;
      u32x4 v = {0u, 0u, 0u, 0u};
      if constexpr (!TB) {  // op(B)(k,j) = B[k + j*ldb]: k contiguous
        const int r = q / CH, c = q % CH;
        const int gj = n0 + r, gk = k0 + c * EPC;
        if (gj < p.N) {
          const T* src = p.B + (int64_t)gj * p.ldb + gk;
          if (p.vecB && gk + EPC <= kend) {
            v = *reinterpret_cast<const u32x4*>(src);
          } else {
            T t[EPC];
#pragma unroll
            for (int e = 0; e < EPC; ++e) t[e] = (gk + e < kend) ? src[e] : T(0);
            v = *reinterpret_cast<u32x4*>(t);
          }
        }
      } else {  // op(B)(k,j) = B[j + k*ldb]: columns of op(B) contiguous
        const int r = q % BN, c = q / BN;
        const int gj = n0 + r, gk = k0 + c * EPC;
        T t[EPC];
#pragma unroll
        for (int e = 0; e < EPC; ++e)
          t[e] = (gj < p.N && gk + e < kend) ? p.B[gj + (int64_t)(gk + e) * p.ldb] : T(0);
        v = *reinterpret_cast<u32x4*>(t);
      }
      rb[j] = v;
    }
  };
  // ---- registers -> LDS buffer -------------------------------------------
  auto store_tiles = [&](int buf) {
    unsigned char* As = smem + buf * (BM + BN) * ROWB;
    unsigned char* Bs = As + BM * ROWB;
#pragma unroll
    for (int j = 0; j < A_CH; ++j) {
      const int q = tid + NT * j;
      const int r = TA ? q / CH : q % BM, c = TA ? q % CH : q / BM;
      *reinterpret_cast<u32x4*>(As + r * ROWB + swz8(r, c) * 16) = ra[j];
    }
#pragma unroll
    for (int j = 0; j < B_CH; ++j) {
      const int q = tid + NT * j;
      const int r = TB ? q % BN : q / CH, c = TB ? q / BN : q % CH;
      *reinterpret_cast<u32x4*>(Bs + r * ROWB + swz8(r, c) * 16) = rb[j];
    }
  };

  typename F::acc_t acc[RM][RN];
#pragma unroll
  for (int i = 0; i < RM; ++i)
#pragma unroll
    for (int j = 0; j < RN; ++j)
#pragma unroll
      for (int r = 0; r < F::ACC; ++r) acc[i][j][r] = T(0);

  const int li = lane % F::MI;  // row of the MFMA operand this lane feeds
  const int lg = lane / F::MI;  // k group of this lane

  if (nkt > 0) {
    load_a(kbeg);
    load_b(kbeg);
    store_tiles(0);
    __syncthreads();
    for (int kt = 0; kt < nkt; ++kt) {
      const int buf = kt & 1;
      if (kt + 1 < nkt) {
        load_a(kbeg + (kt + 1) * BK);
        load_b(kbeg + (kt + 1) * BK);
      }
      const unsigned char* As = smem + buf * (BM + BN) * ROWB;
      const unsigned char* Bs = As + BM * ROWB;
#pragma unroll
      for (int cc = 0; cc < CPL; ++cc) {
        const int c = lg * CPL + cc;
        T a[RM][EPC], b[RN][EPC];
#pragma unroll
        for (int i = 0; i < RM; ++i) {
          const int r = wm * WTM + i * F::MI + li;
          u32x4 v = *reinterpret_cast<const u32x4*>(As + r * ROWB + swz8(r, c) * 16);
          const T* t = reinterpret_cast<const T*>(&v);
#pragma unroll
          for (int e = 0; e < EPC; ++e) a[i][e] = t[e];
        }
#pragma unroll
        for (int j = 0; j < RN; ++j) {
          const int r = wn * WTN + j * F::NI + li;
          u32x4 v = *reinterpret_cast<const u32x4*>(Bs + r * ROWB + swz8(r, c) * 16);
          const T* t = reinterpret_cast<const T*>(&v);
#pragma unroll
          for (int e = 0; e < EPC; ++e) b[j][e] = t[e];
        }
#pragma unroll
        for (int e = 0; e < EPC; ++e)
#pragma unroll
          for (int i = 0; i < RM; ++i)
#pragma unroll
            for (int j = 0; j < RN; ++j) F::mma(a[i][e], b[j][e], acc[i][j]);
      }
      if (kt + 1 < nkt) store_tiles(buf ^ 1);
      __syncthreads();
    }
  }

  // ---- epilogue -------------------------------------------------------------
  const bool partial = gridDim.y > 1;
#pragma unroll
  for (int i = 0; i < RM; ++i)
#pragma unroll
    for (int j = 0; j < RN; ++j) {
      const int col = n0 + wn * WTN + j * F::NI + F::col(lane);
      if (col >= p.N) continue;
#pragma unroll
      for (int r = 0; r < F::ACC; ++r) {
        const int row = m0 + wm * WTM + i * F::MI + F::row(r, lane);
        if (row >= p.M) continue;
        if (partial) {
          p.W[(int64_t)z * p.M * p.N + row + (int64_t)col * p.M] = acc[i][j][r];
        } else {
          T* dst = p.C + row + (int64_t)col * p.ldc;
          T v = acc[i][j][r];
          *dst = (p.beta == T(0)) ? p.alpha * v : fma_t(p.alpha, v, p.beta * *dst);
        }
      }
    }
}

// Deterministic split-K reduction: C = alpha * sum_z W[z] + beta*C.
template <typename T>
__global__ void __launch_bounds__(256) splitk_reduce_kernel(GemmParams<T> p, int splits) {
  const int64_t total = (int64_t)p.M * p.N;
  for (int64_t idx = blockIdx.x * 256ll + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
    T s = T(0);
    for (int z = 0; z < splits; ++z) s += p.W[(int64_t)z * total + idx];
    const int row = (int)(idx % p.M);
    const int64_t col = idx / p.M;
    T* dst = p.C + row + col * p.ldc;
    *dst = (p.beta == T(0)) ? p.alpha * s : fma_t(p.alpha, s, p.beta * *dst);
  }
}

// C = beta*C (beta==0 writes zeros without reading C).  Used when alpha==0 or
// k==0: the product term vanishes and A, B are not read.
template <typename T>
__global__ void __launch_bounds__(256) scale_c_kernel(T* C, int M, int N, int64_t ldc, T beta) {
  const int64_t total = (int64_t)M * N;
  for (int64_t idx = blockIdx.x * 256ll + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
    const int row = (int)(idx % M);
    const int64_t col = idx / M;
    T* dst = C + row + col * ldc;
    *dst = (beta == T(0)) ? T(0) : beta * *dst;
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T, int BM, int BN, bool TA, bool TB>
static hipError_t launch_one(const GemmParams<T>& p, int splits, hipStream_t s) {
  const int ntm = (p.M + BM - 1) / BM, ntn = (p.N + BN - 1) / BN;
  dim3 grid(ntm * ntn, splits);
  hipLaunchKernelGGL((gemm_mfma_kernel<T, BM, BN, 2, 2, TA, TB>), grid, dim3(256), 0, s, p);
  return hipGetLastError();
}

template <typename T, int BM, int BN>
static hipError_t launch_tile(const GemmParams<T>& p, bool ta, bool tb, int splits, hipStream_t s) {
  if (!ta && !tb) return launch_one<T, BM, BN, false, false>(p, splits, s);
  if (ta && !tb) return launch_one<T, BM, BN, true, false>(p, splits, s);
  if (!ta && tb) return launch_one<T, BM, BN, false, true>(p, splits, s);
  return launch_one<T, BM, BN, true, true>(p, splits, s);
}

template <typename T>
hipError_t launch_gemm_mfma(const GemmParams<T>& p, int tile, bool ta, bool tb, int splits, hipStream_t s) {
  switch (tile) {
    case 0: return launch_tile<T, 64, 64>(p, ta, tb, splits, s);
    case 1: return launch_tile<T, 64, 128>(p, ta, tb, splits, s);
    case 2: return launch_tile<T, 128, 64>(p, ta, tb, splits, s);
    case 3: return launch_tile<T, 128, 128>(p, ta, tb, splits, s);
    case 4: return launch_tile<T, 128, 256>(p, ta, tb, splits, s);
    case 5: return launch_tile<T, 256, 128>(p, ta, tb, splits, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t launch_splitk_reduce(const GemmParams<T>& p, int splits, hipStream_t s) {
  const int64_t total = (int64_t)p.M * p.N;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(splitk_reduce_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, p, splits);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_scale_c(T* C, int M, int N, int64_t ldc, T beta, hipStream_t s) {
  const int64_t total = (int64_t)M * N;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(scale_c_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, C, M, N, ldc, beta);
  return hipGetLastError();
}

template hipError_t launch_gemm_mfma<float>(const GemmParams<float>&, int, bool, bool, int, hipStream_t);
template hipError_t launch_gemm_mfma<double>(const GemmParams<double>&, int, bool, bool, int, hipStream_t);
template hipError_t launch_splitk_reduce<float>(const GemmParams<float>&, int, hipStream_t);
template hipError_t launch_splitk_reduce<double>(const GemmParams<double>&, int, hipStream_t);
template hipError_t launch_scale_c<float>(float*, int, int, int64_t, float, hipStream_t);
template hipError_t launch_scale_c<double>(double*, int, int, int64_t, double, hipStream_t);

}  // namespace eh
