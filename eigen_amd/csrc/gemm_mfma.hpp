// gemm_mfma.hpp -- kernel templates of the plain FP32 / FP64 MFMA products.
//
// One workgroup computes a BM x BN tile of C = alpha*op(A)*op(B) + beta*C
// over a K range (the whole K, or one split of it).  This is the device form
// of Eigen's GEBP (Eigen/src/Core/products/GeneralBlockPanelKernel.h):
//   * gemm_pack_lhs / gemm_pack_rhs (:1893-2478) become the global->LDS
//     staging of one BK-deep slice of op(A) and op(B), double buffered, with
//     every LDS row holding 128 contiguous bytes of k for one row of op(A) /
//     one column of op(B) (an XOR-swizzled, bank-conflict-free packed layout);
//   * the register-blocked micro-kernel (gebp_micro_panel_impl, :1352-1456)
//     becomes a WGM x WGN grid of waves (2x2, or 4x2 / 2x4 for two waves per
//     SIMD), each issuing RM x RN MFMA micro-tiles;
//   * the store `C += alpha*acc` (:1446-1455) becomes a fused alpha/beta
//     epilogue (beta == 0 never reads C, matching blas/level3_impl.h:82-87).
// Transposed operands only change which global index is contiguous; the LDS
// layout is the same for all four op combinations.
#pragma once
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
  constexpr uint32_t SZ = sizeof(T);
  const __amdgpu_buffer_rsrc_t rA = make_rsrc(p.A, p.bytesA);
  const __amdgpu_buffer_rsrc_t rB = make_rsrc(p.B, p.bytesB);

  // Loop-invariant per-lane byte offsets of every staged chunk; the k advance
  // goes into the wave-uniform soffset.  Rows / columns past M / N read zeros
  // or in-range values that only reach outputs that are never stored.
  uint32_t voA[A_CH], voB[B_CH];
#pragma unroll
  for (int j = 0; j < A_CH; ++j) {
    const int q = tid + NT * j;
    if constexpr (TA) {  // op(A)(i,k) = A[k + i*lda]: k contiguous
      const int r = q / CH, c = q % CH;
      voA[j] = (uint32_t)(((uint64_t)(m0 + r) * p.lda + c * EPC) * SZ);
    } else {  // op(A)(i,k) = A[i + k*lda]: rows contiguous
      voA[j] = (uint32_t)((m0 + q % BM) * SZ);
    }
  }
#pragma unroll
  for (int j = 0; j < B_CH; ++j) {
    const int q = tid + NT * j;
    if constexpr (!TB) {  // op(B)(k,j) = B[k + j*ldb]: k contiguous
      const int r = q / CH, c = q % CH;
      voB[j] = (uint32_t)(((uint64_t)(n0 + r) * p.ldb + c * EPC) * SZ);
    } else {  // op(B)(k,j) = B[j + k*ldb]
      voB[j] = (uint32_t)((n0 + q % BN) * SZ);
    }
  }

  // k-contiguous chunk: one 16-byte load (or EPC element loads when rows are
  // not 16-byte aligned); on the final partial slice, k >= kend is zeroed.
  auto load_kcont = [&](__amdgpu_buffer_rsrc_t rs, uint32_t vo, int k0, int c, bool vec, bool tail) -> u32x4 {
    u32x4 v;
    if (vec) {
      v = buf_load16(rs, vo, (uint32_t)k0 * SZ);
    } else {
      T t[EPC];
#pragma unroll
      for (int e = 0; e < EPC; ++e) t[e] = buf_load(rs, vo + e * SZ, (uint32_t)k0 * SZ, (T*)nullptr);
      v = *reinterpret_cast<u32x4*>(t);
    }
    if (tail) {
      T* t = reinterpret_cast<T*>(&v);
#pragma unroll
      for (int e = 0; e < EPC; ++e)
        if (k0 + c * EPC + e >= kend) t[e] = T(0);
    }
    return v;
  };
  // rows-contiguous chunk: EPC loads, one per k, coalesced across the wave;
  // k >= K lies past the end of the operand and reads zero.
  auto load_mncont = [&](__amdgpu_buffer_rsrc_t rs, uint32_t vo, int k0, int cu, int64_t ld) -> u32x4 {
    T t[EPC];
#pragma unroll
    for (int e = 0; e < EPC; ++e)
      t[e] = buf_load(rs, vo, (uint32_t)((uint64_t)(k0 + cu * EPC + e) * ld * SZ), (T*)nullptr);
    return *reinterpret_cast<u32x4*>(t);
  };

  // ---- global -> registers for one BK slice starting at k0 ----------------
  auto load_a = [&](int k0, bool tail) {
#pragma unroll
    for (int j = 0; j < A_CH; ++j) {
      const int q = tid + NT * j;
      if constexpr (TA)
        ra[j] = load_kcont(rA, voA[j], k0, q % CH, p.vecA, tail);
      else
        ra[j] = load_mncont(rA, voA[j], k0, __builtin_amdgcn_readfirstlane(q / BM), p.lda);
    }
  };
  auto load_b = [&](int k0, bool tail) {
#pragma unroll
    for (int j = 0; j < B_CH; ++j) {
      const int q = tid + NT * j;
      if constexpr (!TB)
        rb[j] = load_kcont(rB, voB[j], k0, q % CH, p.vecB, tail);
      else
        rb[j] = load_mncont(rB, voB[j], k0, __builtin_amdgcn_readfirstlane(q / BN), p.ldb);
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
    load_a(kbeg, kbeg + BK > kend);
    load_b(kbeg, kbeg + BK > kend);
    store_tiles(0);
    __syncthreads();
    for (int kt = 0; kt < nkt; ++kt) {
      const int buf = kt & 1;
      if (kt + 1 < nkt) {
        const int k1 = kbeg + (kt + 1) * BK;
        load_a(k1, k1 + BK > kend);
        load_b(k1, k1 + BK > kend);
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

}  // namespace eh
