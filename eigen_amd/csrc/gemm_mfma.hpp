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

// A triangular update (tri != 0) of a square result with square tiles
// launches only the tiles on or beside the triangle (tri_tile_coords);
// otherwise the full grid runs and tiles off the triangle return at once.
template <int BM, int BN, typename T>
__host__ __device__ inline bool tri_compact(const GemmParams<T>& p) {
  return p.tri != 0 && BM == BN && p.M == p.N;
}
template <int BM, int BN, typename T>
inline int64_t grid_tiles(const GemmParams<T>& p) {
  const int64_t ntm = (p.M + BM - 1) / BM, ntn = (p.N + BN - 1) / BN;
  return tri_compact<BM, BN>(p) ? ntm * (ntm + 1) / 2 : ntm * ntn;
}

// LDS bytes of one tile body: two stages of (BM + BN) rows of 128 bytes.
template <int BM, int BN>
constexpr int tile_smem_bytes() {
  return 2 * (BM + BN) * 128;
}

// One BM x BN tile of C at (m0, n0) over k in [kbeg, kend): the whole body of
// a product workgroup (gemm_mfma_kernel below).  z / partial: split-K
// slice and whether the result goes to the partial buffer p.W.  smem holds
// tile_smem_bytes<BM, BN>(); the caller synchronises before reusing it.
template <typename T, int BM, int BN, int WGM, int WGN, bool TA, bool TB, int RS>
__device__ __forceinline__ void gemm_tile(GemmParams<T> p, int m0, int n0, int kbeg, int kend, int z,
                                          bool partial, unsigned char* __restrict__ smem) {
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
  static_assert(RS == 1 || RS == 2, "one or two register stages");

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wid = tid >> 6;
  const int wm = wid % WGM, wn = wid / WGM;
  const int nkt = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;

  constexpr uint32_t SZ = sizeof(T);
  // A rows-contiguous operand (k is its column index) gets a buffer range
  // that ends at column kend, so every element with k >= kend reads zero;
  // k-contiguous operands are masked in registers instead.
  const uint32_t bytesA = TA ? p.bytesA : (uint32_t)min((uint64_t)p.bytesA, (uint64_t)kend * p.lda * SZ);
  const uint32_t bytesB = !TB ? p.bytesB : (uint32_t)min((uint64_t)p.bytesB, (uint64_t)kend * p.ldb * SZ);
  const __amdgpu_buffer_rsrc_t rA = make_rsrc(p.A, bytesA);
  const __amdgpu_buffer_rsrc_t rB = make_rsrc(p.B, bytesB);

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

  // One BK slice of both operands in registers.  No load depends on data or
  // on the slice index, so the code between loads and their use has no
  // branches and each wait covers exactly the loads it needs.
  struct Stage {
    u32x4 a[A_CH], b[B_CH];
  };
  // k-contiguous chunk: one 16-byte load (buffer loads need 4-byte alignment
  // only); rows-contiguous chunk: EPC loads, one per k, coalesced across the
  // wave.
  auto load_mncont = [&](__amdgpu_buffer_rsrc_t rs, uint32_t vo, int k0, int cu, int64_t ld) -> u32x4 {
    T t[EPC];
#pragma unroll
    for (int e = 0; e < EPC; ++e)
      t[e] = buf_load(rs, vo, (uint32_t)((uint64_t)(k0 + cu * EPC + e) * ld * SZ), (T*)nullptr);
    return *reinterpret_cast<u32x4*>(t);
  };
  // slice s (clamped to the last one: a prefetch past the end re-reads it)
  auto load = [&](Stage& st, int s) {
    const int k0 = kbeg + min(s, nkt - 1) * BK;
#pragma unroll
    for (int j = 0; j < A_CH; ++j) {
      const int q = tid + NT * j;
      if constexpr (TA)
        st.a[j] = buf_load16(rA, voA[j], (uint32_t)k0 * SZ);
      else
        st.a[j] = load_mncont(rA, voA[j], k0, __builtin_amdgcn_readfirstlane(q / BM), p.lda);
    }
#pragma unroll
    for (int j = 0; j < B_CH; ++j) {
      const int q = tid + NT * j;
      if constexpr (!TB)
        st.b[j] = buf_load16(rB, voB[j], (uint32_t)k0 * SZ);
      else
        st.b[j] = load_mncont(rB, voB[j], k0, __builtin_amdgcn_readfirstlane(q / BN), p.ldb);
    }
  };
  // k-contiguous chunk c of the slice at k0: zero the elements at k >= kend
  // (selects, no branch; a no-op except on the last slice)
  auto mask_kcont = [&](u32x4& v, int k0, int c) {
    T* t = reinterpret_cast<T*>(&v);
#pragma unroll
    for (int e = 0; e < EPC; ++e) t[e] = (k0 + c * EPC + e < kend) ? t[e] : T(0);
  };
  // ---- registers -> LDS buffer -------------------------------------------
  auto store = [&](Stage& st, int s, int buf) {
    const int k0 = kbeg + min(s, nkt - 1) * BK;
    unsigned char* As = smem + buf * (BM + BN) * ROWB;
    unsigned char* Bs = As + BM * ROWB;
#pragma unroll
    for (int j = 0; j < A_CH; ++j) {
      const int q = tid + NT * j;
      const int r = TA ? q / CH : q % BM, c = TA ? q % CH : q / BM;
      if constexpr (TA) mask_kcont(st.a[j], k0, c);
      *reinterpret_cast<u32x4*>(As + r * ROWB + swz8(r, c) * 16) = st.a[j];
    }
#pragma unroll
    for (int j = 0; j < B_CH; ++j) {
      const int q = tid + NT * j;
      const int r = TB ? q % BN : q / CH, c = TB ? q / BN : q % CH;
      if constexpr (!TB) mask_kcont(st.b[j], k0, c);
      *reinterpret_cast<u32x4*>(Bs + r * ROWB + swz8(r, c) * 16) = st.b[j];
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

  // ---- MFMAs over the slice in LDS buffer `buf` ---------------------------
  auto compute = [&](int buf) {
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
  };

  if (nkt > 0) {
    if constexpr (RS == 2) {
      // Two register stages: the loads of slice s + 2 are issued while slice
      // s is multiplied, so every slice has two slices of MFMA time to arrive.
      Stage r0, r1;
      load(r0, 0);
      store(r0, 0, 0);
      load(r1, 1);
      load(r0, 2);
      __syncthreads();
      // at the top of each pair: LDS buffer 0 holds slice kt, r1 slice
      // kt + 1, r0 slice kt + 2
      int kt = 0;
      for (; kt + 2 <= nkt; kt += 2) {
        compute(0);
        store(r1, kt + 1, 1);
        __syncthreads();
        load(r1, kt + 3);
        compute(1);
        store(r0, kt + 2, 0);
        __syncthreads();
        load(r0, kt + 4);
      }
      if (kt < nkt) compute(0);  // odd slice count: the last slice is in buffer 0
    } else {
      Stage r0;
      load(r0, 0);
      store(r0, 0, 0);
      __syncthreads();
      for (int kt = 0; kt < nkt; ++kt) {
        load(r0, kt + 1);
        compute(kt & 1);
        store(r0, kt + 1, (kt + 1) & 1);
        __syncthreads();
      }
    }
  }

  // ---- epilogue -------------------------------------------------------------
#pragma unroll
  for (int i = 0; i < RM; ++i)
#pragma unroll
    for (int j = 0; j < RN; ++j) {
      const int col = n0 + wn * WTN + j * F::NI + F::col(lane);
      if (col >= p.N) continue;
#pragma unroll
      for (int r = 0; r < F::ACC; ++r) {
        const int row = m0 + wm * WTM + i * F::MI + F::row(r, lane);
        if (row >= p.M || (p.tri == 1 && row < col) || (p.tri == 2 && row > col)) continue;
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

template <typename T, int BM, int BN, int WGM, int WGN, bool TA, bool TB, int RS>
__global__ void __launch_bounds__(64 * WGM * WGN) gemm_mfma_kernel(GemmParams<T> p) {
  __shared__ __attribute__((aligned(16))) unsigned char smem[tile_smem_bytes<BM, BN>()];
  const int ntm = (p.M + BM - 1) / BM, ntn = (p.N + BN - 1) / BN;
  int tm, tn;
  if (tri_compact<BM, BN>(p)) {
    // the grid holds only the tiles that meet the triangle (lower: tn <= tm)
    int r, c;
    tri_tile_coords(xcd_remap(blockIdx.x, gridDim.x), ntm, 8, r, c);
    tm = p.tri == 1 ? r : c;
    tn = p.tri == 1 ? c : r;
  } else {
    tile_coords(xcd_remap(blockIdx.x, gridDim.x), ntm, ntn, 8, tm, tn);
  }
  const int m0 = tm * BM, n0 = tn * BN;
  // a tile with no element in the updated triangle has nothing to do
  if ((p.tri == 1 && m0 + BM <= n0) || (p.tri == 2 && n0 + BN <= m0)) return;
  const int z = blockIdx.y;
  const int kbeg = z * p.k_per_split;
  const int kend = min(p.K, kbeg + p.k_per_split);
  gemm_tile<T, BM, BN, WGM, WGN, TA, TB, RS>(p, m0, n0, kbeg, kend, z, gridDim.y > 1, smem);
}

// Deterministic split-K reduction: C = alpha * sum_z W[z] + beta*C.
template <typename T>
__global__ void __launch_bounds__(256) splitk_reduce_kernel(GemmParams<T> p, int splits) {
  const int64_t total = (int64_t)p.M * p.N;
  for (int64_t idx = blockIdx.x * 256ll + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
    const int row = (int)(idx % p.M);
    const int64_t col = idx / p.M;
    if ((p.tri == 1 && row < col) || (p.tri == 2 && row > col)) continue;
    T s = T(0);
    for (int z = 0; z < splits; ++z) s += p.W[(int64_t)z * total + idx];
    T* dst = p.C + row + col * p.ldc;
    *dst = (p.beta == T(0)) ? p.alpha * s : fma_t(p.alpha, s, p.beta * *dst);
  }
}

// C = beta*C on the triangle `tri` (as GemmParams::tri) of C (beta==0 writes
// zeros without reading C).  Used when alpha==0 or k==0: the product term
// vanishes and A, B are not read.
template <typename T>
__global__ void __launch_bounds__(256) scale_c_kernel(T* C, int M, int N, int64_t ldc, T beta, int tri) {
  const int64_t total = (int64_t)M * N;
  for (int64_t idx = blockIdx.x * 256ll + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
    const int row = (int)(idx % M);
    const int64_t col = idx / M;
    if ((tri == 1 && row < col) || (tri == 2 && row > col)) continue;
    T* dst = C + row + col * ldc;
    *dst = (beta == T(0)) ? T(0) : beta * *dst;
  }
}

// Result with an inner stride: C(i,j) = C[i*incc + j*ldc] gets
// alpha*W(i,j) + beta*C(i,j) (W = the unscaled product, m x n contiguous), in
// the epilogue's arithmetic; W == nullptr means the product term vanishes
// (C = beta*C).  beta == 0 never reads C.
template <typename T>
__global__ void __launch_bounds__(256) strided_combine_kernel(const T* W, int M, int N, T* C, int64_t incc,
                                                              int64_t ldc, T alpha, T beta) {
  const int64_t total = (int64_t)M * N;
  for (int64_t idx = blockIdx.x * 256ll + threadIdx.x; idx < total; idx += (int64_t)gridDim.x * 256) {
    const int row = (int)(idx % M);
    const int64_t col = idx / M;
    T* dst = C + row * incc + col * ldc;
    if (W) {
      const T v = W[idx];
      *dst = (beta == T(0)) ? alpha * v : fma_t(alpha, v, beta * *dst);
    } else {
      *dst = (beta == T(0)) ? T(0) : beta * *dst;
    }
  }
}

}  // namespace eh
