// gemm_split.hip -- FP32-accurate sgemm on BF16 MFMA (v_mfma_f32_32x32x16_bf16).
//
// Every FP32 element x of op(A) and op(B) is split, while it is staged from
// registers into LDS, into bf16 terms
//     t0 = bf16(x),  t1 = bf16(x - t0),  t2 = bf16(x - t0 - t1)
// (each residual is exact in FP32).  The product keeps the cross terms
// t_p(A) * t_q(B) with p + q < TERMS -- 6 products for TERMS = 3 (hi*hi,
// hi*mid, mid*hi, hi*lo, lo*hi, mid*mid), 3 for TERMS = 2 -- each an exact
// bf16 x bf16 product summed by the MFMA into one FP32 accumulator, smallest
// terms first within each k slice.  With 3 terms the dropped products are
// below 2^-26 relative, so the result carries FP32-level error while running
// on the BF16 matrix cores.
//
// Rounding of each term is Eigen's float -> bfloat16 round-to-nearest-even
// (Eigen/src/Core/arch/Default/BFloat16.h:434-600) with one deliberate
// exception: a finite x whose bf16 rounding would overflow to infinity keeps
// its truncated leading term instead, so the split stays exact for finite
// inputs up to FLT_MAX.  Non-finite x gives t0 = bf16(x) and zero residuals.
//
// Tiling and staging mirror gemm_mfma.hip (the GEBP packers become LDS
// staging, the micro-kernel becomes MFMA micro-tiles, alpha/beta is fused),
// with the LDS holding TERMS planes of 64-byte bf16 rows per operand.
#include "common.hpp"
#include "kernels.hpp"

namespace eh {

__device__ inline unsigned short bf16_rtne_bits(unsigned int u) {
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)(((u >> 16) & 0x8000u) | 0x7fc0u);
  return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// Leading term: RTNE, except that a finite value rounding up to infinity is
// truncated instead.
__device__ inline unsigned short bf16_lead_bits(unsigned int u) {
  unsigned short h = bf16_rtne_bits(u);
  if ((h & 0x7fffu) == 0x7f80u && (u & 0x7f800000u) != 0x7f800000u) h = (unsigned short)(u >> 16);
  return h;
}

__device__ inline float bf16_to_f32(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

// Split 8 floats into TERMS vectors of 8 bf16.
template <int TERMS>
__device__ inline void split8(const float (&x)[8], u16x8 (&out)[TERMS]) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const unsigned int u = __float_as_uint(x[e]);
    const unsigned short h0 = bf16_lead_bits(u);
    out[0][e] = h0;
    if constexpr (TERMS > 1) {
      const bool finite = (u & 0x7f800000u) != 0x7f800000u;
      const float r1 = finite ? x[e] - bf16_to_f32(h0) : 0.0f;
      const unsigned short h1 = bf16_rtne_bits(__float_as_uint(r1));
      out[1][e] = h1;
      if constexpr (TERMS > 2) {
        const float r2 = r1 - bf16_to_f32(h1);
        out[2][e] = bf16_rtne_bits(__float_as_uint(r2));
      }
    }
  }
}

template <int BM, int BN, int WGM, int WGN, bool TA, bool TB, int TERMS>
__global__ void __launch_bounds__(64 * WGM * WGN) gemm_split_kernel(GemmParams<float> p) {
  constexpr int NT = 64 * WGM * WGN;
  constexpr int BK = 32;                 // k per slice
  constexpr int ROWB = BK * 2;           // bytes of one bf16 plane row
  constexpr int PC = BK / 8;             // 16-byte chunks per plane row (8 bf16 each) = 4
  constexpr int WTM = BM / WGM, WTN = BN / WGN;
  constexpr int RM = WTM / 32, RN = WTN / 32;
  constexpr int A_PC = BM * PC / NT, B_PC = BN * PC / NT;
  constexpr int PLANE = (BM + BN) * ROWB;  // bytes of one term plane (A rows then B rows)
  constexpr int STAGE = TERMS * PLANE;
  static_assert(RM >= 1 && RN >= 1, "wave tile smaller than one MFMA block");
  static_assert(BM * PC % NT == 0 && BN * PC % NT == 0, "tile not divisible by threads");

  __shared__ __attribute__((aligned(16))) unsigned char smem[2 * STAGE];

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

  float ra[A_PC][8], rb[B_PC][8];

  // Load 8 consecutive k of one row of op(X) (X = A or B) into v.
  auto load8 = [&](const float* X, int64_t ld, bool kcont, bool vec, int gi, int rows_valid, int gk,
                   float (&v)[8]) {
    if (kcont) {
      if (gi < rows_valid) {
        const float* src = X + (int64_t)gi * ld + gk;
        if (vec && gk + 8 <= kend) {
          floatx4 lo = *reinterpret_cast<const floatx4*>(src);
          floatx4 hi = *reinterpret_cast<const floatx4*>(src + 4);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            v[e] = lo[e];
            v[e + 4] = hi[e];
          }
        } else {
#pragma unroll
          for (int e = 0; e < 8; ++e) v[e] = (gk + e < kend) ? src[e] : 0.0f;
        }
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = 0.0f;
      }
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e)
        v[e] = (gi < rows_valid && gk + e < kend) ? X[gi + (int64_t)(gk + e) * ld] : 0.0f;
    }
  };

  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int j = 0; j < A_PC; ++j) {
      const int q = tid + NT * j;
      const int r = TA ? q / PC : q % BM, c = TA ? q % PC : q / BM;
      load8(p.A, p.lda, TA, p.vecA, m0 + r, p.M, k0 + c * 8, ra[j]);
    }
#pragma unroll
    for (int j = 0; j < B_PC; ++j) {
      const int q = tid + NT * j;
      const int r = TB ? q % BN : q / PC, c = TB ? q / BN : q % PC;
      load8(p.B, p.ldb, !TB, p.vecB, n0 + r, p.N, k0 + c * 8, rb[j]);
    }
  };

  auto store_tiles = [&](int buf) {
    unsigned char* base = smem + buf * STAGE;
#pragma unroll
    for (int j = 0; j < A_PC; ++j) {
      const int q = tid + NT * j;
      const int r = TA ? q / PC : q % BM, c = TA ? q % PC : q / BM;
      u16x8 t[TERMS];
      split8<TERMS>(ra[j], t);
#pragma unroll
      for (int s = 0; s < TERMS; ++s)
        *reinterpret_cast<u16x8*>(base + s * PLANE + r * ROWB + swz4(r, c) * 16) = t[s];
    }
#pragma unroll
    for (int j = 0; j < B_PC; ++j) {
      const int q = tid + NT * j;
      const int r = TB ? q % BN : q / PC, c = TB ? q / BN : q % PC;
      u16x8 t[TERMS];
      split8<TERMS>(rb[j], t);
      const int rr = BM + r;
#pragma unroll
      for (int s = 0; s < TERMS; ++s)
        *reinterpret_cast<u16x8*>(base + s * PLANE + rr * ROWB + swz4(rr, c) * 16) = t[s];
    }
  };

  floatx16 acc[RM][RN];
#pragma unroll
  for (int i = 0; i < RM; ++i)
#pragma unroll
    for (int j = 0; j < RN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

  const int li = lane & 31;
  const int lg = lane >> 5;

  if (nkt > 0) {
    load_tiles(kbeg);
    store_tiles(0);
    __syncthreads();
    for (int kt = 0; kt < nkt; ++kt) {
      const int buf = kt & 1;
      if (kt + 1 < nkt) load_tiles(kbeg + (kt + 1) * BK);
      const unsigned char* base = smem + buf * STAGE;
#pragma unroll
      for (int s = 0; s < 2; ++s) {  // two 16-deep MFMA steps per slice
        const int c = 2 * s + lg;
        bf16x8 a[TERMS][RM], b[TERMS][RN];
#pragma unroll
        for (int t = 0; t < TERMS; ++t) {
#pragma unroll
          for (int i = 0; i < RM; ++i) {
            const int r = wm * WTM + i * 32 + li;
            a[t][i] = *reinterpret_cast<const bf16x8*>(base + t * PLANE + r * ROWB + swz4(r, c) * 16);
          }
#pragma unroll
          for (int j = 0; j < RN; ++j) {
            const int r = BM + wn * WTN + j * 32 + li;
            b[t][j] = *reinterpret_cast<const bf16x8*>(base + t * PLANE + r * ROWB + swz4(r, c) * 16);
          }
        }
#pragma unroll
        for (int i = 0; i < RM; ++i)
#pragma unroll
          for (int j = 0; j < RN; ++j) {
            if constexpr (TERMS == 3) {
              acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[2][i], b[0][j], acc[i][j], 0, 0, 0);
              acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], b[2][j], acc[i][j], 0, 0, 0);
              acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1][i], b[1][j], acc[i][j], 0, 0, 0);
            }
            if constexpr (TERMS >= 2) {
              acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1][i], b[0][j], acc[i][j], 0, 0, 0);
              acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], b[1][j], acc[i][j], 0, 0, 0);
            }
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], b[0][j], acc[i][j], 0, 0, 0);
          }
      }
      if (kt + 1 < nkt) store_tiles(buf ^ 1);
      __syncthreads();
    }
  }

  const bool partial = gridDim.y > 1;
#pragma unroll
  for (int i = 0; i < RM; ++i)
#pragma unroll
    for (int j = 0; j < RN; ++j) {
      const int col = n0 + wn * WTN + j * 32 + (lane & 31);
      if (col >= p.N) continue;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm * WTM + i * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
        if (row >= p.M) continue;
        if (partial) {
          p.W[(int64_t)z * p.M * p.N + row + (int64_t)col * p.M] = acc[i][j][r];
        } else {
          float* dst = p.C + row + (int64_t)col * p.ldc;
          const float v = acc[i][j][r];
          *dst = (p.beta == 0.0f) ? p.alpha * v : __builtin_fmaf(p.alpha, v, p.beta * *dst);
        }
      }
    }
}

// Device form of the split itself (for parity checks against the host
// restatement): x[n] -> TERMS planes of bf16 bits.
template <int TERMS>
__global__ void __launch_bounds__(256) split_bf16_kernel(const float* x, int64_t n, unsigned short* t0,
                                                        unsigned short* t1, unsigned short* t2) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    float v[8] = {x[i], 0, 0, 0, 0, 0, 0, 0};
    u16x8 t[TERMS];
    split8<TERMS>(v, t);
    t0[i] = t[0][0];
    if constexpr (TERMS > 1) t1[i] = t[1][0];
    if constexpr (TERMS > 2) t2[i] = t[2][0];
  }
}

template <int BM, int BN, int WGM, int WGN, int TERMS>
static hipError_t launch_split_t(const GemmParams<float>& p, bool ta, bool tb, int splits, hipStream_t s) {
  const int ntm = (p.M + BM - 1) / BM, ntn = (p.N + BN - 1) / BN;
  dim3 grid(ntm * ntn, splits), block(64 * WGM * WGN);
  if (!ta && !tb)
    hipLaunchKernelGGL((gemm_split_kernel<BM, BN, WGM, WGN, false, false, TERMS>), grid, block, 0, s, p);
  else if (ta && !tb)
    hipLaunchKernelGGL((gemm_split_kernel<BM, BN, WGM, WGN, true, false, TERMS>), grid, block, 0, s, p);
  else if (!ta && tb)
    hipLaunchKernelGGL((gemm_split_kernel<BM, BN, WGM, WGN, false, true, TERMS>), grid, block, 0, s, p);
  else
    hipLaunchKernelGGL((gemm_split_kernel<BM, BN, WGM, WGN, true, true, TERMS>), grid, block, 0, s, p);
  return hipGetLastError();
}

hipError_t launch_gemm_split(const GemmParams<float>& p, int terms, int tile, bool ta, bool tb, int splits,
                             hipStream_t s) {
  // tile 0: 128x128 (4 waves of 64x64), 1: 256x128 (4 waves of 128x64)
  if (terms == 3) {
    if (tile == 0) return launch_split_t<128, 128, 2, 2, 3>(p, ta, tb, splits, s);
    return launch_split_t<256, 128, 2, 2, 3>(p, ta, tb, splits, s);
  }
  if (terms == 2) {
    if (tile == 0) return launch_split_t<128, 128, 2, 2, 2>(p, ta, tb, splits, s);
    return launch_split_t<256, 128, 2, 2, 2>(p, ta, tb, splits, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_split_bf16(const float* x, int64_t n, int terms, unsigned short* t0, unsigned short* t1,
                             unsigned short* t2, hipStream_t s) {
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  if (terms == 1)
    hipLaunchKernelGGL(split_bf16_kernel<1>, dim3((unsigned)blocks), dim3(256), 0, s, x, n, t0, t1, t2);
  else if (terms == 2)
    hipLaunchKernelGGL(split_bf16_kernel<2>, dim3((unsigned)blocks), dim3(256), 0, s, x, n, t0, t1, t2);
  else
    hipLaunchKernelGGL(split_bf16_kernel<3>, dim3((unsigned)blocks), dim3(256), 0, s, x, n, t0, t1, t2);
  return hipGetLastError();
}

}  // namespace eh
