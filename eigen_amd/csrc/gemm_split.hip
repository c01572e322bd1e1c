// gemm_split.hip -- FP32-accurate sgemm on BF16 MFMA (v_mfma_f32_32x32x16_bf16).
//
// Every FP32 element x of op(A) and op(B) is split once, by a packing pass,
// into bf16 terms
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
// The packing pass plays the role of Eigen's gemm_pack_lhs/gemm_pack_rhs; the
// product kernel streams the packed planes into LDS with direct
// global-to-LDS loads, runs the cross products as MFMA micro-tiles and fuses
// alpha/beta into the store.
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

// ---------------------------------------------------------------------------
// Pack: op(X) (rows x K, FP32) -> TERMS bf16 planes laid out
//     P[ks][t][r][0..31]   ks = k / 32, r < Rp (rows padded to 256), 64 B rows
// with zeros past `rows` and past K.  This is gemm_pack_lhs / gemm_pack_rhs
// (Eigen/src/Core/products/GeneralBlockPanelKernel.h:1893-2478) for the
// split product: each element is split once here instead of once per tile
// that stages it, and every k slice of a tile becomes one contiguous block.
// KCONT: op(X)(r,k) = X[k + r*ld]; otherwise op(X)(r,k) = X[r + k*ld].
// ---------------------------------------------------------------------------
template <int TERMS, bool KCONT>
__global__ void __launch_bounds__(256) pack_split_kernel(const float* __restrict__ X, int64_t ld, int rows, int K,
                                                         int Rp, int KS, u16x8* __restrict__ P) {
  // A workgroup moves a 64-row x 64-k tile through LDS: reads run along the
  // contiguous index of X, writes along r (four 16-byte chunks per 64-byte
  // row, 64 rows contiguous per plane and k slice).
  constexpr int TR = 64, TK = 64;
  __shared__ float tile[TR][TK + 1];
  // consecutive workgroups walk the contiguous index of X, so the ones running
  // together read neighbouring segments: a 2-D grid (row tiles in x) when rows
  // are contiguous, a 1-D grid with the k tiles fastest when k is
  const int ktiles = (K + TK - 1) / TK;
  const int r0 = (KCONT ? blockIdx.x / ktiles : blockIdx.x) * TR;
  const int k0 = (KCONT ? blockIdx.x % ktiles : blockIdx.y) * TK;
  const int tid = threadIdx.x;
  // 16-byte loads of 4 consecutive elements along the contiguous index (a
  // buffer load needs 4-byte alignment only); reads past the operand return
  // zero, elements past `rows` or K are zeroed.
  const uint64_t extent = KCONT ? ((uint64_t)(rows - 1) * ld + K) * 4 : ((uint64_t)(K - 1) * ld + rows) * 4;
  const int64_t base_el = KCONT ? (int64_t)r0 * ld + k0 : r0 + (int64_t)k0 * ld;
  const uint64_t left = extent > (uint64_t)base_el * 4 ? extent - (uint64_t)base_el * 4 : 0;
  const __amdgpu_buffer_rsrc_t rs =
      make_rsrc(X + base_el, (uint32_t)(left < 0xFFFFFFF0ull ? left : 0xFFFFFFF0ull));
#pragma unroll
  for (int i = 0; i < TR * TK / 1024; ++i) {
    const int q = tid + 256 * i;  // one 4-element group
    const int r = KCONT ? q / (TK / 4) : (q % (TR / 4)) * 4;
    const int k = KCONT ? (q % (TK / 4)) * 4 : q / (TR / 4);
    const uint32_t off = (uint32_t)((KCONT ? (int64_t)r * ld + k : r + (int64_t)k * ld) * 4);
    const u32x4 v = buf_load16(rs, off, 0);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int rr = KCONT ? r : r + e, kk = KCONT ? k + e : k;
      const bool ok = r0 + rr < rows && k0 + kk < K;
      tile[rr][kk] = ok ? __uint_as_float(v[e]) : 0.0f;
    }
  }
  __syncthreads();
  // 2 slices x 64 rows x 4 chunks = 512 chunks: chunk fastest, then row
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int q = tid + 256 * i;
    const int c = q & 3, r = (q >> 2) & (TR - 1), sl = q >> 8;
    const int ks = k0 / 32 + sl;
    if (ks >= KS) continue;
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = tile[r][sl * 32 + c * 8 + e];
    u16x8 t[TERMS];
    split8<TERMS>(v, t);
#pragma unroll
    for (int s2 = 0; s2 < TERMS; ++s2) P[(((int64_t)ks * TERMS + s2) * Rp + r0 + r) * 4 + c] = t[s2];
  }
}

__device__ inline void wait_vmcnt0() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Barrier between LDS stages that leaves later stages' direct loads in
// flight.  __syncthreads' release fence waits for every outstanding vector
// memory operation (vmcnt(0)), including the direct-to-LDS loads of stages
// still ahead, which cuts the pipeline to one stage in flight.  Here each
// wave has already waited (vmcnt) for its own loads of the stage about to be
// read, so after the barrier every wave's part of that stage is in LDS; its
// own LDS reads of the previous stage are complete (lgkmcnt(0)) before it
// arrives, so the buffer refilled next is free.  The empty asm statements'
// memory clobbers keep the compiler from moving LDS reads or loads across
// it; the wait is the builtin so that the compiler's own wait insertion
// sees it.
__device__ inline void stage_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt, expcnt unconstrained
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}
template <int N>
__device__ inline void wait_vmcnt() {
  // gfx9 encoding: vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] | vmcnt[5:4] << 14
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

// ---------------------------------------------------------------------------
// Product over packed planes.  Each stage of LDS holds KD of k (a whole
// 32-deep packed slice, or half of one): TERMS planes of [BM rows of A | BN
// rows of B] x 2*KD bytes.  Stages are filled by global_load_lds_dwordx4:
// every wave instruction moves one 1 KB block straight from HBM into LDS,
// each lane fetching the global chunk that belongs at its LDS position.
// NSTAGE - 1 stages are in flight while one is consumed.
//   KD = 32: 64-byte rows stored in pairs under swz8 (lds_pair_offset).
//   KD = 16: per 16 plane rows and 16-byte chunk h, one 256-byte run of the
//            16 rows' chunks in row order (lds_half_offset), so the 8 lanes
//            a ds_read_b128 serves together read 128 contiguous bytes.
// Half stages let a 256x256 tile with three terms keep three stages in LDS
// (48 KB each): half the L2-to-LDS bytes and 3/4 of the LDS reads per MFMA
// of the 256x128 tile.
// ---------------------------------------------------------------------------
// F16: the planes hold IEEE binary16 operands (one term, no split) and the
// product runs on v_mfma_f32_32x32x16_f16 instead of the BF16 instruction.
__device__ __host__ inline int lds_half_offset(int row, int h) { return ((row >> 4) * 2 + h) * 256 + (row & 15) * 16; }

// Wait until at most `ahead` stages of LPW direct loads each are still in
// flight (capped at A, the most a loop ever has ahead of the stage it reads).
template <int A, int LPW>
__device__ inline void wait_stages_behind(int ahead) {
  if constexpr (A <= 0) {
    wait_vmcnt0();
  } else {
    if (ahead >= A)
      wait_vmcnt<A * LPW>();
    else
      wait_stages_behind<A - 1, LPW>(ahead);
  }
}

// Scheduling groups of one pipelined step: one MFMA, then one LDS read of
// the next step, until the NREAD reads are placed, then the other MFMAs.
template <int NREAD, int NMMA, int Q = 0>
__device__ inline void interleave_order() {
  if constexpr (Q < NREAD) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
    __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // LDS read
    interleave_order<NREAD, NMMA, Q + 1>();
  } else {
    __builtin_amdgcn_sched_group_barrier(0x008, NMMA - Q, 0);
  }
}

template <int BM, int BN, int WGM, int WGN, int TERMS, int NSTAGE, bool F16 = false, int KD = 32, bool PIPE = false>
__global__ void __launch_bounds__(64 * WGM * WGN) gemm_split_packed_kernel(SplitParams p) {
  static_assert(KD == 32 || KD == 16, "stage depth");
  constexpr int NW = WGM * WGN;
  constexpr int ROWB = 64;                   // bytes per row of a packed 32-deep slice
  constexpr int RB = 2 * KD;                 // LDS bytes per plane row in a stage
  constexpr int SPS = 32 / KD;               // stages per packed slice
  constexpr int RPB = 1024 / RB;             // plane rows per 1 KB block
  constexpr int PROWS = BM + BN;             // rows of one plane in a stage
  constexpr int PLANE = PROWS * RB;
  constexpr int STAGE = TERMS * PLANE;
  constexpr int BLOCKS = STAGE / 1024;       // 1 KB blocks per stage
  constexpr int LPW = BLOCKS / NW;           // direct loads per wave per stage
  static_assert(PROWS % RPB == 0, "blocks must not straddle planes");
  constexpr int WTM = BM / WGM, WTN = BN / WGN;
  constexpr int RM = WTM / 32, RN = WTN / 32;
  static_assert(BLOCKS % NW == 0, "blocks must divide over waves");
  static_assert(RM >= 1 && RN >= 1, "wave tile smaller than one MFMA block");

  __shared__ __attribute__((aligned(1024))) unsigned char smem[NSTAGE * STAGE];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wid % WGM, wn = wid / WGM;

  const int ntm = (p.M + BM - 1) / BM, ntn = (p.N + BN - 1) / BN;
  int tm, tn;
  tile_coords(xcd_remap(blockIdx.x, gridDim.x), ntm, ntn, 8, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int z = blockIdx.y;
  const int ks0 = z * p.ks_per_split;
  const int nkt = SPS * (min(p.KS, ks0 + p.ks_per_split) - ks0);  // stages

  // Direct loads: block b of a stage covers RPB plane rows [blk0, blk0 + RPB)
  // of one plane, and lane l fills LDS slot l of the block.
  //   KD = 32: 8 LDS rows of 128 bytes (plane rows 2R and 2R + 1 side by
  //   side, see lds_pair_offset): LDS row blk0/2 + l/8, physical chunk l&7,
  //   which holds logical chunk V = (l&7) ^ swizzle(row), i.e. plane row
  //   2R + V/4, 16-byte chunk V&3.
  //   KD = 16: four 256-byte runs (lds_half_offset): plane row
  //   blk0 + 16*(l/32) + (l&15), chunk (l/16)&1 of the stage's half slice.
  // The per-lane global offset is computed per block once; the block bases
  // are wave-uniform and advance by one slice per SPS stages.  An operand
  // read in place (one term, ldAd / ldBd set) has the same rows and chunks
  // at its own addresses: row r of op(X) at r * ld elements, a slice 64
  // bytes further along it; rows past the matrix re-read its last row.
  const unsigned char* srcA = reinterpret_cast<const unsigned char*>(p.Ap);
  const unsigned char* srcB = reinterpret_cast<const unsigned char*>(p.Bp);
  const int64_t sliceA = (int64_t)TERMS * p.Mp * ROWB, sliceB = (int64_t)TERMS * p.Np * ROWB;
  const unsigned char* gbase[LPW];
  uint32_t loff[LPW];
  int64_t gstride[LPW];
#pragma unroll
  for (int i = 0; i < LPW; ++i) {
    const int b = wid + NW * i;                // block index within the stage (wave-uniform)
    const int t = b / (PROWS / RPB);             // plane
    const int blk0 = (b % (PROWS / RPB)) * RPB;  // first plane row of the block
    int prow, pchunk;                            // plane row within the block and 16-byte chunk of this lane
    if constexpr (KD == 32) {
      const int R = (blk0 >> 1) + (lane >> 3);   // LDS row (pair of plane rows)
      const int V = swz8(R, lane & 7);           // logical chunk stored at this LDS slot
      prow = 2 * (lane >> 3) + (V >> 2);
      pchunk = V & 3;
    } else {
      prow = 16 * (lane >> 5) + (lane & 15);
      pchunk = (lane >> 4) & 1;
    }
    const bool isA = blk0 < BM;
    const int64_t ldd = TERMS == 1 ? (isA ? p.ldAd : p.ldBd) : 0;
    if (ldd != 0) {
      const int row = min((isA ? m0 + blk0 : n0 + blk0 - BM) + prow, (isA ? p.M : p.N) - 1);
      gbase[i] = (isA ? srcA : srcB) + (int64_t)ks0 * ROWB;
      loff[i] = (uint32_t)((int64_t)row * ldd * 2 + pchunk * 16);
      gstride[i] = ROWB;
    } else {
      loff[i] = (uint32_t)(prow * ROWB + pchunk * 16);
      if (isA) {
        gbase[i] = srcA + (int64_t)ks0 * sliceA + ((int64_t)t * p.Mp + m0 + blk0) * ROWB;
        gstride[i] = sliceA;
      } else {
        gbase[i] = srcB + (int64_t)ks0 * sliceB + ((int64_t)t * p.Np + n0 + blk0 - BM) * ROWB;
        gstride[i] = sliceB;
      }
    }
  }
  auto issue = [&](int kt, int buf) {
#pragma unroll
    for (int i = 0; i < LPW; ++i) {
      const int b = wid + NW * i;
      __attribute__((address_space(3))) void* l =
          (__attribute__((address_space(3))) void*)(smem + buf * STAGE + b * 1024);
      const int64_t g = (SPS == 1 ? kt : kt / SPS) * gstride[i] + (SPS == 1 ? 0 : (kt % SPS) * 32);
      __builtin_amdgcn_global_load_lds((const void*)(gbase[i] + g + loff[i]), l, 16, 0, 0);
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

  // Fragments travel as 32-bit integer vectors (bf16 vectors crossing
  // control flow are legalised element by element) and are bit-cast at the
  // MFMA.
  constexpr int NSTEP = KD / 16;  // 16-deep MFMA steps per stage
  auto read_step = [&](const unsigned char* base, int s, u32x4(&fa)[TERMS][RM], u32x4(&fb)[TERMS][RN]) {
#pragma unroll
    for (int t = 0; t < TERMS; ++t) {
#pragma unroll
      for (int i = 0; i < RM; ++i) {
        const int r = wm * WTM + i * 32 + li;
        const int off = KD == 32 ? lds_pair_offset(r, 2 * s + lg) : lds_half_offset(r, lg);
        fa[t][i] = *reinterpret_cast<const u32x4*>(base + t * PLANE + off);
      }
#pragma unroll
      for (int j = 0; j < RN; ++j) {
        const int r = BM + wn * WTN + j * 32 + li;
        const int off = KD == 32 ? lds_pair_offset(r, 2 * s + lg) : lds_half_offset(r, lg);
        fb[t][j] = *reinterpret_cast<const u32x4*>(base + t * PLANE + off);
      }
    }
  };
  // cross products with p + q < TERMS, smallest first; consecutive MFMAs go
  // to different accumulators (one term: the single product of bf16 inputs)
  auto mma_step = [&](const u32x4(&fa)[TERMS][RM], const u32x4(&fb)[TERMS][RN]) {
    constexpr int NPROD = TERMS == 3 ? 6 : (TERMS == 2 ? 3 : 1);
    constexpr int PA[6] = {TERMS == 3 ? 2 : (TERMS == 2 ? 1 : 0), 0, TERMS == 3 ? 1 : 0, 1, 0, 0};
    constexpr int PB[6] = {0, TERMS == 3 ? 2 : (TERMS == 2 ? 1 : 0), TERMS == 3 ? 1 : 0, 0, 1, 0};
#pragma unroll
    for (int pq = 0; pq < NPROD; ++pq)
#pragma unroll
      for (int i = 0; i < RM; ++i)
#pragma unroll
        for (int j = 0; j < RN; ++j) {
          if constexpr (F16)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, fa[PA[pq]][i]),
                                                               __builtin_bit_cast(f16x8, fb[PB[pq]][j]), acc[i][j], 0,
                                                               0, 0);
          else
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, fa[PA[pq]][i]),
                                                                __builtin_bit_cast(bf16x8, fb[PB[pq]][j]), acc[i][j],
                                                                0, 0, 0);
        }
  };

  if constexpr (PIPE) {
    // Software-pipelined over the two 16-deep steps of each stage: the
    // fragments of the next step are read from LDS into a second register
    // set while the MFMAs of the current one run, one read after each MFMA
    // (scheduling groups).  Entering stage st + 1, a wave waits for its own
    // loads of it and meets the others at the barrier; by then every wave's
    // reads of stage st have returned, so its buffer is refilled NSTAGE
    // stages ahead.  Fills past the last stage reload the last stage (the
    // same bytes) into a free buffer, so every transition waits for the same
    // count of loads and the loop has no branch: a branch inside it makes the
    // compiler's LDS waits before the MFMAs cover the reads just issued.
    static_assert(NSTEP == 2, "the pipelined loop takes two steps per stage");
    auto fill = [&](int st, int buf) { issue(st < nkt ? st : nkt - 1, buf); };
    if (nkt > 0) {
#pragma unroll
      for (int s = 0; s < NSTAGE; ++s) fill(s, s);
      wait_vmcnt<(NSTAGE - 1) * LPW>();
      stage_barrier();
      u32x4 fa0[TERMS][RM], fb0[TERMS][RN], fa1[TERMS][RM], fb1[TERMS][RN];
      read_step(smem, 0, fa0, fb0);
      constexpr int NREAD = TERMS * (RM + RN), NMMA = (TERMS == 3 ? 6 : (TERMS == 2 ? 3 : 1)) * RM * RN;
      static_assert(NMMA >= NREAD, "not enough MFMAs to interleave the reads with");
      for (int st = 0; st < nkt; ++st) {
        read_step(smem + (st % NSTAGE) * STAGE, 1, fa1, fb1);
        mma_step(fa0, fb0);
        interleave_order<NREAD, NMMA>();
        __builtin_amdgcn_sched_barrier(0);
        // stage st + 1 must have landed; NSTAGE - 2 fills after it may not
        wait_vmcnt<(NSTAGE - 2) * LPW>();
        stage_barrier();
        fill(st + NSTAGE, st % NSTAGE);
        // no LDS read is outstanding here (the barrier waited for them): this
        // wait costs nothing and keeps the compiler from counting the direct
        // loads just issued against the reads below
        __builtin_amdgcn_s_waitcnt(0xC07F);
        const int stn = st + 1 < nkt ? st + 1 : st;  // past the end: an unused re-read
        read_step(smem + (stn % NSTAGE) * STAGE, 0, fa0, fb0);
        mma_step(fa1, fb1);
        interleave_order<NREAD, NMMA>();
        __builtin_amdgcn_sched_barrier(0);
      }
      wait_vmcnt0();  // no direct load may still target LDS when the waves end
    }
  } else {
    // Per stage: wait for it, barrier, refill the buffer freed a stage ago,
    // read all its fragments, multiply.  (Forcing all reads of both 16-deep
    // steps ahead of the MFMAs measured slower than the compiler's
    // interleaving.)
#pragma unroll
    for (int s = 0; s < NSTAGE - 1; ++s)
      if (s < nkt) issue(s, s);
    for (int kt = 0; kt < nkt; ++kt) {
      // stage kt must have landed; the stages issued after it (at most
      // NSTAGE - 2, fewer near the end) may still be in flight
      wait_stages_behind<NSTAGE - 2, LPW>(nkt - 1 - kt);
      stage_barrier();
      if (kt + NSTAGE - 1 < nkt) issue(kt + NSTAGE - 1, (kt + NSTAGE - 1) % NSTAGE);
      const unsigned char* base = smem + (kt % NSTAGE) * STAGE;
      u32x4 a[NSTEP][TERMS][RM], b[NSTEP][TERMS][RN];
#pragma unroll
      for (int s = 0; s < NSTEP; ++s) read_step(base, s, a[s], b[s]);
#pragma unroll
      for (int s = 0; s < NSTEP; ++s) mma_step(a[s], b[s]);
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

// Pack of a bf16 operand (one plane, no split): op(X) (rows x K, bf16 bits)
// -> P[ks][r][0..31] with zeros past `rows` and past K.  Same layout as the
// one-term split planes, so the product kernel is shared.  Loads are 2-byte
// buffer loads by consecutive lanes along the contiguous index (any element
// alignment is valid for bf16 operands).
// WIDE: 16-byte loads of 8 elements along the contiguous index, valid when
// every chunk starts on a 4-byte boundary (even element offset: an even
// leading dimension and a 4-byte aligned base) and the operand ends on one
// (the range check is per dword); otherwise 2-byte loads.
template <bool KCONT, bool WIDE>
__global__ void __launch_bounds__(256) pack_bf16_kernel(const unsigned short* __restrict__ X, int64_t ld, int rows,
                                                        int K, int Rp, int KS, u16x8* __restrict__ P) {
  constexpr int TR = 64, TK = 64;
  __shared__ unsigned short tile[TR][TK + 2];
  const int ktiles = (K + TK - 1) / TK;  // grid shapes as in pack_split_kernel
  const int r0 = (KCONT ? blockIdx.x / ktiles : blockIdx.x) * TR;
  const int k0 = (KCONT ? blockIdx.x % ktiles : blockIdx.y) * TK;
  const int tid = threadIdx.x;
  const uint64_t extent = KCONT ? ((uint64_t)(rows - 1) * ld + K) * 2 : ((uint64_t)(K - 1) * ld + rows) * 2;
  const int64_t base_el = KCONT ? (int64_t)r0 * ld + k0 : r0 + (int64_t)k0 * ld;
  const uint64_t left = extent > (uint64_t)base_el * 2 ? extent - (uint64_t)base_el * 2 : 0;
  const __amdgpu_buffer_rsrc_t rs =
      make_rsrc(X + base_el, (uint32_t)(left < 0xFFFFFFF0ull ? left : 0xFFFFFFF0ull));
  if constexpr (WIDE) {
#pragma unroll
    for (int i = 0; i < TR * TK / (256 * 8); ++i) {
      const int q = tid + 256 * i;  // one 8-element chunk
      const int r = KCONT ? q / (TK / 8) : (q % (TR / 8)) * 8;
      const int k = KCONT ? (q % (TK / 8)) * 8 : q / (TR / 8);
      const uint32_t off = (uint32_t)((KCONT ? (int64_t)r * ld + k : r + (int64_t)k * ld) * 2);
      const u32x4 v = buf_load16(rs, off, 0);
      const unsigned short* e8 = reinterpret_cast<const unsigned short*>(&v);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int rr = KCONT ? r : r + e, kk = KCONT ? k + e : k;
        tile[rr][kk] = (r0 + rr < rows && k0 + kk < K) ? e8[e] : (unsigned short)0;
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < TR * TK / 256; ++i) {
      const int q = tid + 256 * i;
      const int r = KCONT ? q / TK : q % TR;  // lanes walk the contiguous index
      const int k = KCONT ? q % TK : q / TR;
      const uint32_t off = (uint32_t)((KCONT ? (int64_t)r * ld + k : r + (int64_t)k * ld) * 2);
      const unsigned short v = __builtin_amdgcn_raw_buffer_load_b16(rs, (int)off, 0, 0);
      tile[r][k] = (r0 + r < rows && k0 + k < K) ? v : (unsigned short)0;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int q = tid + 256 * i;
    const int c = q & 3, r = (q >> 2) & (TR - 1), sl = q >> 8;
    const int ks = k0 / 32 + sl;
    if (ks >= KS) continue;
    u16x8 t;
#pragma unroll
    for (int e = 0; e < 8; ++e) t[e] = tile[r][sl * 32 + c * 8 + e];
    P[((int64_t)ks * Rp + r0 + r) * 4 + c] = t;
  }
}

hipError_t launch_pack_bf16(const unsigned short* X, int64_t ld, bool kcont, int rows, int K, void* P,
                            hipStream_t s) {
  const int Rp = split_rows_padded(rows), KS = split_kslices(K);
  dim3 grid(Rp / 64, (K + 63) / 64);
  // the buffer range check works per dword: the operand must also end on a
  // dword boundary (an even count of contiguous elements per row or column)
  // so that the dword holding its last element is inside the range
  const bool wide = (reinterpret_cast<uintptr_t>(X) % 4) == 0 && ld % 2 == 0 && (kcont ? K % 2 == 0 : rows % 2 == 0);
  auto* out = static_cast<u16x8*>(P);
  const dim3 kgrid(grid.x * grid.y);  // k-contiguous operands: 1-D, k tiles fastest
  if (kcont && wide)
    hipLaunchKernelGGL((pack_bf16_kernel<true, true>), kgrid, dim3(256), 0, s, X, ld, rows, K, Rp, KS, out);
  else if (kcont)
    hipLaunchKernelGGL((pack_bf16_kernel<true, false>), kgrid, dim3(256), 0, s, X, ld, rows, K, Rp, KS, out);
  else if (wide)
    hipLaunchKernelGGL((pack_bf16_kernel<false, true>), grid, dim3(256), 0, s, X, ld, rows, K, Rp, KS, out);
  else
    hipLaunchKernelGGL((pack_bf16_kernel<false, false>), grid, dim3(256), 0, s, X, ld, rows, K, Rp, KS, out);
  return hipGetLastError();
}

template <int TERMS>
static hipError_t launch_pack_t(const float* X, int64_t ld, bool kcont, int rows, int K, int Rp, int KS, void* P,
                                hipStream_t s) {
  dim3 grid(Rp / 64, (K + 63) / 64);
  if (kcont)
    hipLaunchKernelGGL((pack_split_kernel<TERMS, true>), dim3(grid.x * grid.y), dim3(256), 0, s, X, ld, rows, K, Rp, KS,
                       static_cast<u16x8*>(P));
  else
    hipLaunchKernelGGL((pack_split_kernel<TERMS, false>), grid, dim3(256), 0, s, X, ld, rows, K, Rp, KS,
                       static_cast<u16x8*>(P));
  return hipGetLastError();
}

hipError_t launch_pack_split(const float* X, int64_t ld, bool kcont, int rows, int K, int terms, void* P,
                             hipStream_t s) {
  const int Rp = split_rows_padded(rows), KS = split_kslices(K);
  if (terms == 3) return launch_pack_t<3>(X, ld, kcont, rows, K, Rp, KS, P, s);
  if (terms == 2) return launch_pack_t<2>(X, ld, kcont, rows, K, Rp, KS, P, s);
  return hipErrorInvalidValue;
}

template <int BM, int BN, int WGM, int WGN, int TERMS, int NSTAGE, bool F16 = false, int KD = 32, bool PIPE = false>
static hipError_t launch_packed_t(const SplitParams& p, int splits, hipStream_t s) {
  const int ntm = (p.M + BM - 1) / BM, ntn = (p.N + BN - 1) / BN;
  hipLaunchKernelGGL((gemm_split_packed_kernel<BM, BN, WGM, WGN, TERMS, NSTAGE, F16, KD, PIPE>),
                     dim3(ntm * ntn, splits),
                     dim3(64 * WGM * WGN), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_gemm_split_packed(const SplitParams& p, int terms, int tile, int splits, hipStream_t s, bool f16) {
  if (f16) {  // IEEE half operands: the one-term tiles on the FP16 instruction
    if (terms != 1) return hipErrorInvalidValue;
    if (tile == 0) return launch_packed_t<128, 128, 2, 2, 1, 4, true, 32, true>(p, splits, s);
    return launch_packed_t<256, 256, 2, 4, 1, 3, true, 32, true>(p, splits, s);
  }
  // tile 0: 128x128, 4 waves (3 stages); 1: 256x128, 4 waves of 128x64 (2 stages);
  // 2: 256x128, 8 waves of 64x64 -- two waves per SIMD (2 stages, 3 for two terms);
  // 3: 256x256, 8 waves of 128x64, 16-deep half-slice stages (3 stages, 4 for two terms).
  // Tiles 0 and 2 run the software-pipelined loop: against the plain loop,
  // 6-8% faster at 2048^3 (tile 0: 122-124 against 133-134 us for three
  // terms), 4-6% at 3072^3 and 1-2% at 4096x2048x4096 (tile 2: 446 against
  // 468 us, 366 against 373 us), alternating on one box.
  if (terms == 3) {
    if (tile == 0) return launch_packed_t<128, 128, 2, 2, 3, 3, false, 32, true>(p, splits, s);
    if (tile == 2) return launch_packed_t<256, 128, 4, 2, 3, 2, false, 32, true>(p, splits, s);
    if (tile == 3) return launch_packed_t<256, 256, 2, 4, 3, 3, false, 16>(p, splits, s);
    return launch_packed_t<256, 128, 2, 2, 3, 2>(p, splits, s);
  }
  if (terms == 2) {
    if (tile == 0) return launch_packed_t<128, 128, 2, 2, 2, 3, false, 32, true>(p, splits, s);
    if (tile == 2) return launch_packed_t<256, 128, 4, 2, 2, 3, false, 32, true>(p, splits, s);
    if (tile == 3) return launch_packed_t<256, 256, 2, 4, 2, 4, false, 16>(p, splits, s);
    return launch_packed_t<256, 128, 2, 2, 2, 2>(p, splits, s);
  }
  if (terms == 1) {
    // bf16 inputs: one product per slice, so the LDS reads per MFMA decide
    // the rate.  tile 0: 128x128, 4 waves of 64x64 (4 stages); tile 1:
    // 256x256, 8 waves of 128x64 (two waves per SIMD): 3/4 of the LDS reads
    // per MFMA of 64x64 wave tiles, 903 TFLOP/s at 8192^3 against 637 for
    // the 256x128 8-wave tile and 866 for 4 waves of 128x128 (one wave per
    // SIMD, accumulators in AGPRs); 814 vs 776 at 4096^3.  Its loop is the
    // software-pipelined one with three stages: 4% faster than the plain
    // loop (1.075-1.079 ms against 1.114-1.128 ms at 8192^3, alternating
    // on one box).  Below 256 tiles of 256x256 the 128x128 tile wins; its
    // pipelined loop is 6-9% faster at 2048^3 and 3072^3 (47-49 against 52
    // us, 114 against 121 us), 2% at 1024^3.
    if (tile == 0) return launch_packed_t<128, 128, 2, 2, 1, 4, false, 32, true>(p, splits, s);
    return launch_packed_t<256, 256, 2, 4, 1, 3, false, 32, true>(p, splits, s);
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
