// common.hpp -- shared device/host definitions for the gfx950 product kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cplx.hpp"

namespace eh {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// MFMA micro-tile traits.  The k index a lane feeds to one instruction is a
// free permutation (k is summed), so kernels choose it to make each lane's
// k values contiguous in LDS; only the output layout is fixed by hardware.
template <typename T>
struct Mfma;

// v_mfma_f32_32x32x2_f32: A lane l -> (row l%32, k l/32); 16 accumulators.
template <>
struct Mfma<float> {
  static constexpr int MI = 32, NI = 32;
  static constexpr int KG = 2;  // k groups across lanes (lane / 32)
  static constexpr int ACC = 16;
  using acc_t = floatx16;
  __device__ static inline void mma(float a, float b, acc_t& c) {
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  }
  // row (within the 32x32 block) of accumulator r held by lane l; column is l%32
  __device__ static inline int row(int r, int l) { return (r >> 2) * 8 + (l >> 5) * 4 + (r & 3); }
  __device__ static inline int col(int l) { return l & 31; }
};

// v_mfma_f64_16x16x4_f64: A lane l -> (row l%16, k l/16); 4 accumulators,
// accumulator r of lane l holds row l/16 + 4r (rows interleaved across lanes,
// unlike the FP32 16x16 layout).
template <>
struct Mfma<double> {
  static constexpr int MI = 16, NI = 16;
  static constexpr int KG = 4;
  static constexpr int ACC = 4;
  using acc_t = doublex4;
  __device__ static inline void mma(double a, double b, acc_t& c) {
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int r, int l) { return (l >> 4) + 4 * r; }
  __device__ static inline int col(int l) { return l & 15; }
};

__device__ inline float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename R>
__device__ inline cplx<R> fma_t(cplx<R> a, cplx<R> b, cplx<R> c) {
  return {fma_t(a.re, b.re, fma_t(-a.im, b.im, c.re)), fma_t(a.re, b.im, fma_t(a.im, b.re, c.im))};
}

// XOR swizzle of 16-byte chunks inside an LDS row.  For 8 chunks per row
// (128-byte rows) any 8 consecutive rows at the same logical chunk land on
// distinct chunks modulo 8, and any 16 consecutive rows on distinct chunks
// modulo 16, so ds_read_b128 / ds_write_b128 by consecutive lanes walking
// rows do not collide on banks whether the LDS serves 128 or 256 bytes per
// clock.
__device__ __host__ inline int swz8(int row, int chunk) { return chunk ^ ((row ^ (row >> 3)) & 7); }
// 64-byte rows (4 chunks) are stored in pairs: rows 2R and 2R + 1 share the
// 128-byte LDS row R, as logical chunks 0-3 and 4-7 under swz8.  32 lanes
// reading 32 consecutive rows at one chunk then touch every bank group of
// each 128-byte phase exactly once (a plain per-row XOR over 4 chunks costs
// the same reads four extra LDS cycles each, measured with
// SQ_LDS_BANK_CONFLICT on gfx950).
__device__ __host__ inline int lds_pair_offset(int row, int chunk) {
  const int R = row >> 1;
  return R * 128 + swz8(R, ((row & 1) << 2) | chunk) * 16;
}

// Raw buffer resource over [base, base + bytes): loads past the end return 0
// in hardware, which is how the product kernels handle matrix edges without
// per-element branches.  Word 3 is the gfx9 default (32-bit data format).
__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ inline u32x4 buf_load16(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0);
}
__device__ inline float buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float*) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ inline double buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, double*) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0);
  return __hiloint2double((int)v.y, (int)v.x);
}

// Map a linear workgroup id to a logical tile id so that each XCD (workgroups
// are dealt round-robin over the 8 XCDs) owns one contiguous range of tiles,
// which keeps the A row panels and B column panels it reuses in its own L2.
__device__ inline int xcd_remap(int bid, int nblocks) {
  constexpr int kXcd = 8;
  if (nblocks % kXcd != 0) return bid;
  return (bid % kXcd) * (nblocks / kXcd) + bid / kXcd;
}

// Grouped tile order: walk GM tile-rows at a time so neighbouring tiles share
// B column panels in L2.
__device__ inline void tile_coords(int logical, int ntm, int ntn, int gm, int& tm, int& tn) {
  int group = gm * ntn;
  int g = logical / group;
  int first = g * gm;
  int rows = ntm - first < gm ? ntm - first : gm;
  int in = logical - g * group;
  tm = first + in % rows;
  tn = in / rows;
}

// Tile (r, c), c <= r, of the lower triangle of an nt x nt tile grid for
// logical index L in [0, nt*(nt+1)/2): the triangle is walked in groups of
// gm tile rows, column by column inside a group (the order tile_coords uses
// for a full grid), so a contiguous range of indices -- what xcd_remap gives
// one XCD -- covers a compact block of tiles that share operand panels, and
// every XCD gets the same number of tiles.  Wave-uniform integer arithmetic.
__host__ __device__ inline void tri_tile_coords(int L, int nt, int gm, int& r, int& c) {
  // tiles before group g (all groups before it are full)
  auto prefix = [gm](int64_t g) { return gm * gm * g * (g - 1) / 2 + g * gm * (gm + 1) / 2; };
  const int groups = (nt + gm - 1) / gm;
  int g = (int)(sqrtf(2.0f * (float)L) / (float)gm);
  if (g > groups - 1) g = groups - 1;
  while (g + 1 < groups && prefix(g + 1) <= L) ++g;
  while (g > 0 && prefix(g) > L) --g;
  int o = L - (int)prefix(g);
  const int base = g * gm;
  const int R = nt - base < gm ? nt - base : gm;  // rows in this group
  if (o < R * base) {  // the rectangle left of the diagonal block
    c = o / R;
    r = base + o % R;
    return;
  }
  o -= R * base;  // the R x R lower triangle on the diagonal, column c' has R - c' tiles
  int cc = 0;
  while (o >= R - cc) {
    o -= R - cc;
    ++cc;
  }
  c = base + cc;
  r = base + cc + o;
}

}  // namespace eh
