/*
 * oracle.c -- CPU restatement of Eigen's dense matrix-product path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the eigen_amd product links, loads or
 * calls this file; only tests/, the smoke check and the CPU-baseline leg of
 * the benchmark use it, and only as the checker.
 *
 * What it restates (reference paths are relative to the Eigen source tree):
 *   - orc_gemm_check   : argument validation order and xerbla info codes of
 *                        sgemm_/dgemm_           blas/level3_impl.h:61-78
 *   - orc_gemm_gebp_*  : the BLAS gemm semantics (quick returns, beta
 *                        pre-scale, alpha==0 early-out) and the Goto/GEBP
 *                        loop nest with packed panels and the C += alpha*acc
 *                        epilogue          blas/level3_impl.h:80-95,
 *                        Eigen/src/Core/products/GeneralMatrixMatrix.h:126-245,
 *                        Eigen/src/Core/products/GeneralBlockPanelKernel.h:1352-1456
 *   - orc_gemm_ref     : long-double brute-force product (the FP64 / extended
 *                        oracle the error metric is measured against)
 *   - orc_bf16_rtne    : float -> bfloat16 round-to-nearest-even with NaN
 *                        squashing      Eigen/src/Core/arch/Default/BFloat16.h:434-600
 *   - orc_bf16_split   : hi/mid/lo bf16 split of a float (each term the RTNE
 *                        of the remaining residual)
 *   - orc_gemm_split_ref : long-double sum of the kept bf16 cross products
 *   - orc_fill_*       : the seeded splitmix64 input generators
 *
 * Pinning (see tests/test_oracle.py): golden bf16 bit patterns from
 * test/bfloat16_float.cpp:59-96; Eigen itself compiled from the reference
 * tree into oracle/_ref (GEBP results, bf16 rounding, BLAS info codes);
 * closed forms (identity, all-ones, small-integer exact products) and
 * brute force on tiny inputs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* splitmix64 generator: element i of a stream is mix(seed + (i+1)*gamma),   */
/* so every element is independent of the others (the device generator in   */
/* the product computes the same sequence in parallel).                      */
/* ------------------------------------------------------------------------ */
static const uint64_t kGamma = 0x9E3779B97F4A7C15ull;

ORC_API uint64_t orc_splitmix64_at(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * kGamma;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double u01(uint64_t u) { return (double)(u >> 11) * 0x1.0p-53; }

enum { ORC_UNIFORM = 0, ORC_NORMAL = 1, ORC_WIDE = 2, ORC_INT = 3 };

/* One value of distribution `dist` for element i, computed in double. */
static double orc_value(uint64_t seed, uint64_t i, int dist, double p0, double p1) {
  uint64_t ua = orc_splitmix64_at(seed, 2 * i);
  switch (dist) {
    case ORC_UNIFORM:
      return fma(p1 - p0, u01(ua), p0);
    case ORC_NORMAL: {
      uint64_t ub = orc_splitmix64_at(seed, 2 * i + 1);
      double r1 = ((double)(ua >> 11) + 1.0) * 0x1.0p-53; /* (0,1] */
      double r2 = u01(ub);
      return p0 + p1 * sqrt(-2.0 * log(r1)) * cos(6.283185307179586 * r2);
    }
    case ORC_WIDE: {
      /* sign * U(1,2) * 2^U(p0,p1) */
      uint64_t ub = orc_splitmix64_at(seed, 2 * i + 1);
      double sgn = (ua & 1) ? -1.0 : 1.0;
      double mant = 1.0 + u01(ua);
      double e = fma(p1 - p0, u01(ub), p0);
      return sgn * mant * exp2(e);
    }
    case ORC_INT:
    default:
      return p0 + floor(u01(ua) * (p1 - p0 + 1.0));
  }
}

ORC_API void orc_fill_f64(double* x, int64_t n, uint64_t seed, int dist, double p0, double p1) {
  for (int64_t i = 0; i < n; ++i) x[i] = orc_value(seed, (uint64_t)i, dist, p0, p1);
}

ORC_API void orc_fill_f32(float* x, int64_t n, uint64_t seed, int dist, double p0, double p1) {
  for (int64_t i = 0; i < n; ++i) x[i] = (float)orc_value(seed, (uint64_t)i, dist, p0, p1);
}

/* ------------------------------------------------------------------------ */
/* BLAS ?gemm_ argument validation, blas/level3_impl.h:61-78.               */
/* ------------------------------------------------------------------------ */
enum { OP_N = 0, OP_T = 1, OP_C = 2, OP_BAD = 0xff };

static int op_code(char x) {
  if (x == 'N' || x == 'n') return OP_N;
  if (x == 'T' || x == 't') return OP_T;
  if (x == 'C' || x == 'c') return OP_C;
  return OP_BAD;
}

static int imax(int a, int b) { return a > b ? a : b; }

ORC_API int orc_gemm_check(char ta, char tb, int m, int n, int k, int lda, int ldb, int ldc) {
  int oa = op_code(ta), ob = op_code(tb);
  if (oa == OP_BAD) return 1;
  if (ob == OP_BAD) return 2;
  if (m < 0) return 3;
  if (n < 0) return 4;
  if (k < 0) return 5;
  if (lda < imax(1, oa == OP_N ? m : k)) return 8;
  if (ldb < imax(1, ob == OP_N ? k : n)) return 10;
  if (ldc < imax(1, m)) return 13;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* GEBP restatement.  The loop nest follows GeneralMatrixMatrix.h:29-62      */
/* (pack A' per (i2,k2), pack B' per j2, gebp on the packed panels) and the  */
/* micro-kernel follows GeneralBlockPanelKernel.h:1352-1456: an mr x nr      */
/* register block accumulates sum_k a*b over one kc slice with fused         */
/* multiply-adds, then the store does C += alpha * acc.                      */
/* Transposed operands are read through their strides exactly as             */
/* const_blas_data_mapper does (BlasUtil.h:175-443): no copy is made.        */
/* ------------------------------------------------------------------------ */
#define ORC_MR 8
#define ORC_NR 4

#define DEFINE_GEBP(T, SUF, FMA)                                                                          \
  static T opA_##SUF(const T* A, int64_t lda, int ta, int64_t i, int64_t k) {                             \
    return ta ? A[k + i * lda] : A[i + k * lda];                                                           \
  }                                                                                                        \
  static T opB_##SUF(const T* B, int64_t ldb, int tb, int64_t k, int64_t j) {                             \
    return tb ? B[j + k * ldb] : B[k + j * ldb];                                                           \
  }                                                                                                        \
  ORC_API int orc_gemm_gebp_##SUF(char cta, char ctb, int m, int n, int k, T alpha, const T* A, int lda,  \
                                  const T* B, int ldb, T beta, T* C, int ldc, int kc, int mc, int nc) {     \
    int info = orc_gemm_check(cta, ctb, m, n, k, lda, ldb, ldc);                                          \
    if (info) return info;                                                                                 \
    if (m == 0 || n == 0) return 0;                                                                        \
    int ta = op_code(cta) != OP_N, tb = op_code(ctb) != OP_N;                                              \
    if (beta != (T)1) {                                                                                    \
      for (int64_t j = 0; j < n; ++j)                                                                      \
        for (int64_t i = 0; i < m; ++i)                                                                    \
          C[i + j * (int64_t)ldc] = (beta == (T)0) ? (T)0 : C[i + j * (int64_t)ldc] * beta;               \
    }                                                                                                      \
    if (k == 0) return 0;                                                                                  \
    if (alpha == (T)0) return 0; /* GeneralMatrixMatrix.h:130: A and B are not read */                     \
    if (kc <= 0) kc = k;                                                                                   \
    if (mc <= 0) mc = m;                                                                                   \
    if (nc <= 0) nc = n;                                                                                   \
    T* packA = (T*)malloc(sizeof(T) * (size_t)(mc + ORC_MR) * (size_t)kc);                                 \
    T* packB = (T*)malloc(sizeof(T) * (size_t)(nc + ORC_NR) * (size_t)kc);                                 \
    for (int64_t i2 = 0; i2 < m; i2 += mc) {                                                               \
      int64_t amc = (i2 + mc < m ? i2 + mc : m) - i2;                                                      \
      for (int64_t k2 = 0; k2 < k; k2 += kc) {                                                             \
        int64_t akc = (k2 + kc < k ? k2 + kc : k) - k2;                                                    \
        /* pack A': mr-row micro panels, k-major inside a panel (GeneralBlockPanelKernel.h:1879-1892) */   \
        for (int64_t p = 0; p < amc; p += ORC_MR)                                                          \
          for (int64_t kk = 0; kk < akc; ++kk)                                                             \
            for (int64_t r = 0; r < ORC_MR; ++r)                                                           \
              packA[p * akc + kk * ORC_MR + r] =                                                           \
                  (p + r < amc) ? opA_##SUF(A, lda, ta, i2 + p + r, k2 + kk) : (T)0;                       \
        for (int64_t j2 = 0; j2 < n; j2 += nc) {                                                           \
          int64_t anc = (j2 + nc < n ? j2 + nc : n) - j2;                                                  \
          /* pack B': nr-column panels, row-interleaved (GeneralBlockPanelKernel.h:2178-2184) */           \
          for (int64_t q = 0; q < anc; q += ORC_NR)                                                        \
            for (int64_t kk = 0; kk < akc; ++kk)                                                           \
              for (int64_t c = 0; c < ORC_NR; ++c)                                                         \
                packB[q * akc + kk * ORC_NR + c] =                                                         \
                    (q + c < anc) ? opB_##SUF(B, ldb, tb, k2 + kk, j2 + q + c) : (T)0;                     \
          /* gebp: register block mr x nr, accumulate over the kc slice, then C += alpha*acc */           \
          for (int64_t q = 0; q < anc; q += ORC_NR)                                                        \
            for (int64_t p = 0; p < amc; p += ORC_MR) {                                                    \
              T acc[ORC_MR][ORC_NR];                                                                       \
              memset(acc, 0, sizeof(acc));                                                                 \
              const T* pa = packA + p * akc;                                                               \
              const T* pb = packB + q * akc;                                                               \
              for (int64_t kk = 0; kk < akc; ++kk)                                                         \
                for (int r = 0; r < ORC_MR; ++r)                                                           \
                  for (int c = 0; c < ORC_NR; ++c)                                                         \
                    acc[r][c] = FMA(pa[kk * ORC_MR + r], pb[kk * ORC_NR + c], acc[r][c]);                  \
              for (int c = 0; c < ORC_NR && q + c < anc; ++c)                                              \
                for (int r = 0; r < ORC_MR && p + r < amc; ++r) {                                          \
                  T* dst = &C[(i2 + p + r) + (j2 + q + c) * (int64_t)ldc];                                 \
                  *dst = FMA(alpha, acc[r][c], *dst);                                                      \
                }                                                                                          \
            }                                                                                              \
        }                                                                                                  \
      }                                                                                                    \
    }                                                                                                      \
    free(packA);                                                                                           \
    free(packB);                                                                                           \
    return 0;                                                                                              \
  }

DEFINE_GEBP(float, f32, fmaf)
DEFINE_GEBP(double, f64, fma)

/* ------------------------------------------------------------------------ */
/* Long-double brute-force reference on a subset of rows.                    */
/*   out[r*n + j] = alpha * sum_k op(A)(rows[r],k) op(B)(k,j) + beta*C0      */
/*   absout[r*n+j] = |alpha| * sum_k |op(A)| |op(B)| (for error bounds)      */
/* beta == 0 means C0 is not read (blas/level3_impl.h:82-87).                */
/* dtype 0 = float, 1 = double.  rows == NULL means rows 0..nrows-1.         */
/* ------------------------------------------------------------------------ */
static long double ld_at(const void* p, int dtype, int64_t idx) {
  return dtype == 0 ? (long double)((const float*)p)[idx] : (long double)((const double*)p)[idx];
}

ORC_API void orc_gemm_ref(int dtype, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha,
                          const void* A, int64_t lda, const void* B, int64_t ldb, double beta, const void* C0,
                          int64_t ldc, const int64_t* rows, int64_t nrows, double* out, double* absout) {
  (void)m;
  long double* colB = (long double*)malloc(sizeof(long double) * (size_t)(k > 0 ? k : 1));
  for (int64_t j = 0; j < n; ++j) {
    for (int64_t kk = 0; kk < k; ++kk) colB[kk] = ld_at(B, dtype, tb ? j + kk * ldb : kk + j * ldb);
    for (int64_t r = 0; r < nrows; ++r) {
      int64_t i = rows ? rows[r] : r;
      long double s = 0.0L, sa = 0.0L;
      for (int64_t kk = 0; kk < k; ++kk) {
        long double a = ld_at(A, dtype, ta ? kk + i * lda : i + kk * lda);
        s += a * colB[kk];
        sa += fabsl(a * colB[kk]);
      }
      long double v = (long double)alpha * s;
      if (beta != 0.0) v += (long double)beta * ld_at(C0, dtype, i + j * ldc);
      out[r * n + j] = (double)v;
      if (absout) absout[r * n + j] = (double)(fabsl((long double)alpha) * sa);
    }
  }
  free(colB);
}

/* ------------------------------------------------------------------------ */
/* bfloat16: BFloat16.h:434-600 (NaN squashed to a signed quiet NaN,         */
/* otherwise add 0x7fff + lsb and truncate) and BFloat16.h:624 (widening).   */
/* ------------------------------------------------------------------------ */
ORC_API uint16_t orc_bf16_rtne(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if (isnan(f)) return (u >> 31) ? 0xFFC0 : 0x7FC0;
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}

ORC_API float orc_bf16_to_f32(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

/* Split x into `terms` bf16 values: t0 = rtne(x), t1 = rtne(x - t0),
 * t2 = rtne(x - t0 - t1).  The residuals are exact in float (a float minus
 * its own bf16 rounding has at most 16 significant bits left).
 * Two rules keep the split exact at the ends of the range: a finite x whose
 * RTNE would overflow to infinity keeps its truncated leading term, and a
 * non-finite x gets zero residuals (t0 alone carries the inf / NaN).      */
ORC_API void orc_bf16_split(const float* x, int64_t n, int terms, uint16_t* t0, uint16_t* t1, uint16_t* t2) {
  for (int64_t i = 0; i < n; ++i) {
    float r = x[i];
    uint32_t u;
    memcpy(&u, &r, 4);
    int finite = (u & 0x7f800000u) != 0x7f800000u;
    uint16_t h = orc_bf16_rtne(r);
    if (finite && (h & 0x7fffu) == 0x7f80u) h = (uint16_t)(u >> 16);
    t0[i] = h;
    if (terms < 2) continue;
    if (!finite) {
      t1[i] = 0;
      if (terms > 2) t2[i] = 0;
      continue;
    }
    r = r - orc_bf16_to_f32(h);
    h = orc_bf16_rtne(r);
    t1[i] = h;
    if (terms < 3) continue;
    r = r - orc_bf16_to_f32(h);
    t2[i] = orc_bf16_rtne(r);
  }
}

/* Long-double value of a split product restricted to a set of cross
 * products.  Bit p*3+q of `pairs` keeps the product t_p(op(A)) * t_q(op(B))
 * (p indexes the terms of A, q those of B).  The split product of
 * gemm_split.hip keeps the pairs with p + q < terms (hi*hi, hi*mid, mid*hi
 * for 2 terms; plus hi*lo, lo*hi, mid*mid for 3 terms): this is the value it
 * computes before its own FP32 accumulation rounding.  Other masks give the
 * values a kernel with a wrong kept set would approach.
 *   out[r*n+j] = alpha * sum_k sum_{(p,q) kept} ta_p(i,k) tb_q(k,j) + beta*C0 */
ORC_API void orc_gemm_split_pairs_ref(int terms, int pairs, int ta, int tb, int64_t m, int64_t n, int64_t k,
                                      double alpha, const float* A, int64_t lda, const float* B, int64_t ldb,
                                      double beta, const float* C0, int64_t ldc, const int64_t* rows,
                                      int64_t nrows, double* out) {
  (void)m;
  long double* bt = (long double*)malloc(sizeof(long double) * 3 * (size_t)(k > 0 ? k : 1));
  for (int64_t j = 0; j < n; ++j) {
    for (int64_t kk = 0; kk < k; ++kk) {
      float x = tb ? B[j + kk * ldb] : B[kk + j * ldb];
      uint16_t s0 = 0, s1 = 0, s2 = 0;
      orc_bf16_split(&x, 1, terms, &s0, &s1, &s2);
      bt[3 * kk + 0] = orc_bf16_to_f32(s0);
      bt[3 * kk + 1] = terms > 1 ? orc_bf16_to_f32(s1) : 0.0L;
      bt[3 * kk + 2] = terms > 2 ? orc_bf16_to_f32(s2) : 0.0L;
    }
    for (int64_t r = 0; r < nrows; ++r) {
      int64_t i = rows ? rows[r] : r;
      long double s = 0.0L;
      for (int64_t kk = 0; kk < k; ++kk) {
        float x = ta ? A[kk + i * lda] : A[i + kk * lda];
        uint16_t s0 = 0, s1 = 0, s2 = 0;
        orc_bf16_split(&x, 1, terms, &s0, &s1, &s2);
        long double a[3] = {orc_bf16_to_f32(s0), terms > 1 ? orc_bf16_to_f32(s1) : 0.0L,
                            terms > 2 ? orc_bf16_to_f32(s2) : 0.0L};
        for (int p = 0; p < 3; ++p)
          for (int q = 0; q < 3; ++q)
            if (pairs >> (3 * p + q) & 1) s += a[p] * bt[3 * kk + q];
      }
      long double v = (long double)alpha * s;
      if (beta != 0.0) v += (long double)beta * (long double)C0[i + j * ldc];
      out[r * n + j] = (double)v;
    }
  }
  free(bt);
}

/* The kept set of the split product with `terms` terms: pairs p + q < terms. */
ORC_API int orc_split_kept_pairs(int terms) {
  int mask = 0;
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q)
      if (p + q < terms) mask |= 1 << (3 * p + q);
  return mask;
}

ORC_API void orc_gemm_split_ref(int terms, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha,
                                const float* A, int64_t lda, const float* B, int64_t ldb, double beta,
                                const float* C0, int64_t ldc, const int64_t* rows, int64_t nrows, double* out) {
  orc_gemm_split_pairs_ref(terms, orc_split_kept_pairs(terms), ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C0,
                           ldc, rows, nrows, out);
}

ORC_API void orc_bf16_rtne_n(const float* x, int64_t n, uint16_t* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = orc_bf16_rtne(x[i]);
}

/* ------------------------------------------------------------------------ */
/* IEEE binary16: float -> half round to nearest, ties to even (the           */
/* conversion Eigen::half(float) performs, float_to_half_rtne in               */
/* Eigen/src/Core/arch/Default/Half.h:683), restated from the IEEE 754         */
/* format: 5 exponent bits (bias 15), 10 fraction bits, subnormals down to     */
/* 2^-24, overflow to infinity at and above 65520 (the tie between 65504 and   */
/* 2^16 rounds to the even 2^16), NaN kept quiet with its top fraction bits.   */
/* ------------------------------------------------------------------------ */
ORC_API uint16_t orc_half_rtne(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  const uint16_t sign = (uint16_t)((u >> 16) & 0x8000u);
  const uint32_t a = u & 0x7fffffffu;
  if (a > 0x7f800000u) return (uint16_t)(sign | 0x7e00u | ((a >> 13) & 0x3ffu)); /* NaN */
  if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);                         /* inf or overflow */
  const uint32_t e = a >> 23;                                                       /* biased float exponent */
  uint32_t q, rest, half_ulp;
  if (e >= 113) { /* normal half: exponent e - 112, fraction = top 10 of 23 bits */
    q = ((e - 112) << 10) | ((a >> 13) & 0x3ffu);
    rest = a & 0x1fffu;
    half_ulp = 0x1000u;
  } else { /* subnormal half (or zero): count units of 2^-24 */
    if (e < 102) return sign; /* below 2^-25: rounds to zero (2^-25 itself is e = 102) */
    const uint32_t m = (a & 0x7fffffu) | 0x800000u;
    const uint32_t shift = 126 - e; /* value = m * 2^(e - 150) = (m >> shift) units of 2^-24 */
    q = m >> shift;
    rest = m & ((1u << shift) - 1u);
    half_ulp = 1u << (shift - 1);
  }
  if (rest > half_ulp || (rest == half_ulp && (q & 1u))) ++q; /* may carry into the exponent */
  return (uint16_t)(sign | q);
}

ORC_API float orc_half_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1fu, frac = h & 0x3ffu;
  uint32_t u;
  if (e == 0x1f) {
    u = sign | 0x7f800000u | (frac << 13);
  } else if (e != 0) {
    u = sign | ((e + 112) << 23) | (frac << 13);
  } else {
    float v = (float)frac * 5.9604644775390625e-08f; /* frac * 2^-24, exact */
    memcpy(&u, &v, 4);
    u |= sign;
  }
  float f;
  memcpy(&f, &u, 4);
  return f;
}

ORC_API void orc_half_rtne_n(const float* x, int64_t n, uint16_t* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = orc_half_rtne(x[i]);
}

ORC_API void orc_half_to_f32_n(const uint16_t* h, int64_t n, float* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = orc_half_to_f32(h[i]);
}
