/*
 * eigen_amd.h -- C-ABI of the MI355X (gfx950) dense matrix-product library.
 *
 * Plain pointers and sizes only.  Matrices are column-major; A, B, C are
 * device pointers (hipMalloc'd or any device-accessible memory); alpha and
 * beta point to ONE host scalar of the call's dtype; `stream` is a
 * hipStream_t (NULL = the default stream).  Calls are asynchronous on the
 * stream unless stated otherwise and return a status code, never throw.
 *
 * Reference interfaces replaced (paths inside the Eigen source tree):
 *   eh_gemm / eh_gemm_config
 *       Eigen::internal::general_matrix_matrix_product<...>::run
 *       (Eigen/src/Core/products/GeneralMatrixMatrix.h:117-246) together with
 *       gemm_blocking_space / computeProductBlockingSizes
 *       (GeneralMatrixMatrix.h:282-416, GeneralBlockPanelKernel.h:201-452):
 *       the whole GEBP product C = alpha*op(A)*op(B) + beta*C, where the
 *       reference's beta handling lives one level up in ?gemm_
 *       (blas/level3_impl.h:80-95) and is fused here.
 *   eh_gemv
 *       ?gemv_ (blas/level2_impl.h:27-84) / general_matrix_vector_product
 *       (GeneralMatrixVector.h:71-289).
 *   eh_gemm_complex
 *       cgemm_/zgemm_ and Eigen's complex GEBP (GeneralBlockPanelKernel.h:631-1050).
 *   eh_gemm_bf16
 *       the sbgemm redirect of bfloat16 products
 *       (Eigen/src/Core/products/GeneralMatrixMatrix_BLAS.h:120-202).
 *   eh_gemm_f16
 *       the product of Eigen::half matrices (GEBP on half scalars,
 *       GeneralBlockPanelKernel.h with packet_traits<half>).
 *   eh_gemm_strided
 *       the same driver with a result increment (resIncr,
 *       GeneralMatrixMatrix.h:98-137) for destinations with an inner stride.
 *   eh_syrk
 *       ?syrk_ (blas/level3_impl.h:594-689) /
 *       general_matrix_matrix_triangular_product (GeneralMatrixMatrixTriangular.h).
 *   eh_gemmtr / eh_syr2k / eh_symm / eh_trmm / eh_trsm
 *       ?gemmtr_ / ?syr2k_ / ?symm_ / ?trmm_ / ?trsm_ (blas/level3_impl.h:98-758)
 *       and the level-3 kernels they run (GeneralMatrixMatrixTriangular.h,
 *       SelfadjointMatrixMatrix.h, TriangularMatrixMatrix.h,
 *       TriangularSolverMatrix.h), on the same MFMA product kernels.
 *   eh_plan / eh_set_tune_mode
 *       computeProductBlockingSizes + evaluateProductBlockingSizesHeuristic
 *       (GeneralBlockPanelKernel.h:201-452): the blocking choice, here a
 *       tile/split-K configuration picked by a heuristic, a trained MLP or an
 *       exhaustive timing table.
 *   eh_split_bf16
 *       float -> bfloat16 RTNE (Eigen/src/Core/arch/Default/BFloat16.h:434)
 *       applied term by term: the split used by the FP32-accurate BF16 path.
 *   eh_grouped_*, eh_gemm_batched_strided
 *       no single reference function: Eigen runs many small products as a
 *       host loop of general_matrix_matrix_product / lazy products
 *       (Eigen/src/Core/ProductEvaluators.h:606-692); here one persistent
 *       kernel walks a device work list.
 *   eh_mlp_*
 *       no reference counterpart (the learned replacement of the blocking
 *       heuristic).
 * The BLAS entry points sgemm_/dgemm_ are declared in eigen_amd_blas.h.
 */
#ifndef EIGEN_AMD_H
#define EIGEN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum eh_status {
  EH_SUCCESS = 0,
  EH_ERR_INVALID = 1,     /* bad argument (see eh_last_error) */
  EH_ERR_HIP = 2,         /* a HIP runtime call failed */
  EH_ERR_UNSUPPORTED = 3, /* valid request this build does not handle */
  EH_ERR_ALLOC = 4        /* device allocation failed */
} eh_status;

typedef enum eh_dtype { EH_F32 = 0, EH_F64 = 1 } eh_dtype;

/* op(X): N = X, T = X^T, C = X^H (== X^T for real scalars). */
typedef enum eh_op { EH_OP_N = 0, EH_OP_T = 1, EH_OP_C = 2 } eh_op;

/* Arithmetic of an FP32 product.  AUTO = the library default for FP32
 * (eh_set_sgemm_precision, or EIGEN_AMD_SGEMM_PRECISION=plain|split2|split3
 * in the environment; PLAIN when unset).  FP64 products always use FP64
 * MFMA and accept AUTO or PLAIN only. */
typedef enum eh_precision {
  EH_PREC_AUTO = 0,
  EH_PREC_PLAIN = 1,  /* FP32 MFMA (v_mfma_f32_32x32x2_f32) */
  EH_PREC_SPLIT2 = 2, /* 2 bf16 terms, 3 cross products on BF16 MFMA */
  EH_PREC_SPLIT3 = 3  /* 3 bf16 terms, 6 cross products on BF16 MFMA */
} eh_precision;

/* Tile / schedule selection for the plain kernels.  Default: MLP when the
 * library was built with a trained selector (eigen_amd/csrc/selector_weights.inc),
 * else HEURISTIC; EIGEN_AMD_TUNE=heuristic|mlp|table overrides. */
typedef enum eh_tune_mode {
  EH_TUNE_HEURISTIC = 0, /* closed-form rule (the no-MLP method) */
  EH_TUNE_MLP = 1,       /* trained selector (falls back to HEURISTIC without weights) */
  EH_TUNE_TABLE = 2      /* exhaustive timing table (tunes unseen shapes on first use) */
} eh_tune_mode;

/* ---- products ------------------------------------------------------------ */

/* C = alpha*op(A)*op(B) + beta*C.  m x n result, inner dimension k.
 * beta == 0: C is written without being read (NaN/Inf in C are discarded).
 * alpha == 0 or k == 0: A and B are not read; C = beta*C.
 * m == 0 or n == 0: nothing happens. */
eh_status eh_gemm(eh_dtype dtype, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, const void* alpha,
                  const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc,
                  eh_precision prec, void* stream);

/* eh_gemm with no device allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under hipGraph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * dispatch, kernels and results (bit for bit) as eh_gemm; the workspace
 * eh_gemm would take from its per-stream pool (split-K partial sums, the
 * packed bf16 planes of a split-precision product, matrix-vector partial
 * sums) is instead one fixed device array of EH_NOALLOC_SCRATCH_BYTES per
 * GPU, part of the library and resident until the process ends.  A product
 * that needs more scratch than that returns EH_ERR_UNSUPPORTED without
 * launching anything (e.g. a split-BF16 product needs 6 bytes per element of
 * op(A) and op(B) with 3 terms, 4 with 2).  The array is shared by all calls
 * on one GPU: calls that overlap in time on the same GPU race on it, so the
 * caller serialises them (one stream, or an event between streams).  The
 * planner never tunes inside this call (EH_TUNE_TABLE falls back to the
 * network or the heuristic for shapes it has not timed). */
#define EH_NOALLOC_SCRATCH_BYTES (64u << 20)
eh_status eh_gemm_noalloc(eh_dtype dtype, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, const void* alpha,
                          const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C,
                          int64_t ldc, eh_precision prec, void* stream);

/* Same, with an explicit configuration index (config < 0 = let the planner
 * choose): in [0, eh_num_configs()) for the plain kernels, in
 * [0, eh_num_packed_configs(EH_PACKED_SPLIT2 / SPLIT3)) when prec selects
 * the split-BF16 product. */
eh_status eh_gemm_config(eh_dtype dtype, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, const void* alpha,
                         const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C,
                         int64_t ldc, eh_precision prec, int config, void* stream);

/* C(i, j) at C[i*incc + j*ldc]: a destination whose consecutive rows are
 * incc elements apart (incc >= 1, ldc >= (m-1)*incc + 1).  This is the
 * result increment of Eigen's GEMM driver (resIncr of
 * general_matrix_matrix_product<...>::run, GeneralMatrixMatrix.h:98-137,
 * addressed through blas_data_mapper<..., Incr>, BlasUtil.h:266-443), which
 * Eigen uses for destinations with a non-unit inner stride
 * (test/product.h:345-357).  incc == 1 is eh_gemm.  Otherwise the product
 * op(A)*op(B) is formed in a stream-ordered temporary and combined into C
 * with the same fused alpha/beta arithmetic as eh_gemm's epilogue, so the
 * result is bit-identical to eh_gemm on a contiguous copy of C. */
eh_status eh_gemm_strided(eh_dtype dtype, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, const void* alpha,
                          const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C,
                          int64_t incc, int64_t ldc, eh_precision prec, void* stream);

/* Matrix-vector product with strided vectors (the BLAS ?gemv_ contract,
 * blas/level2_impl.h:27-84): A is m x n;
 *   trans = N:  y = alpha*A*x + beta*y     (x: n elements, y: m)
 *   trans = T:  y = alpha*A^T*x + beta*y   (x: m elements, y: n)
 * x(i) = x[i*incx] and y(i) = y[i*incy]; a negative increment starts at the
 * last element, as in BLAS.  beta == 0 does not read y; alpha == 0 does not
 * read A or x.  This is also the kernel Eigen's products with one output
 * column or row reach (GeneralMatrixMatrix.h:480-490 ->
 * GeneralMatrixVector.h:71-289), which eh_gemm takes for n == 1 or m == 1. */
eh_status eh_gemv(eh_dtype dtype, eh_op trans, int64_t m, int64_t n, const void* alpha, const void* A, int64_t lda,
                  const void* x, int64_t incx, const void* beta, void* y, int64_t incy, void* stream);

/* bfloat16 inputs, FP32 result: C = alpha*op(A)*op(B) + beta*C with A and B
 * bfloat16 (raw 16-bit patterns, column-major, any element alignment) and
 * alpha, beta, C float.  Products of bf16 values are exact in FP32 and are
 * accumulated in FP32 on the BF16 matrix cores (v_mfma_f32_32x32x16_bf16).
 * This is the sbgemm contract that Eigen binds for bfloat16 products under
 * EIGEN_USE_OPENBLAS_BFLOAT16 (GeneralMatrixMatrix_BLAS.h:120-202), which
 * then rounds the FP32 result to bfloat16 itself; beta == 0 does not read C,
 * alpha == 0 or k == 0 does not read A or B.  Operands must span less than
 * 4 GB (EH_ERR_UNSUPPORTED otherwise).  Operands are repacked into a
 * stream workspace, except those whose op(X) rows are k-contiguous (A with
 * opa T, B with opb N) when k is a multiple of 32 and the rows are 16-byte
 * aligned (8 | ld, 16-byte aligned base): those are read in place. */
eh_status eh_gemm_bf16(eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, const float* alpha, const void* A,
                       int64_t lda, const void* B, int64_t ldb, const float* beta, float* C, int64_t ldc,
                       void* stream);

/* IEEE half (binary16) inputs, FP32 result: C = alpha*op(A)*op(B) + beta*C
 * with A and B half bit patterns (column-major, any element alignment) and
 * alpha, beta, C float.  Products of half values are exact in FP32 and are
 * accumulated in FP32 on the FP16 matrix cores (v_mfma_f32_32x32x16_f16), on
 * the same packed planes and tiles as eh_gemm_bf16.  Replaces the product of
 * Eigen::half matrices (GEBP over half scalars, exercised by
 * test/product_large.cpp:366), which on hosts without native half arithmetic
 * rounds every multiply-add to half (pmadd<Packet8h>,
 * arch/AVX/PacketMath.h:2328); here only the final FP32 result is rounded,
 * and a caller that wants half output converts it (Eigen::half(float),
 * Half.h:683).  beta == 0 does not read C; alpha == 0 or k == 0 does not read
 * A or B.  Operands must span less than 4 GB (EH_ERR_UNSUPPORTED otherwise);
 * they are repacked or read in place as for eh_gemm_bf16. */
eh_status eh_gemm_f16(eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, const float* alpha, const void* A,
                      int64_t lda, const void* B, int64_t ldb, const float* beta, float* C, int64_t ldc,
                      void* stream);

/* Complex product C = alpha*op(A)*op(B) + beta*C on interleaved complex
 * storage (std::complex<float> for EH_F32, std::complex<double> for EH_F64;
 * leading dimensions in complex elements; alpha and beta point to one host
 * complex scalar, i.e. two reals).  op = C conjugates.  Replaces
 * cgemm_/zgemm_ (blas/level3_impl.h:12-96 instantiated for complex scalars)
 * and Eigen's complex GEBP (GeneralBlockPanelKernel.h:631-1050).  Runs as ONE
 * real product of order (2m) x n x (2k) on the FP32 / FP64 MFMA kernels: each
 * complex entry of alpha*op(A) is repacked as the 2x2 real block
 * [Re -Im; Im Re], B (when op(B) = B) and C are used in place as real
 * matrices with twice the leading dimension.  FP32 follows `prec` like
 * eh_gemm.  beta == 0 does not read C; alpha == 0 or k == 0 does not read A
 * or B.  m and k must be below 2^30. */
eh_status eh_gemm_complex(eh_dtype dtype, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, const void* alpha,
                          const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C,
                          int64_t ldc, eh_precision prec, void* stream);

/* Complex matrix-vector product with strided vectors (the cgemv_/zgemv_
 * contract, blas/level2_impl.h:27-84 for complex scalars): A m x n
 * interleaved complex, y = alpha*op(A)*x + beta*y, op = C conjugates, alpha
 * and beta point to one host complex scalar.  A is read once: its real view
 * (2m x n, leading dimension 2*lda) multiplies two real right-hand sides
 * built from x on the FP32 / FP64 MFMA kernels, and a combine pass forms y.
 * Increments as eh_gemv; m and n must be below 2^30. */
eh_status eh_gemv_complex(eh_dtype dtype, eh_op trans, int64_t m, int64_t n, const void* alpha, const void* A,
                          int64_t lda, const void* x, int64_t incx, const void* beta, void* y, int64_t incy,
                          void* stream);

/* Triangle of a symmetric result. */
typedef enum { EH_UPPER = 0, EH_LOWER = 1 } eh_uplo;

/* Symmetric rank-k update of one triangle of C (n x n):
 *   trans = N:  C = alpha*A*A^T + beta*C   (A n x k)
 *   trans = T:  C = alpha*A^T*A + beta*C   (A k x n)
 * Only the `uplo` triangle of C (diagonal included) is read or written.
 * Replaces ?syrk_ (blas/level3_impl.h:594-689) and the rank update it runs,
 * general_matrix_matrix_triangular_product
 * (Eigen/src/Core/products/GeneralMatrixMatrixTriangular.h:38-330), which
 * Eigen's selfadjointView<>().rankUpdate() and triangularView<>() += A*A^T
 * reach, and which GeneralMatrixMatrixTriangular_BLAS.h:80-106 redirects to
 * ?syrk_ under EIGEN_USE_BLAS.  Runs the plain FP32 / FP64 MFMA kernels with
 * the tiles outside the triangle skipped and the epilogue masked; alpha == 0
 * or k == 0 scales the triangle without reading A. */
eh_status eh_syrk(eh_dtype dtype, eh_uplo uplo, eh_op trans, int64_t n, int64_t k, const void* alpha, const void* A,
                  int64_t lda, const void* beta, void* C, int64_t ldc, void* stream);

/* ---- other level-3 routines on the product kernels -----------------------
 * All take device pointers and host scalars, run on `stream` and return
 * without synchronising.  Temporaries (a dense copy of a triangle, a copy of
 * a block of B) come from the stream-ordered allocator.  FP32 runs on FP32
 * MFMA (the split-BF16 precision applies to eh_gemm only). */

typedef enum { EH_LEFT = 0, EH_RIGHT = 1 } eh_side;
typedef enum { EH_NON_UNIT = 0, EH_UNIT = 1 } eh_diag;

/* Product with a triangle of C: C = alpha*op(A)*op(B) + beta*C on the `uplo`
 * triangle of the n x n result only (op(A) n x k, op(B) k x n).  Replaces
 * ?gemmtr_ (blas/level3_impl.h:98-228) and the
 * general_matrix_matrix_triangular_product it runs
 * (GeneralMatrixMatrixTriangular.h:38-330). */
eh_status eh_gemmtr(eh_dtype dtype, eh_uplo uplo, eh_op opa, eh_op opb, int64_t n, int64_t k, const void* alpha,
                    const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc,
                    void* stream);

/* Symmetric rank-2k update of one triangle of C (n x n):
 *   trans = N:  C = alpha*A*B^T + alpha*B*A^T + beta*C   (A, B n x k)
 *   trans = T:  C = alpha*A^T*B + alpha*B^T*A + beta*C   (A, B k x n)
 * Replaces ?syr2k_ (blas/level3_impl.h:690-758), which evaluates the two
 * products into the triangle (triangularView<>() += ...); here two masked
 * products, the second accumulating onto the first. */
eh_status eh_syr2k(eh_dtype dtype, eh_uplo uplo, eh_op trans, int64_t n, int64_t k, const void* alpha, const void* A,
                   int64_t lda, const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc, void* stream);

/* Symmetric product, A symmetric and stored in its `uplo` triangle:
 *   side = LEFT:   C = alpha*A*B + beta*C   (A m x m)
 *   side = RIGHT:  C = alpha*B*A + beta*C   (A n x n)
 * Replaces ?symm_ (blas/level3_impl.h:505-592) and
 * product_selfadjoint_matrix (SelfadjointMatrixMatrix.h:323-520), whose
 * packing mirrors the stored triangle (symm_pack_lhs/rhs, :20-250); here a
 * tiled kernel writes the mirrored dense matrix once (HBM-bound, order^2
 * elements) and the product kernels take it as a plain operand. */
eh_status eh_symm(eh_dtype dtype, eh_side side, eh_uplo uplo, int64_t m, int64_t n, const void* alpha, const void* A,
                  int64_t lda, const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc, void* stream);

/* Triangular product in place, A triangular in its `uplo` triangle (unit
 * diagonal not read when diag = UNIT):
 *   side = LEFT:   B = alpha*op(A)*B   (A m x m)
 *   side = RIGHT:  B = alpha*B*op(A)   (A n x n)
 * Replaces ?trmm_ (blas/level3_impl.h:366-503) and
 * product_triangular_matrix_matrix (TriangularMatrixMatrix.h:84-420).
 * Recursive 2x2 blocking: the off-diagonal blocks are plain products
 * accumulated in place; a diagonal block of order <= 256 is made dense (zeros
 * off the triangle) and multiplied from a copy of its rows (columns) of B.
 * alpha == 0 sets B to zero without reading A (the netlib rule). */
eh_status eh_trmm(eh_dtype dtype, eh_side side, eh_uplo uplo, eh_op transa, eh_diag diag, int64_t m, int64_t n,
                  const void* alpha, const void* A, int64_t lda, void* B, int64_t ldb, void* stream);

/* Triangular solve in place, A triangular in its `uplo` triangle:
 *   side = LEFT:   op(A)*X = alpha*B   (A m x m)
 *   side = RIGHT:  X*op(A) = alpha*B   (A n x n)
 * X overwrites B.  Replaces ?trsm_ (blas/level3_impl.h:228-364) and
 * triangular_solve_matrix (TriangularSolverMatrix.h:20-330).  Recursive 2x2
 * blocking: a block of unknowns is solved, its contribution is subtracted
 * from the remaining right-hand sides by a product on the MFMA kernels.  The
 * 64 x 64 diagonal blocks are inverted up front by one batched launch
 * (substitution, multiplying by the reciprocal of the diagonal as the
 * reference does); for real types the inverses are then doubled to order
 * 128 and 256 while that order divides the triangle's (two batched products
 * per doubling), and a leaf multiplies its right-hand sides by the inverse
 * of its block on the MFMA kernels, from a copy of them.
 * alpha is applied to the right-hand sides before the solve (the reference
 * scales the solution afterwards: equal up to rounding); alpha == 0 sets B
 * to zero without reading A.  No check for a zero diagonal: as in the
 * reference, it produces infinities. */
eh_status eh_trsm(eh_dtype dtype, eh_side side, eh_uplo uplo, eh_op transa, eh_diag diag, int64_t m, int64_t n,
                  const void* alpha, const void* A, int64_t lda, void* B, int64_t ldb, void* stream);

/* ---- complex level-3 routines ---------------------------------------------
 * The complex counterparts of the routines above, on interleaved complex
 * storage (std::complex<float> for EH_F32, std::complex<double> for EH_F64;
 * leading dimensions in complex elements; alpha and beta point to one host
 * complex scalar -- two reals -- unless stated "real").  Replace Eigen's
 * complex BLAS instantiations of blas/level3_impl.h (?trsm_ :228-364, ?trmm_
 * :366-503, ?symm_ :505-592, ?syrk_ :594-689, ?syr2k_ :690-758, ?gemmtr_
 * :98-228, ?hemm_ :760-833, ?herk_ :835-912, ?her2k_ :915-986) and the
 * Eigen products they run (triangularView<>().solve(), selfadjointView<>()
 * products and rank updates on complex matrices).  Every product runs as one
 * real product over 2x2 real blocks on the FP32 / FP64 MFMA kernels
 * (eh_gemm_complex); op C conjugates.  Same device / stream / temporary
 * rules as the real routines; dimensions below 2^30.
 *   eh_trsm_complex, eh_trmm_complex: as eh_trsm / eh_trmm (transa C: op(A)
 *     = A^H); the TRSM leaves invert 32 x 32 diagonal blocks.
 *   eh_symm_complex: A complex symmetric (mirror not conjugated).
 *   eh_hemm: A Hermitian, stored in its uplo triangle (mirror conjugated,
 *     the imaginary part of its diagonal not read).
 *   eh_syrk_complex:  C = alpha*op(A)*op(A)^T + beta*C, trans N or T.
 *   eh_herk:   C = alpha*op(A)*op(A)^H + beta*C, trans N or C, alpha and
 *     beta REAL (one host real each); the imaginary parts of C's diagonal
 *     are set to zero unless the call leaves C unchanged.
 *   eh_syr2k_complex: C = alpha*op(A)*op(B)^T + alpha*op(B)*op(A)^T + beta*C,
 *     trans N or T.
 *   eh_her2k: C = alpha*op(A)*op(B)^H + conj(alpha)*op(B)*op(A)^H + beta*C,
 *     trans N or C, beta REAL; diagonal imaginary parts as eh_herk.
 *   eh_gemmtr_complex: C = alpha*op(A)*op(B) + beta*C on one triangle.
 * The rank updates and eh_gemmtr_complex form the full n x n product in a
 * stream-ordered temporary and combine its uplo triangle into C (one pass);
 * k == 0 or alpha == 0 scales the triangle without reading A or B. */
eh_status eh_trsm_complex(eh_dtype dtype, eh_side side, eh_uplo uplo, eh_op transa, eh_diag diag, int64_t m,
                          int64_t n, const void* alpha, const void* A, int64_t lda, void* B, int64_t ldb,
                          void* stream);
eh_status eh_trmm_complex(eh_dtype dtype, eh_side side, eh_uplo uplo, eh_op transa, eh_diag diag, int64_t m,
                          int64_t n, const void* alpha, const void* A, int64_t lda, void* B, int64_t ldb,
                          void* stream);
eh_status eh_symm_complex(eh_dtype dtype, eh_side side, eh_uplo uplo, int64_t m, int64_t n, const void* alpha,
                          const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C,
                          int64_t ldc, void* stream);
eh_status eh_hemm(eh_dtype dtype, eh_side side, eh_uplo uplo, int64_t m, int64_t n, const void* alpha, const void* A,
                  int64_t lda, const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc, void* stream);
eh_status eh_syrk_complex(eh_dtype dtype, eh_uplo uplo, eh_op trans, int64_t n, int64_t k, const void* alpha,
                          const void* A, int64_t lda, const void* beta, void* C, int64_t ldc, void* stream);
eh_status eh_herk(eh_dtype dtype, eh_uplo uplo, eh_op trans, int64_t n, int64_t k, const void* alpha, const void* A,
                  int64_t lda, const void* beta, void* C, int64_t ldc, void* stream);
eh_status eh_syr2k_complex(eh_dtype dtype, eh_uplo uplo, eh_op trans, int64_t n, int64_t k, const void* alpha,
                           const void* A, int64_t lda, const void* B, int64_t ldb, const void* beta, void* C,
                           int64_t ldc, void* stream);
eh_status eh_her2k(eh_dtype dtype, eh_uplo uplo, eh_op trans, int64_t n, int64_t k, const void* alpha, const void* A,
                   int64_t lda, const void* B, int64_t ldb, const void* beta, void* C, int64_t ldc, void* stream);
eh_status eh_gemmtr_complex(eh_dtype dtype, eh_uplo uplo, eh_op opa, eh_op opb, int64_t n, int64_t k,
                            const void* alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
                            const void* beta, void* C, int64_t ldc, void* stream);

/* ---- planning ------------------------------------------------------------ */

int eh_num_configs(void);
/* Tile (tile_m x tile_n) and split-K factor of configuration `config`. */
eh_status eh_config_info(int config, int* tile_m, int* tile_n, int* split_k);
/* Configuration the planner would use for this shape (synchronous). */
eh_status eh_plan(eh_dtype dtype, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, int* config);
eh_status eh_set_tune_mode(eh_tune_mode mode);
eh_tune_mode eh_get_tune_mode(void);
eh_status eh_set_sgemm_precision(eh_precision prec);
eh_precision eh_get_sgemm_precision(void);

/* ---- grouped products ------------------------------------------------------ */

/* One product of a batch: C = alpha*op(A)*op(B) + beta*C, column-major,
 * device pointers.  For EH_F32 products alpha and beta are rounded to float. */
typedef struct eh_gemm_problem {
  int32_t dtype; /* eh_dtype */
  int32_t opa;   /* eh_op */
  int32_t opb;   /* eh_op */
  int32_t reserved;
  int64_t m, n, k;
  int64_t lda, ldb, ldc;
  const void* A;
  const void* B;
  void* C;
  double alpha, beta;
} eh_gemm_problem;

typedef struct eh_grouped eh_grouped; /* opaque batch handle */

/* Validates `count` descriptors (host array), builds the work list on the
 * host once -- one descriptor per product grouped by operand type, the first
 * output tile of each product, and one [begin, end) run of tiles of one
 * operand type and equal estimated cost per resident workgroup -- and
 * uploads it to the device.  Products whose operands span 2^31 bytes or
 * more are kept aside for the single-product path.  Products must not write
 * overlapping C.  Under EH_TUNE_TABLE the call also times every schedule
 * (below) on the batch -- reading A and B, writing a scratch copy of the
 * results, never C -- and keeps the fastest; otherwise schedule 0. */
eh_status eh_grouped_create(const eh_gemm_problem* problems, int64_t count, eh_grouped** out);
/* All products of the work list in one launch of the persistent grouped
 * kernel, then the set-aside large products one eh_gemm each. */
eh_status eh_grouped_run(eh_grouped* g, void* stream);
/* The same launches captured once in a hipGraph (kernel nodes only) and
 * replayed on every call. */
eh_status eh_grouped_graph_launch(eh_grouped* g, void* stream);
/* Reference path: one eh_gemm launch per product. */
eh_status eh_grouped_run_per_product(eh_grouped* g, void* stream);
/* Number of 32x32 output tiles and of non-empty products. */
eh_status eh_grouped_info(eh_grouped* g, int64_t* tiles, int64_t* products);
/* Schedules of the grouped kernel: `waves_per_group` waves per workgroup
 * compute a block of one product's 32x32 tiles at a time, `band` tile
 * columns wide (0: 16 waves, 4; 1: 8 waves, 2; 2: 4 waves, 2).  Setting a
 * schedule rebuilds the runs (and drops a captured graph), so it must not
 * overlap a run of the same handle; results are bit-identical across
 * schedules. */
int eh_grouped_num_schedules(void);
eh_status eh_grouped_schedule_info(int schedule, int* waves_per_group, int* band);
eh_status eh_grouped_set_schedule(eh_grouped* g, int schedule);
eh_status eh_grouped_get_schedule(eh_grouped* g, int* schedule);
eh_status eh_grouped_destroy(eh_grouped* g);

/* `count` products of one shape with operands at fixed element strides
 * (product p: A + p*stride_a, B + p*stride_b, C + p*stride_c), e.g. arrays of
 * fixed-size matrices: the batched form of Eigen's small and fixed-size
 * products, which Eigen evaluates one at a time (coefficient-based below
 * ProductEvaluators.h:606-692's threshold, GEBP above).  Shapes up to
 * 4x4x4 run one thread per product (whole columns for arrays of dense 2x2
 * and 4x4 matrices, nine-element runs for dense 3x3 ones), shapes up to 16x16 with depth up to 64 (FP64) or 32
 * (FP32) one wave and one 16x16 MFMA block per product: one launch each.
 * Larger shapes: the work list is written on the device (one small kernel:
 * descriptors and the equal-cost split of the count*tiles 32x32 tiles over
 * the resident waves) and the grouped kernel runs it: two launches, no host
 * loop, no allocation beyond the per-stream workspace.  Results must not
 * overlap (stride_c >= (n-1)*ldc + m when count > 1).  Products whose
 * operands span 2 GB or more run one product at a time. */
eh_status eh_gemm_batched_strided(eh_dtype dtype, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k,
                                  const void* alpha, const void* A, int64_t lda, int64_t stride_a, const void* B,
                                  int64_t ldb, int64_t stride_b, const void* beta, void* C, int64_t ldc,
                                  int64_t stride_c, int64_t count, void* stream);

/* ---- autotuner ------------------------------------------------------------- */

/* Median time (ms) of `reps` runs, after one warm-up, of configuration
 * `config` for C = A*B (beta = 0) on the caller's device buffers (packed:
 * lda/ldb/ldc = rows).  dtype / opa / opb take eh_dtype / eh_op values. */
eh_status eh_time_config(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, const void* A, const void* B,
                         void* C, int config, int reps, float* ms);
/* The exhaustive table used by EH_TUNE_TABLE: shape -> best configuration. */
eh_status eh_tuner_record(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, int config);
int64_t eh_tuner_count(void);
eh_status eh_tuner_clear(void);
/* Text file: a header "# eigen_amd tuning table v2 arch=<gcnArchName>
 * kernels=<version of the configuration sets>", then one line per shape
 * "kind dtype opa opb m n k config" (kind 0 = plain kernels, else an
 * eh_packed_kind).  Loading refuses a table made for another architecture or
 * kernel set; v1 files (lines "dtype opa opb m n k config", plain kernels
 * only, no stamp) still load. */
eh_status eh_tuner_save(const char* path);
eh_status eh_tuner_load(const char* path);

/* Packed products -- the split-BF16 FP32 products (eh_gemm with
 * EH_PREC_SPLIT2 / SPLIT3) and the 16-bit input products (eh_gemm_bf16,
 * eh_gemm_f16) -- have their own configurations: tile * 4 + split-K index
 * (splits {1, 2, 4, 8}); tiles for the split kinds 0 = 128x128, 1 = 256x128
 * (4 waves), 2 = 256x128 (8 waves), 3 = 256x256 (8 waves, 16-deep stages);
 * for the 16-bit kinds 0 = 128x128, 1 = 256x256.  EH_TUNE_TABLE times every
 * configuration of a new shape on first use (scratch operands) and
 * remembers the fastest; the other modes use the fixed rule (the largest
 * tile that fills the chip, split K below one wave of workgroups).  No counterpart in the reference: Eigen's blocking
 * choice (GeneralBlockPanelKernel.h:201-452) has no per-kernel tuning. */
typedef enum eh_packed_kind {
  EH_PACKED_SPLIT2 = 2,
  EH_PACKED_SPLIT3 = 3,
  EH_PACKED_BF16 = 4,
  EH_PACKED_F16 = 5
} eh_packed_kind;
int eh_num_packed_configs(eh_packed_kind kind); /* 0 for an unknown kind */
eh_status eh_packed_config_info(eh_packed_kind kind, int config, int* tile_m, int* tile_n, int* split_k);
/* Configuration the planner would use (may tune, synchronously, in TABLE mode). */
eh_status eh_plan_packed(eh_packed_kind kind, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k, int* config);
/* As eh_time_config for a packed product: A and B are FP32 for the split
 * kinds, 16-bit (bf16 / IEEE half bits) for EH_PACKED_BF16 / F16; C FP32. */
eh_status eh_time_config_packed(eh_packed_kind kind, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k,
                                const void* A, const void* B, void* C, int config, int reps, float* ms);
eh_status eh_tuner_record_packed(eh_packed_kind kind, eh_op opa, eh_op opb, int64_t m, int64_t n, int64_t k,
                                 int config);

/* ---- learned selector ------------------------------------------------------ */

/* A 7 -> 64 (ReLU) -> 24 (softmax) network over shape features, with device
 * buffers for batches of up to `capacity` samples.  Parameters: 2072 floats,
 * W1[64][7], b1[64], W2[24][64], b2[24]. */
typedef struct eh_mlp eh_mlp;
eh_status eh_mlp_create(int inputs, int hidden, int outputs, int64_t capacity, eh_mlp** out);
eh_status eh_mlp_destroy(eh_mlp* m);
int64_t eh_mlp_num_params(eh_mlp* m);
eh_status eh_mlp_set_params(eh_mlp* m, const float* host_params);
eh_status eh_mlp_get_params(eh_mlp* m, float* host_params);
/* Features of a shape: log2 m/14, log2 n/14, log2 k/14, [f32, f64], ta, tb. */
eh_status eh_mlp_features(int dtype, int opa, int opb, int64_t m, int64_t n, int64_t k, float* x7);
/* Device X[B][7] -> device probs[B][24] (probs may be NULL). */
eh_status eh_mlp_forward(eh_mlp* m, const float* X, int64_t B, float* probs, void* stream);
/* Forward + backward of the mean cross entropy against soft targets
 * Q[B][24]: device grads[2072] and loss[1] (either may be NULL). */
eh_status eh_mlp_backward(eh_mlp* m, const float* X, const float* Q, int64_t B, float* grads, float* loss,
                          void* stream);
/* One Adam step (decoupled weight decay on weights) with the given device
 * gradients (NULL = those of the last backward). */
eh_status eh_mlp_adam_step(eh_mlp* m, const float* grads, float lr, float beta1, float beta2, float eps,
                           float weight_decay, void* stream);
/* `epochs` full-batch Adam steps, cosine-decayed learning rate; optional
 * device loss_curve[epochs].  Synchronous. */
eh_status eh_mlp_train(eh_mlp* m, const float* X, const float* Q, int64_t B, int epochs, float lr,
                       float weight_decay, float* loss_curve, void* stream);
/* Device X[B][7] -> device configs[B] (argmax). */
eh_status eh_mlp_predict(eh_mlp* m, const float* X, int64_t B, int32_t* configs, void* stream);
/* Use the network's current parameters for EH_TUNE_MLP planning (NULL
 * restores the compiled-in default network, or none). */
eh_status eh_planner_set_mlp(eh_mlp* m, void* stream);
eh_status eh_planner_set_mlp_params(const float* host_params);

/* ---- utilities ----------------------------------------------------------- */

/* Seeded synthetic data on the device: element i from splitmix64 at
 * (seed, 2i[, 2i+1]); dist 0 uniform[p0,p1], 1 normal(p0,p1),
 * 2 sign*U(1,2)*2^U(p0,p1), 3 integers in [p0,p1]. */
eh_status eh_fill(eh_dtype dtype, void* x, int64_t n, uint64_t seed, int dist, double p0, double p1, void* stream);

/* The FP32 -> bf16 term split the split-precision products apply while
 * staging: t0 = bf16(x), t1 = bf16(x - t0), t2 = bf16(x - t0 - t1)
 * (t1/t2 may be NULL when terms < 2/3).  Device pointers. */
eh_status eh_split_bf16(const float* x, int64_t n, int terms, uint16_t* t0, uint16_t* t1, uint16_t* t2,
                        void* stream);

/* Event timing of kernel launches, for measurement (no reference
 * counterpart; the role of a profiler's kernel trace, inside the process).
 * While enabled, every launch of a kernel of the listed classes that is not
 * being captured into a graph is bracketed by two HIP events recorded on the
 * stream it is launched on.  eh_kernel_timing(1) clears earlier records and
 * enables, eh_kernel_timing(0) clears and disables; eh_kernel_timing_read
 * waits for the recorded events of one class, returns the sum of their
 * elapsed times and the number of launches, and clears them. */
typedef enum eh_kernel_class {
  EH_KC_PRODUCT = 0,        /* plain FP32 / FP64 MFMA product kernels */
  EH_KC_PACKED_PRODUCT = 1, /* split-BF16 and bf16 products over packed planes */
  EH_KC_PACK = 2,           /* split / bf16 operand packing */
  EH_KC_GROUPED = 3,        /* persistent grouped kernel, fixed-size batched kernel */
  EH_KC_GEMV = 4,           /* matrix-vector kernels */
  EH_KC_COUNT = 5
} eh_kernel_class;
eh_status eh_kernel_timing(int enable);
eh_status eh_kernel_timing_read(eh_kernel_class cls, double* total_ms, int64_t* launches);

eh_status eh_synchronize(void* stream);
const char* eh_status_string(eh_status s);
/* Message of the last failing call on this thread ("" if none). */
const char* eh_last_error(void);
const char* eh_version(void);

#ifdef __cplusplus
}
#endif

#endif /* EIGEN_AMD_H */
