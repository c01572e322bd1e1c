/*
 * eigen_amd_blas.h -- BLAS level-3 entry points of the MI355X library.
 *
 * Same symbols, argument lists, validation order, error codes and quick
 * returns as Eigen's BLAS (blas/level3_impl.h:12-96, declared in
 * Eigen/src/misc/blas.h:61-64 and blas/blas.h): Fortran calling convention,
 * every argument by pointer, 32-bit integers, column-major storage.
 *
 *   info 1: transa not N/T/C      info 2: transb not N/T/C
 *   info 3: m < 0   info 4: n < 0   info 5: k < 0
 *   info 8: lda < max(1, transa==N ? m : k)
 *   info 10: ldb < max(1, transb==N ? k : n)
 *   info 13: ldc < max(1, m)
 * reported through xerbla_("SGEMM " / "DGEMM ", &info, 6) -- a weak symbol,
 * so callers (e.g. the netlib testers) can supply their own.
 *
 * a, b, c may be device pointers (the product runs in place on the GPU) or
 * host pointers (operands are staged through device buffers).  The call runs
 * on the stream set by eh_blas_set_stream (default stream if unset) and
 * returns when C is complete, as BLAS callers expect.
 *
 * sgemm_ runs on FP32 MFMA by default; eh_set_sgemm_precision (or
 * EIGEN_AMD_SGEMM_PRECISION=split3) switches it to the BF16-split product.
 */
#ifndef EIGEN_AMD_BLAS_H
#define EIGEN_AMD_BLAS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

void sgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k, const float* alpha,
            const float* a, const int* lda, const float* b, const int* ldb, const float* beta, float* c,
            const int* ldc);
void dgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k, const double* alpha,
            const double* a, const int* lda, const double* b, const int* ldb, const double* beta, double* c,
            const int* ldc);
/* Symmetric rank-k update of the uplo triangle of C: C = alpha*op(A)*op(A)^T + beta*C
 * (trans 'N': A n x k; 'T'/'C': A k x n).  Same symbols, arguments, validation
 * order and info codes (1 uplo, 2 trans, 3 n, 4 k, 7 lda, 10 ldc) as Eigen's
 * ?syrk_ (blas/level3_impl.h:594-689), which Eigen's rank updates reach under
 * EIGEN_USE_BLAS (GeneralMatrixMatrixTriangular_BLAS.h:80-106).  Device or
 * host pointers, as for ?gemm_. */
void ssyrk_(const char* uplo, const char* trans, const int* n, const int* k, const float* alpha, const float* a,
            const int* lda, const float* beta, float* c, const int* ldc);
void dsyrk_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha, const double* a,
            const int* lda, const double* beta, double* c, const int* ldc);
/* The other real level-3 routines, with the symbols, arguments, validation
 * order, info codes and quick returns of Eigen's BLAS; device or host
 * pointers as for ?gemm_, the call returns when the output is complete.
 *   ?gemmtr_  blas/level3_impl.h:98-228   info 1 uplo, 2 transa, 3 transb, 4 n,
 *             5 k, 8 lda, 10 ldb, 13 ldc (name "SGEMMTR", 7 characters)
 *   ?syr2k_   blas/level3_impl.h:690-758  info 1 uplo, 2 trans, 3 n, 4 k,
 *             7 lda, 9 ldb, 12 ldc
 *   ?symm_    blas/level3_impl.h:505-592  info 1 side, 2 uplo, 3 m, 4 n,
 *             7 lda, 9 ldb, 12 ldc
 *   ?trmm_    blas/level3_impl.h:366-503  info 1 side, 2 uplo, 3 transa,
 *   ?trsm_    blas/level3_impl.h:228-364  4 diag, 5 m, 6 n, 9 lda, 11 ldb
 * Only the uplo triangle of C is written by ?gemmtr_ / ?syr2k_; host C
 * travels whole, so the other triangle comes back unchanged. */
void sgemmtr_(const char* uplo, const char* transa, const char* transb, const int* n, const int* k,
              const float* alpha, const float* a, const int* lda, const float* b, const int* ldb, const float* beta,
              float* c, const int* ldc);
void dgemmtr_(const char* uplo, const char* transa, const char* transb, const int* n, const int* k,
              const double* alpha, const double* a, const int* lda, const double* b, const int* ldb,
              const double* beta, double* c, const int* ldc);
void ssyr2k_(const char* uplo, const char* trans, const int* n, const int* k, const float* alpha, const float* a,
             const int* lda, const float* b, const int* ldb, const float* beta, float* c, const int* ldc);
void dsyr2k_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha, const double* a,
             const int* lda, const double* b, const int* ldb, const double* beta, double* c, const int* ldc);
void ssymm_(const char* side, const char* uplo, const int* m, const int* n, const float* alpha, const float* a,
            const int* lda, const float* b, const int* ldb, const float* beta, float* c, const int* ldc);
void dsymm_(const char* side, const char* uplo, const int* m, const int* n, const double* alpha, const double* a,
            const int* lda, const double* b, const int* ldb, const double* beta, double* c, const int* ldc);
void strmm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const float* alpha, const float* a, const int* lda, float* b, const int* ldb);
void dtrmm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const double* alpha, const double* a, const int* lda, double* b, const int* ldb);
void strsm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const float* alpha, const float* a, const int* lda, float* b, const int* ldb);
void dtrsm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const double* alpha, const double* a, const int* lda, double* b, const int* ldb);
/* Matrix-vector product y = alpha*op(A)*x + beta*y (blas/level2_impl.h:27-84,
 * the routine GeneralMatrixVector_BLAS.h:125-126 redirects Eigen's
 * matrix-vector products to under EIGEN_USE_BLAS): info 1 trans, 2 m, 3 n,
 * 6 lda, 8 incx, 11 incy (names "SGEMV " / "DGEMV "); quick return for
 * m == 0, n == 0, or alpha == 0 with beta == 1; negative increments walk
 * the vectors backwards.  Device or host pointers as for ?gemm_. */
void sgemv_(const char* trans, const int* m, const int* n, const float* alpha, const float* a, const int* lda,
            const float* x, const int* incx, const float* beta, float* y, const int* incy);
void dgemv_(const char* trans, const int* m, const int* n, const double* alpha, const double* a, const int* lda,
            const double* x, const int* incx, const double* beta, double* y, const int* incy);
/* Complex GEMM: alpha, beta, a, b, c are complex (pairs of float / double,
 * Fortran COMPLEX / COMPLEX*16 layout).  Same validation order and info codes
 * as ?gemm_ (names "CGEMM " / "ZGEMM "); op 'C' conjugates. */
void cgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k, const void* alpha,
            const void* a, const int* lda, const void* b, const int* ldb, const void* beta, void* c, const int* ldc);
void zgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k, const void* alpha,
            const void* a, const int* lda, const void* b, const int* ldb, const void* beta, void* c, const int* ldc);
/* Complex level-3 routines (Fortran COMPLEX / COMPLEX*16 arguments as
 * pairs of float / double), with the symbols, arguments, validation order,
 * info codes and quick returns of Eigen's complex BLAS (blas/level3_impl.h):
 *   c/z trsm_, trmm_, symm_, syr2k_, gemmtr_: as the real routines;
 *   c/z syrk_ (:594-689): trans 'C' is rejected (info 2), as in the reference;
 *   c/z hemm_ (:760-833): info 1 side, 2 uplo, 3 m, 4 n, 7 lda, 9 ldb, 12 ldc;
 *   c/z herk_ (:835-912): alpha and beta REAL; trans 'T' rejected (info 2);
 *       info 1 uplo, 3 n, 4 k, 7 lda, 10 ldc;
 *   c/z her2k_ (:915-986): beta REAL; trans 'T' rejected; info 1 uplo, 2 trans,
 *       3 n, 4 k, 7 lda, 9 ldb, 12 ldc.
 * Device or host pointers as for ?gemm_; the call returns when the output
 * is complete. */
void ctrsm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const void* alpha, const void* a, const int* lda, void* b, const int* ldb);
void ztrsm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const void* alpha, const void* a, const int* lda, void* b, const int* ldb);
void ctrmm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const void* alpha, const void* a, const int* lda, void* b, const int* ldb);
void ztrmm_(const char* side, const char* uplo, const char* transa, const char* diag, const int* m, const int* n,
            const void* alpha, const void* a, const int* lda, void* b, const int* ldb);
void csymm_(const char* side, const char* uplo, const int* m, const int* n, const void* alpha, const void* a,
            const int* lda, const void* b, const int* ldb, const void* beta, void* c, const int* ldc);
void zsymm_(const char* side, const char* uplo, const int* m, const int* n, const void* alpha, const void* a,
            const int* lda, const void* b, const int* ldb, const void* beta, void* c, const int* ldc);
void chemm_(const char* side, const char* uplo, const int* m, const int* n, const void* alpha, const void* a,
            const int* lda, const void* b, const int* ldb, const void* beta, void* c, const int* ldc);
void zhemm_(const char* side, const char* uplo, const int* m, const int* n, const void* alpha, const void* a,
            const int* lda, const void* b, const int* ldb, const void* beta, void* c, const int* ldc);
void csyrk_(const char* uplo, const char* trans, const int* n, const int* k, const void* alpha, const void* a,
            const int* lda, const void* beta, void* c, const int* ldc);
void zsyrk_(const char* uplo, const char* trans, const int* n, const int* k, const void* alpha, const void* a,
            const int* lda, const void* beta, void* c, const int* ldc);
void cherk_(const char* uplo, const char* trans, const int* n, const int* k, const float* alpha, const void* a,
            const int* lda, const float* beta, void* c, const int* ldc);
void zherk_(const char* uplo, const char* trans, const int* n, const int* k, const double* alpha, const void* a,
            const int* lda, const double* beta, void* c, const int* ldc);
void csyr2k_(const char* uplo, const char* trans, const int* n, const int* k, const void* alpha, const void* a,
             const int* lda, const void* b, const int* ldb, const void* beta, void* c, const int* ldc);
void zsyr2k_(const char* uplo, const char* trans, const int* n, const int* k, const void* alpha, const void* a,
             const int* lda, const void* b, const int* ldb, const void* beta, void* c, const int* ldc);
void cher2k_(const char* uplo, const char* trans, const int* n, const int* k, const void* alpha, const void* a,
             const int* lda, const void* b, const int* ldb, const float* beta, void* c, const int* ldc);
void zher2k_(const char* uplo, const char* trans, const int* n, const int* k, const void* alpha, const void* a,
             const int* lda, const void* b, const int* ldb, const double* beta, void* c, const int* ldc);
void cgemmtr_(const char* uplo, const char* transa, const char* transb, const int* n, const int* k,
              const void* alpha, const void* a, const int* lda, const void* b, const int* ldb, const void* beta,
              void* c, const int* ldc);
void zgemmtr_(const char* uplo, const char* transa, const char* transb, const int* n, const int* k,
              const void* alpha, const void* a, const int* lda, const void* b, const int* ldb, const void* beta,
              void* c, const int* ldc);
/* Complex matrix-vector product (the ?gemv_ contract on complex elements,
 * names "CGEMV " / "ZGEMV "); op 'C' conjugates. */
void cgemv_(const char* trans, const int* m, const int* n, const void* alpha, const void* a, const int* lda,
            const void* x, const int* incx, const void* beta, void* y, const int* incy);
void zgemv_(const char* trans, const int* m, const int* n, const void* alpha, const void* a, const int* lda,
            const void* x, const int* incx, const void* beta, void* y, const int* incy);
/* bfloat16 product with an FP32 result (the OpenBLAS sbgemm prototype that
 * Eigen declares and calls for bfloat16 products under
 * EIGEN_USE_OPENBLAS_BFLOAT16, GeneralMatrixMatrix_BLAS.h:127-131): a and b
 * hold bfloat16 values (16-bit patterns), alpha, beta and c are float.
 * Validation order and info codes as ?gemm_ (name "SBGEMM"); device or host
 * pointers as for ?gemm_. */
void sbgemm_(const char* transa, const char* transb, const int* m, const int* n, const int* k, const float* alpha,
             const void* a, const int* lda, const void* b, const int* ldb, const float* beta, float* c,
             const int* ldc);
/* Error hook (blas/xerbla.cpp:18): weak default prints the routine and info. */
void xerbla_(const char* name, int* info, size_t len);

/* Stream the BLAS entry points run on (a hipStream_t; NULL = default stream). */
void eh_blas_set_stream(void* stream);
void* eh_blas_get_stream(void);

#ifdef __cplusplus
}
#endif

#endif /* EIGEN_AMD_BLAS_H */
