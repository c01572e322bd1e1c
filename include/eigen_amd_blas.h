/*
 * eigen_amd_blas.h -- BLAS level-3 GEMM entry points of the MI355X library.
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
/* Error hook (blas/xerbla.cpp:18): weak default prints the routine and info. */
void xerbla_(const char* name, int* info, size_t len);

/* Stream the BLAS entry points run on (a hipStream_t; NULL = default stream). */
void eh_blas_set_stream(void* stream);
void* eh_blas_get_stream(void);

#ifdef __cplusplus
}
#endif

#endif /* EIGEN_AMD_BLAS_H */
