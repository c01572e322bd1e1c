/* netlib_stubs.c -- TEST INFRASTRUCTURE ONLY.
 * The netlib level-3 testers (blas/testing/sblat3.f, dblat3.f) link against
 * every real level-3 routine, and the library now provides all of them
 * (GEMM, SYMM, TRMM, TRSM, SYRK, SYR2K), so nothing is stubbed.  The object
 * is kept so the link recipe stays the same. */
void eigen_amd_netlib_stubs_unused(void) {}
