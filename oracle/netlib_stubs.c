/* netlib_stubs.c -- TEST INFRASTRUCTURE ONLY.
 * The netlib level-3 testers (blas/testing/sblat3.f, dblat3.f) reference
 * every level-3 routine; the runs here switch on only the routines the
 * library provides (GEMM, SYRK), so these entry points must exist but are
 * never called. */
#include <stdio.h>
#include <stdlib.h>

#define NEVER(name)                                                     \
  void name(void) {                                                     \
    fprintf(stderr, "netlib tester called " #name " (switched off)\n"); \
    abort();                                                            \
  }
NEVER(ssymm_)
NEVER(ssyr2k_)
NEVER(strmm_)
NEVER(strsm_)
NEVER(dsymm_)
NEVER(dsyr2k_)
NEVER(dtrmm_)
NEVER(dtrsm_)
