/* netlib_stubs.c -- TEST INFRASTRUCTURE ONLY.
 * The netlib level-3 testers link against every level-3 routine of their
 * precision.  The library provides all six real routines, and of the
 * complex ones (blas/testing/cblat3.f, zblat3.f) only CGEMM / ZGEMM; the
 * complex decks switch the others off, so these entry points must exist but
 * are never called. */
#include <stdio.h>
#include <stdlib.h>

#define NEVER(name)                                                     \
  void name(void) {                                                     \
    fprintf(stderr, "netlib tester called " #name " (switched off)\n"); \
    abort();                                                            \
  }
NEVER(chemm_)
NEVER(csymm_)
NEVER(ctrmm_)
NEVER(ctrsm_)
NEVER(cherk_)
NEVER(csyrk_)
NEVER(cher2k_)
NEVER(csyr2k_)
NEVER(zhemm_)
NEVER(zsymm_)
NEVER(ztrmm_)
NEVER(ztrsm_)
NEVER(zherk_)
NEVER(zsyrk_)
NEVER(zher2k_)
NEVER(zsyr2k_)
