/* blas_ref_hooks.c -- TEST INFRASTRUCTURE ONLY.
 * A strong xerbla_ for the reference BLAS built into oracle/_ref: Eigen's own
 * xerbla_ is weak (blas/xerbla.cpp:9,18) precisely so that testers can
 * override it; this one records the info code and routine name instead of
 * printing, so tests can compare error exits code for code. */
#include <stddef.h>
#include <string.h>

static int g_last_info = 0;
static char g_last_name[8];

__attribute__((visibility("default"))) void xerbla_(const char* name, int* info, size_t len) {
  g_last_info = *info;
  size_t n = len < 7 ? len : 7;
  memcpy(g_last_name, name, n);
  g_last_name[n] = 0;
}

__attribute__((visibility("default"))) int ref_blas_last_info(void) { return g_last_info; }
__attribute__((visibility("default"))) const char* ref_blas_last_name(void) { return g_last_name; }
__attribute__((visibility("default"))) void ref_blas_reset(void) {
  g_last_info = 0;
  g_last_name[0] = 0;
}
