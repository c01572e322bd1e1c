// Eigen's BLAS redirects under EIGEN_USE_BLAS for the products beyond GEMM,
// pointed at libeigen_amd_blas, on unchanged Eigen expressions over host Maps:
//   y.noalias() += a*A*x                          -> dgemv_  (GeneralMatrixVector_BLAS.h)
//   A.triangularView<UnitLower>().solveInPlace(B) -> dtrsm_  (TriangularSolverMatrix_BLAS.h)
//   A.triangularView<UnitUpper>().solveInPlace<OnTheRight>(B)
//   C.noalias() = A.triangularView<Lower>() * B   -> dtrmm_  (TriangularMatrixMatrix_BLAS.h)
//   C.noalias() += A.selfadjointView<Lower>() * B -> dsymm_  (SelfadjointMatrixMatrix_BLAS.h)
//   C.selfadjointView<Lower>().rankUpdate(A, a)   -> dsyrk_  (GeneralMatrixMatrixTriangular_BLAS.h)
//   complex<double> C.noalias() += A*B            -> zgemm_  (GeneralMatrixMatrix_BLAS.h)
// Each BLAS symbol is interposed here to count its calls before forwarding to
// the library (dlsym(RTLD_NEXT)), so the test proves the route was taken;
// every result is checked bit for bit against an exact integer loop (the
// triangles are unit and banded with entries in {-1, 0, 1}, so solutions stay
// small integers).
#include <dlfcn.h>

#include <Eigen/Core>
#include <complex>
#include <cstdint>

#include "common.hpp"

using Eigen::Dynamic;
using Eigen::Index;
using Mat = Eigen::Matrix<double, Dynamic, Dynamic>;
using Vec = Eigen::Matrix<double, Dynamic, 1>;
using CMat = Eigen::Matrix<std::complex<double>, Dynamic, Dynamic>;

enum { GEMV, TRSM, TRMM, SYMM, SYRK, ZGEMM, NCALLS };
static int g_calls[NCALLS];

template <typename F>
static F next_symbol(const char* name) {
  void* p = dlsym(RTLD_NEXT, name);
  if (!p) {
    std::fprintf(stderr, "no %s after this program\n", name);
    std::exit(2);
  }
  return reinterpret_cast<F>(p);
}

extern "C" {
void dgemv_(const char* t, const int* m, const int* n, const double* al, const double* a, const int* lda,
            const double* x, const int* ix, const double* be, double* y, const int* iy) {
  ++g_calls[GEMV];
  static auto f = next_symbol<decltype(&dgemv_)>("dgemv_");
  f(t, m, n, al, a, lda, x, ix, be, y, iy);
}
void dtrsm_(const char* s, const char* u, const char* t, const char* d, const int* m, const int* n, const double* al,
            const double* a, const int* lda, double* b, const int* ldb) {
  ++g_calls[TRSM];
  static auto f = next_symbol<decltype(&dtrsm_)>("dtrsm_");
  f(s, u, t, d, m, n, al, a, lda, b, ldb);
}
void dtrmm_(const char* s, const char* u, const char* t, const char* d, const int* m, const int* n, const double* al,
            const double* a, const int* lda, double* b, const int* ldb) {
  ++g_calls[TRMM];
  static auto f = next_symbol<decltype(&dtrmm_)>("dtrmm_");
  f(s, u, t, d, m, n, al, a, lda, b, ldb);
}
void dsymm_(const char* s, const char* u, const int* m, const int* n, const double* al, const double* a,
            const int* lda, const double* b, const int* ldb, const double* be, double* c, const int* ldc) {
  ++g_calls[SYMM];
  static auto f = next_symbol<decltype(&dsymm_)>("dsymm_");
  f(s, u, m, n, al, a, lda, b, ldb, be, c, ldc);
}
void dsyrk_(const char* u, const char* t, const int* n, const int* k, const double* al, const double* a,
            const int* lda, const double* be, double* c, const int* ldc) {
  ++g_calls[SYRK];
  static auto f = next_symbol<decltype(&dsyrk_)>("dsyrk_");
  f(u, t, n, k, al, a, lda, be, c, ldc);
}
// (the prototype Eigen declares, Eigen/src/misc/blas.h:70-72: complex as pairs of doubles)
void zgemm_(const char* ta, const char* tb, const int* m, const int* n, const int* k, const double* al,
            const double* a, const int* lda, const double* b, const int* ldb, const double* be, double* c,
            const int* ldc) {
  ++g_calls[ZGEMM];
  static auto f = next_symbol<decltype(&zgemm_)>("zgemm_");
  f(ta, tb, m, n, k, al, a, lda, b, ldb, be, c, ldc);
}
}

// Unit triangle (lower or upper) with a band of 3 off-diagonals in {-1, 0, 1};
// the other triangle holds junk the routines must not read.
static Mat unit_band(Index n, bool lower, unsigned seed) {
  Mat A(n, n);
  fill_int(A, seed);
  for (Index j = 0; j < n; ++j)
    for (Index i = 0; i < n; ++i) {
      const Index d = lower ? i - j : j - i;
      if (d == 0)
        A(i, j) = 1;
      else if (d > 0 && d <= 3)
        A(i, j) = double(((i * 7 + j * 3 + seed) % 3)) - 1.0;
      else if (d > 3)
        A(i, j) = 0;
      else
        A(i, j) = 1e300;  // other triangle: never read
    }
  return A;
}

static Mat tri_of(const Mat& A, bool lower, bool unit) {
  Mat T = Mat::Zero(A.rows(), A.cols());
  for (Index j = 0; j < A.cols(); ++j)
    for (Index i = 0; i < A.rows(); ++i)
      if (i == j)
        T(i, j) = unit ? 1.0 : A(i, j);
      else if (lower ? i > j : i < j)
        T(i, j) = A(i, j);
  return T;
}

static Mat loop_mul(const Mat& A, const Mat& B) {
  Mat C = Mat::Zero(A.rows(), B.cols());
  for (Index j = 0; j < B.cols(); ++j)
    for (Index p = 0; p < A.cols(); ++p)
      for (Index i = 0; i < A.rows(); ++i) C(i, j) += A(i, p) * B(p, j);
  return C;
}

int main() {
  int fails = 0;
  // GEMV: y += 2*A*x and y = A^T x
  {
    Mat A(70, 45);
    Vec x(45), y0(70), x2(70);
    fill_int(A, 11);
    fill_int(x, 12);
    fill_int(y0, 13);
    fill_int(x2, 14);
    Vec y = y0;
    y.noalias() += 2.0 * A * x;
    Vec want = y0 + 2.0 * loop_mul(A, x);
    fails += report("gemv/N", y == want);
    Vec z(45);
    z.noalias() = A.transpose() * x2;
    fails += report("gemv/T", z == Vec(loop_mul(A.transpose(), x2)));
  }
  // TRSM left (unit lower) and right (unit upper)
  {
    const Index n = 90, r = 37;
    Mat L = unit_band(n, true, 3), B(n, r);
    fill_int(B, 21);
    Mat X = B;
    L.triangularView<Eigen::UnitLower>().solveInPlace(X);
    fails += report("trsm/left_unit_lower", loop_mul(tri_of(L, true, true), X) == B);
    Mat U = unit_band(n, false, 5), B2(r, n);
    fill_int(B2, 22);
    Mat X2 = B2;
    U.triangularView<Eigen::UnitUpper>().solveInPlace<Eigen::OnTheRight>(X2);
    fails += report("trsm/right_unit_upper", loop_mul(X2, tri_of(U, false, true)) == B2);
  }
  // TRMM: C = tril(A) * B
  {
    const Index n = 80, r = 30;
    Mat A(n, n), B(n, r);
    fill_int(A, 31);
    fill_int(B, 32);
    Mat C(n, r);
    C.noalias() = A.triangularView<Eigen::Lower>() * B;
    fails += report("trmm/lower", C == loop_mul(tri_of(A, true, false), B));
  }
  // SYMM: C += sym(lower of A) * B
  {
    const Index n = 60, r = 25;
    Mat A(n, n), B(n, r), C0(n, r);
    fill_int(A, 41);
    fill_int(B, 42);
    fill_int(C0, 43);
    Mat S = tri_of(A, true, false);
    S += tri_of(A, true, false).transpose();
    S.diagonal() = A.diagonal();
    Mat C = C0;
    C.noalias() += A.selfadjointView<Eigen::Lower>() * B;
    fails += report("symm/lower", C == C0 + loop_mul(S, B));
  }
  // SYRK: lower of C += 2 A A^T, upper untouched
  {
    const Index n = 50, k = 35;
    Mat A(n, k), C0(n, n);
    fill_int(A, 51);
    fill_int(C0, 52);
    Mat C = C0;
    C.selfadjointView<Eigen::Lower>().rankUpdate(A, 2.0);
    Mat full = C0 + 2.0 * loop_mul(A, A.transpose());
    bool ok = true;
    for (Index j = 0; j < n; ++j)
      for (Index i = 0; i < n; ++i) ok = ok && C(i, j) == (i >= j ? full(i, j) : C0(i, j));
    fails += report("syrk/lower", ok);
  }
  // complex GEMM
  {
    const Index m = 40, n = 30, k = 25;
    Mat Ar(m, k), Ai(m, k), Br(k, n), Bi(k, n);
    fill_int(Ar, 61);
    fill_int(Ai, 62);
    fill_int(Br, 63);
    fill_int(Bi, 64);
    CMat A(m, k), B(k, n), C = CMat::Zero(m, n);
    for (Index j = 0; j < k; ++j)
      for (Index i = 0; i < m; ++i) A(i, j) = {Ar(i, j), Ai(i, j)};
    for (Index j = 0; j < n; ++j)
      for (Index i = 0; i < k; ++i) B(i, j) = {Br(i, j), Bi(i, j)};
    C.noalias() += A * B;
    const Mat Re = loop_mul(Ar, Br) - loop_mul(Ai, Bi), Im = loop_mul(Ar, Bi) + loop_mul(Ai, Br);
    bool ok = true;
    for (Index j = 0; j < n; ++j)
      for (Index i = 0; i < m; ++i) ok = ok && C(i, j) == std::complex<double>(Re(i, j), Im(i, j));
    fails += report("zgemm", ok);
  }
  static const char* names[NCALLS] = {"dgemv_", "dtrsm_", "dtrmm_", "dsymm_", "dsyrk_", "zgemm_"};
  for (int c = 0; c < NCALLS; ++c) {
    char name[64];
    std::snprintf(name, sizeof name, "reached %s (%d calls)", names[c], g_calls[c]);
    fails += report(name, g_calls[c] > 0);
  }
  std::printf("%s (%d failures)\n", fails ? "SOME FAILED" : "ALL PASSED", fails);
  return fails ? 1 : 0;
}
