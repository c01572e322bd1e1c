// Level-3 views on eigen_amd::DeviceMap, against Eigen's CPU evaluation of
// the same expressions on host copies:
//   A.triangularView<M>().solveInPlace(X), solveInPlace<OnTheRight>(X)  (trsm)
//   X = A.triangularView<M>() * X, X = X * A.triangularView<M>()         (trmm)
//   C.triangularView<M>() = / += / -= A * B                              (gemmtr)
//   C = / += / -= A.selfadjointView<M>() * B, B * A.selfadjointView<M>() (symm / hemm)
//   C.selfadjointView<M>().rankUpdate(U, a), rankUpdate(U, V, a)         (syrk / herk, syr2k / her2k)
// with transposed and adjoint views, scaled views, unit diagonals and
// row-major storage, for float, double, complex<float>, complex<double>.
// Operands are small integers; the triangles of the solves are unit-banded
// (one off-diagonal of entries in {-1, 0, 1}, so their inverses have entries
// in {-1, 0, 1}), so every result is exact and compared bit for bit.  The
// triangles not referenced hold values that would change the result if read.
#include <complex>
#include <type_traits>

#include "../../include/eigen_amd_eigen.h"
#include "common.hpp"

using Eigen::Dynamic;
using Eigen::Index;
using eigen_amd::DeviceMap;

template <typename T>
struct is_cplx : std::false_type {};
template <typename R>
struct is_cplx<std::complex<R>> : std::true_type {};

template <typename T>
T make(int re, int im) {
  if constexpr (is_cplx<T>::value)
    return T(re, im);
  else
    return T(re);
}

struct Rng {
  unsigned s;
  explicit Rng(unsigned seed) : s(seed * 2654435761u + 99u) {}
  int next(int lo, int hi) {
    s = s * 1664525u + 1013904223u;
    return lo + int((s >> 16) % unsigned(hi - lo + 1));
  }
};

// integers in [-8, 8] (real and imaginary parts)
template <typename Mat>
void fill(Mat& m, unsigned seed) {
  Rng r(seed);
  for (Index j = 0; j < m.cols(); ++j)
    for (Index i = 0; i < m.rows(); ++i) {
      const int re = r.next(-8, 8), im = r.next(-8, 8);
      m(i, j) = make<typename Mat::Scalar>(re, im);
    }
}

// A banded triangle for solves: diagonal +-1 (or, for a unit diagonal, an
// unreferenced 7), one off-diagonal of entries in {-1, 0, 1} on the `lower`
// side, zeros elsewhere in that triangle and 9s in the other (unreferenced)
// one.
template <typename Mat>
void fill_banded(Mat& m, bool lower, bool unit, unsigned seed) {
  using T = typename Mat::Scalar;
  Rng r(seed);
  for (Index j = 0; j < m.cols(); ++j)
    for (Index i = 0; i < m.rows(); ++i) {
      if (i == j)
        m(i, j) = unit ? make<T>(7, 3) : make<T>(r.next(0, 1) ? 1 : -1, 0);
      else if ((i > j) == lower)
        m(i, j) = (lower ? i == j + 1 : j == i + 1) ? make<T>(r.next(-1, 1), 0) : T(0);
      else
        m(i, j) = make<T>(9, -9);
    }
}

// A host matrix and its device copy (leading dimension with padding).
template <typename Plain>
struct Dev {
  using T = typename Plain::Scalar;
  static constexpr bool kRow = Plain::IsRowMajor;
  Plain host;
  Index ld;
  DeviceBuffer<T> buf;
  explicit Dev(const Plain& h, Index pad = 1)
      : host(h), ld((kRow ? h.cols() : h.rows()) + pad), buf((size_t)ld * (kRow ? h.rows() : h.cols())) {
    std::vector<T> v(buf.n, make<T>(-99, 1));
    for (Index j = 0; j < h.cols(); ++j)
      for (Index i = 0; i < h.rows(); ++i) v[kRow ? j + i * ld : i + j * ld] = h(i, j);
    buf.upload(v.data());
  }
  DeviceMap<Plain> map() { return DeviceMap<Plain>(buf.ptr, host.rows(), host.cols(), ld); }
  Plain fetch() const {
    std::vector<T> v(buf.n);
    buf.download(v.data());
    Plain out(host.rows(), host.cols());
    for (Index j = 0; j < host.cols(); ++j)
      for (Index i = 0; i < host.rows(); ++i) out(i, j) = v[kRow ? j + i * ld : i + j * ld];
    return out;
  }
};

template <typename Plain>
Plain random_mat(Index r, Index c, unsigned seed) {
  Plain m(r, c);
  fill(m, seed);
  return m;
}

// Hermitian / symmetric test matrix data: real diagonal (the imaginary part
// of a Hermitian diagonal is not referenced by the BLAS routines).
template <typename Plain>
Plain sa_mat(Index n, unsigned seed) {
  Plain m = random_mat<Plain>(n, n, seed);
  for (Index i = 0; i < n; ++i) m(i, i) = make<typename Plain::Scalar>(int(Eigen::numext::real(m(i, i))), 0);
  return m;
}

template <typename T, int Order>
int run_all(const char* tname) {
  using Mat = Eigen::Matrix<T, Dynamic, Dynamic, Order>;
  using CMat = Eigen::Matrix<T, Dynamic, Dynamic>;              // column-major
  using RMat = Eigen::Matrix<T, Dynamic, Dynamic, Eigen::RowMajor>;
  int fails = 0;
  char name[160];
  auto tag = [&](const char* what) {
    std::snprintf(name, sizeof name, "%s/%s/%s", tname, Order == Eigen::RowMajor ? "rowmajor" : "colmajor", what);
    return name;
  };
  eigen_amd::HipDevice& dev = eigen_amd::default_device();
  const Index n = 37, m = 29, k = 23;
  constexpr bool cplx = is_cplx<T>::value;

  // ---- solves (trsm) ----
  {
    Mat a(n, n);
    fill_banded(a, true, true, 1);
    Dev<Mat> A(a);
    Dev<Mat> X(random_mat<Mat>(n, m, 2));
    A.map().template triangularView<Eigen::UnitLower>().solveInPlace(X.map());
    Mat want = X.host;
    A.host.template triangularView<Eigen::UnitLower>().solveInPlace(want);
    fails += report(tag("solve_unitlower_left"), X.fetch() == want && dev.status == EH_SUCCESS);
  }
  {
    Mat a(n, n);
    fill_banded(a, false, false, 3);
    Dev<Mat> A(a);
    Dev<Mat> X(random_mat<Mat>(m, n, 4));
    A.map().template triangularView<Eigen::Upper>().template solveInPlace<Eigen::OnTheRight>(X.map());
    Mat want = X.host;
    A.host.template triangularView<Eigen::Upper>().template solveInPlace<Eigen::OnTheRight>(want);
    fails += report(tag("solve_upper_right"), X.fetch() == want);
  }
  {  // transposed view, left; a scaled view divides
    Mat a(n, n);
    fill_banded(a, true, false, 5);
    Dev<Mat> A(a);
    Dev<Mat> X(random_mat<Mat>(n, m, 6));
    (T(2) * A.map().template triangularView<Eigen::Lower>()).transpose().solveInPlace(X.map());
    Mat want = X.host;
    A.host.template triangularView<Eigen::Lower>().transpose().solveInPlace(want);
    want /= T(2);
    fails += report(tag("solve_scaled_transposed"), X.fetch() == want);
  }
  if constexpr (cplx) {  // adjoint view (column-major X only: a row-major X would need conj without transpose)
    CMat a(n, n);
    fill_banded(a, false, true, 7);
    Dev<CMat> A(a);
    Dev<CMat> X(random_mat<CMat>(n, m, 8));
    A.map().template triangularView<Eigen::UnitUpper>().adjoint().solveInPlace(X.map());
    CMat want = X.host;
    A.host.template triangularView<Eigen::UnitUpper>().adjoint().solveInPlace(want);
    fails += report(tag("solve_adjoint"), X.fetch() == want);
  }

  // ---- triangular products in place (trmm) ----
  {
    Dev<Mat> A(random_mat<Mat>(n, n, 11));
    Dev<Mat> X(random_mat<Mat>(n, m, 12));
    auto x = X.map();
    x = A.map().template triangularView<Eigen::Lower>() * x;
    Mat want = A.host.template triangularView<Eigen::Lower>() * X.host;
    fails += report(tag("trmm_lower_left"), X.fetch() == want);
  }
  {
    Dev<Mat> A(random_mat<Mat>(n, n, 13));
    Dev<Mat> X(random_mat<Mat>(m, n, 14));
    auto x = X.map();
    x = x * (T(3) * A.map().template triangularView<Eigen::UnitUpper>()).transpose();
    Mat want = T(3) * (X.host * A.host.template triangularView<Eigen::UnitUpper>().transpose());
    fails += report(tag("trmm_scaled_unitupper_transposed_right"), X.fetch() == want);
  }
  if constexpr (cplx) {
    CMat a = random_mat<CMat>(n, n, 15);
    Dev<CMat> A(a);
    Dev<CMat> X(random_mat<CMat>(n, m, 16));
    auto x = X.map();
    x = A.map().template triangularView<Eigen::Upper>().adjoint() * x;
    CMat want = A.host.template triangularView<Eigen::Upper>().adjoint() * X.host;
    fails += report(tag("trmm_adjoint_left"), X.fetch() == want);
  }

  // ---- product into a triangle (gemmtr) ----
  {
    Dev<Mat> A(random_mat<Mat>(n, k, 21)), B(random_mat<Mat>(k, n, 22)), C(random_mat<Mat>(n, n, 23));
    C.map().template triangularView<Eigen::Lower>() = A.map() * B.map();
    Mat want = C.host;
    want.template triangularView<Eigen::Lower>() = A.host * B.host;
    fails += report(tag("gemmtr_lower_assign"), C.fetch() == want);
  }
  {
    Dev<Mat> A(random_mat<Mat>(k, n, 24)), B(random_mat<Mat>(n, k, 25)), C(random_mat<Mat>(n, n, 26));
    C.map().template triangularView<Eigen::Upper>() -= T(2) * A.map().transpose() * B.map().transpose();
    Mat want = C.host;
    want.template triangularView<Eigen::Upper>() -= T(2) * A.host.transpose() * B.host.transpose();
    fails += report(tag("gemmtr_upper_sub_transposed"), C.fetch() == want);
  }

  // ---- self-adjoint products (symm / hemm): A, B, C in one storage order ----
  {
    Dev<Mat> A(sa_mat<Mat>(n, 31)), B(random_mat<Mat>(n, m, 32)), C(random_mat<Mat>(n, m, 33));
    auto c = C.map();
    c = A.map().template selfadjointView<Eigen::Lower>() * B.map();
    Mat want = A.host.template selfadjointView<Eigen::Lower>() * B.host;
    fails += report(tag("symm_lower_left_assign"), C.fetch() == want);
    c += T(2) * A.map().template selfadjointView<Eigen::Lower>() * B.map();
    want += T(2) * (A.host.template selfadjointView<Eigen::Lower>() * B.host);
    fails += report(tag("symm_scaled_left_add"), C.fetch() == want);
  }
  {
    Dev<Mat> A(sa_mat<Mat>(n, 34)), B(random_mat<Mat>(m, n, 35)), C(random_mat<Mat>(m, n, 36));
    auto c = C.map();
    c -= B.map() * A.map().template selfadjointView<Eigen::Upper>();
    Mat want = C.host - B.host * A.host.template selfadjointView<Eigen::Upper>();
    fails += report(tag("symm_upper_right_sub"), C.fetch() == want);
  }

  // ---- rank updates (syrk / herk, syr2k / her2k) ----
  if constexpr (!cplx || Order == Eigen::ColMajor) {
    using RealT = typename Eigen::NumTraits<T>::Real;
    {
      Dev<Mat> C(sa_mat<Mat>(n, 41));
      Dev<CMat> U(random_mat<CMat>(n, k, 42));
      C.map().template selfadjointView<Eigen::Lower>().rankUpdate(U.map(), RealT(2));
      Mat want = C.host;
      want.template selfadjointView<Eigen::Lower>().rankUpdate(U.host, RealT(2));
      fails += report(tag("rank_update_lower"), C.fetch() == want);
    }
    {
      Dev<Mat> C(sa_mat<Mat>(n, 43));
      Dev<CMat> U(random_mat<CMat>(k, n, 44));
      if constexpr (cplx) {
        C.map().template selfadjointView<Eigen::Upper>().rankUpdate(U.map().adjoint(), RealT(-1));
        Mat want = C.host;
        want.template selfadjointView<Eigen::Upper>().rankUpdate(U.host.adjoint(), RealT(-1));
        fails += report(tag("rank_update_upper_adjoint"), C.fetch() == want);
      } else {
        C.map().template selfadjointView<Eigen::Upper>().rankUpdate(U.map().transpose(), RealT(-1));
        Mat want = C.host;
        want.template selfadjointView<Eigen::Upper>().rankUpdate(U.host.transpose(), RealT(-1));
        fails += report(tag("rank_update_upper_transposed"), C.fetch() == want);
      }
    }
    {
      Dev<Mat> C(sa_mat<Mat>(n, 45));
      Dev<CMat> U(random_mat<CMat>(n, k, 46)), V(random_mat<CMat>(n, k, 47));
      const T alpha = make<T>(2, -1);
      C.map().template selfadjointView<Eigen::Lower>().rankUpdate(U.map(), V.map(), alpha);
      // Eigen's own rankUpdate(u, v) takes vectors; for n x k factors the
      // update is the rank-2k sum, written out
      Mat want = C.host;
      want.template triangularView<Eigen::Lower>() +=
          (alpha * U.host * V.host.adjoint() + Eigen::numext::conj(alpha) * V.host * U.host.adjoint()).eval();
      fails += report(tag("rank2k_update_lower"), C.fetch() == want);
      // and with vectors, against Eigen's rank-2 update itself
      Dev<Mat> C2(sa_mat<Mat>(n, 48));
      Dev<CMat> u(random_mat<CMat>(n, 1, 49)), v(random_mat<CMat>(n, 1, 50));
      C2.map().template selfadjointView<Eigen::Upper>().rankUpdate(u.map(), v.map(), alpha);
      Mat want2 = C2.host;
      want2.template selfadjointView<Eigen::Upper>().rankUpdate(u.host.col(0), v.host.col(0), alpha);
      fails += report(tag("rank2_update_upper_vectors"), C2.fetch() == want2);
    }
  }
  return fails;
}

int main() {
  int fails = run_all<float, Eigen::ColMajor>("float") + run_all<double, Eigen::ColMajor>("double") +
              run_all<float, Eigen::RowMajor>("float") + run_all<double, Eigen::RowMajor>("double") +
              run_all<std::complex<float>, Eigen::ColMajor>("cfloat") +
              run_all<std::complex<double>, Eigen::ColMajor>("cdouble") +
              run_all<std::complex<double>, Eigen::RowMajor>("cdouble");
  std::printf("%s (%d failures)\n", fails ? "SOME FAILED" : "ALL PASSED", fails);
  return fails ? 1 : 0;
}
