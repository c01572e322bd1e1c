// Products on eigen_amd::DeviceMap, with no `.device()` at the call site:
//   C = A * B;  C.noalias() = A * B;  C += expr;  C -= expr;
// for matrix, vector, row-vector, tiny and row-major shapes, real and
// complex scalars; and complex products through the `.device()` seam.
// Each case runs the same expression on host matrices with Eigen's CPU path
// and compares bit for bit (small-integer operands: every partial sum is
// exact).
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

// Integers in [-8, 8] in the real and (for complex scalars) imaginary parts.
template <typename Mat>
void fill_exact(Mat& m, unsigned seed) {
  using T = typename Mat::Scalar;
  unsigned s = seed * 2654435761u + 777u;
  auto next = [&]() {
    s = s * 1664525u + 1013904223u;
    return int((s >> 16) % 17) - 8;
  };
  for (Index j = 0; j < m.cols(); ++j)
    for (Index i = 0; i < m.rows(); ++i) {
      if constexpr (is_cplx<T>::value) {
        const int re = next(), im = next();
        m(i, j) = T(re, im);
      } else {
        m(i, j) = T(next());
      }
    }
}

// A host matrix and its copy in device memory (leading dimension ld, padding
// filled with -99), reachable as a DeviceMap.
template <typename Plain>
struct Dev {
  using T = typename Plain::Scalar;
  static constexpr bool kRow = Plain::IsRowMajor;
  Plain host;
  Index ld;
  DeviceBuffer<T> buf;
  Dev(Index r, Index c, unsigned seed, Index pad = 0)
      : host(r, c), ld((kRow ? c : r) + pad), buf((size_t)ld * (kRow ? r : c)) {
    fill_exact(host, seed);
    std::vector<T> h(buf.n, T(-99));
    for (Index j = 0; j < c; ++j)
      for (Index i = 0; i < r; ++i) h[kRow ? j + i * ld : i + j * ld] = host(i, j);
    buf.upload(h.data());
  }
  DeviceMap<Plain> map() { return DeviceMap<Plain>(buf.ptr, host.rows(), host.cols(), ld); }
  Plain fetch() const {
    std::vector<T> h(buf.n);
    buf.download(h.data());
    Plain out(host.rows(), host.cols());
    for (Index j = 0; j < host.cols(); ++j)
      for (Index i = 0; i < host.rows(); ++i) out(i, j) = h[kRow ? j + i * ld : i + j * ld];
    return out;
  }
};

template <typename T>
int run_all(const char* tname) {
  using Mat = Eigen::Matrix<T, Dynamic, Dynamic>;
  using RowMat = Eigen::Matrix<T, Dynamic, Dynamic, Eigen::RowMajor>;
  using Vec = Eigen::Matrix<T, Dynamic, 1>;
  using RowVec = Eigen::Matrix<T, 1, Dynamic>;
  int fails = 0;
  char name[128];
  auto tag = [&](const char* what) {
    std::snprintf(name, sizeof name, "%s/%s", tname, what);
    return name;
  };
  eigen_amd::HipDevice& dev = eigen_amd::default_device();
  const Index m = 70, n = 45, k = 33;
  {  // C = A * B: no host temporary, no host setZero
    Dev<Mat> A(m, k, 1), B(k, n, 2), C(m, n, 3, 4);
    auto c = C.map();
    c = A.map() * B.map();
    Mat want = A.host * B.host;
    fails += report(tag("assign"), C.fetch() == want && dev.status == EH_SUCCESS);
  }
  {  // C.noalias() = A * B
    Dev<Mat> A(m, k, 4), B(k, n, 5), C(m, n, 6);
    auto c = C.map();
    c.noalias() = A.map() * B.map();
    Mat want = A.host * B.host;
    fails += report(tag("noalias_assign"), C.fetch() == want);
  }
  {  // C += 2 * A^T * B, A stored k x m with padding
    Dev<Mat> A(k, m, 7, 3), B(k, n, 8), C(m, n, 9);
    auto c = C.map();
    c += T(2) * A.map().transpose() * B.map();
    Mat want = C.host + T(2) * A.host.transpose() * B.host;
    fails += report(tag("add_scaled_transposed"), C.fetch() == want);
  }
  {  // C.noalias() -= A * B^T
    Dev<Mat> A(m, k, 10), B(n, k, 11), C(m, n, 12);
    auto c = C.map();
    c.noalias() -= A.map() * B.map().transpose();
    Mat want = C.host - A.host * B.host.transpose();
    fails += report(tag("noalias_sub"), C.fetch() == want);
  }
  {  // y = A * x (matrix-vector shape)
    Dev<Mat> A(m, k, 13);
    Dev<Vec> x(k, 1, 14), y(m, 1, 15);
    DeviceMap<Vec> yv(y.buf.ptr, m), xv(x.buf.ptr, k);
    yv = A.map() * xv;
    Vec want = A.host * x.host;
    fails += report(tag("gemv"), y.fetch() == want);
    yv -= A.map() * xv;  // back to zero
    fails += report(tag("gemv_sub"), y.fetch() == Vec::Zero(m));
  }
  {  // r = x^T * A (row-vector result)
    Dev<Mat> A(k, n, 16);
    Dev<RowVec> x(1, k, 17), r(1, n, 18);
    DeviceMap<RowVec> rv(r.buf.ptr, n), xv(x.buf.ptr, k);
    rv = xv * A.map();
    RowVec want = x.host * A.host;
    fails += report(tag("row_vector"), r.fetch() == want);
  }
  {  // tiny: 2x3 times 3x1, and 1x1 results (an inner product)
    Dev<Mat> a(2, 3, 19), b(3, 1, 20), c(2, 1, 21);
    auto cm = c.map();
    cm = a.map() * b.map();
    Mat want = a.host * b.host;
    fails += report(tag("tiny_2x3x1"), c.fetch() == want);
    Dev<Mat> u(5, 1, 22), v(5, 1, 23), s(1, 1, 24);
    auto sm = s.map();
    sm = u.map().transpose() * v.map();
    Mat want2 = u.host.transpose() * v.host;
    fails += report(tag("tiny_1x1"), s.fetch() == want2);
  }
  {  // a product into a sub-block of a DeviceMap (column- and row-major)
    Dev<Mat> A(m, k, 37), B(k, n, 38), C(m, n, 39, 2);
    auto c = C.map();
    c.block(3, 4, 40, 30) = A.map().block(5, 2, 40, 20) * B.map().block(1, 6, 20, 30);
    Mat want = C.host;
    want.block(3, 4, 40, 30) = A.host.block(5, 2, 40, 20) * B.host.block(1, 6, 20, 30);
    fails += report(tag("block_assign"), C.fetch() == want);
    Dev<RowMat> R(m, n, 40, 3);
    auto r = R.map();
    r.block(7, 1, 20, 40) -= A.map().block(0, 0, 20, 10) * B.map().block(0, 0, 10, 40);
    RowMat want2 = R.host;
    want2.block(7, 1, 20, 40) -= A.host.block(0, 0, 20, 10) * B.host.block(0, 0, 10, 40);
    fails += report(tag("block_sub_row_major"), R.fetch() == want2);
  }
  {  // row-major destination and operand
    Dev<RowMat> A(m, k, 25), C(m, n, 27, 2);
    Dev<Mat> B(k, n, 26);
    auto c = C.map();
    c += A.map() * B.map();
    RowMat want = C.host + A.host * B.host;
    fails += report(tag("row_major"), C.fetch() == want);
  }
  if constexpr (is_cplx<T>::value) {
    {  // C = A^H * B (conjugate transpose on the device)
      Dev<Mat> A(k, m, 28), B(k, n, 29), C(m, n, 30);
      auto c = C.map();
      c = A.map().adjoint() * B.map();
      Mat want = A.host.adjoint() * B.host;
      fails += report(tag("adjoint"), C.fetch() == want);
    }
    {  // C -= alpha * A * B^H with a complex alpha
      Dev<Mat> A(m, k, 31), B(n, k, 32), C(m, n, 33);
      const T alpha(2, -1);
      auto c = C.map();
      c.noalias() -= alpha * A.map() * B.map().adjoint();
      Mat want = C.host - alpha * A.host * B.host.adjoint();
      fails += report(tag("sub_alpha_adjoint"), C.fetch() == want);
    }
    {  // through the .device() seam, on plain Maps, row-major conjugated operand
      Dev<RowMat> A(m, k, 34);
      Dev<Mat> B(k, n, 35), C(m, n, 36);
      eigen_amd::HipDevice d2;
      Eigen::Map<Mat, 0, Eigen::OuterStride<>> c(C.buf.ptr, m, n, Eigen::OuterStride<>(C.ld));
      Eigen::Map<RowMat, 0, Eigen::OuterStride<>> a(A.buf.ptr, m, k, Eigen::OuterStride<>(A.ld));
      Eigen::Map<Mat, 0, Eigen::OuterStride<>> b(B.buf.ptr, k, n, Eigen::OuterStride<>(B.ld));
      c.device(d2).noalias() += a.conjugate() * b;  // row-major conj(A): op C of its storage
      Mat want = C.host + A.host.conjugate() * B.host;
      fails += report(tag("seam_conj_row_major"), C.fetch() == want && d2.status == EH_SUCCESS);
    }
  }
  return fails;
}

int main() {
  int fails = run_all<float>("float") + run_all<double>("double") + run_all<std::complex<float>>("cfloat") +
              run_all<std::complex<double>>("cdouble");
  std::printf("%s (%d failures)\n", fails ? "SOME FAILED" : "ALL PASSED", fails);
  return fails ? 1 : 0;
}
