// Eigen's own BLAS redirect (EIGEN_USE_BLAS, GeneralMatrixMatrix_BLAS.h:53-117)
// pointed at libeigen_amd_blas: an unchanged `C.noalias() += alpha*A*B`
// becomes a call of our sgemm_/dgemm_.  With host Maps the operands are
// staged through device buffers; with Maps over device memory the product
// runs in place on the GPU.  (Shapes are at least 20 on every side and not
// vectors, so Eigen's host-side coefficient and GEMV fallbacks are not
// taken, GeneralMatrixMatrix.h:445-447, 480-490.)
#include <Eigen/Core>

#include "common.hpp"

// GEMV routines Eigen's redirect references but these products never reach
// (the GEMV branch is taken only for vector shapes).
extern "C" {
void sgemv_(const char*, const int*, const int*, const float*, const float*, const int*, const float*, const int*,
            const float*, float*, const int*) {
  std::fprintf(stderr, "unexpected call of sgemv_\n");
  std::abort();
}
void dgemv_(const char*, const int*, const int*, const double*, const double*, const int*, const double*,
            const int*, const double*, double*, const int*) {
  std::fprintf(stderr, "unexpected call of dgemv_\n");
  std::abort();
}
}

using Eigen::Dynamic;
using Eigen::Index;
using Eigen::OuterStride;

template <typename T>
using Mat = Eigen::Matrix<T, Dynamic, Dynamic>;
template <typename T>
using CMap = Eigen::Map<Mat<T>, 0, OuterStride<>>;

template <typename T>
Mat<T> loop_product(const Mat<T>& A, const Mat<T>& B, bool ta) {
  const Index m = ta ? A.cols() : A.rows(), k = ta ? A.rows() : A.cols(), n = B.cols();
  Mat<T> C = Mat<T>::Zero(m, n);
  for (Index j = 0; j < n; ++j)
    for (Index p = 0; p < k; ++p)
      for (Index i = 0; i < m; ++i) C(i, j) += (ta ? A(p, i) : A(i, p)) * B(p, j);
  return C;
}

template <typename T>
int run_all(const char* tname) {
  int fails = 0;
  char name[96];
  const Index m = 70, n = 45, k = 33;
  for (int ta = 0; ta < 2; ++ta) {
    Mat<T> A(ta ? k : m, ta ? m : k), B(k, n), C0(m, n);
    fill_int(A, 1 + ta);
    fill_int(B, 3 + ta);
    fill_int(C0, 5 + ta);
    Mat<T> want = C0 + T(2) * loop_product(A, B, ta);
    // host Maps
    Mat<T> C = C0;
    CMap<T> a(A.data(), A.rows(), A.cols(), OuterStride<>(A.rows())), b(B.data(), k, n, OuterStride<>(k)),
        c(C.data(), m, n, OuterStride<>(m));
    if (ta)
      c.noalias() += T(2) * a.transpose() * b;
    else
      c.noalias() += T(2) * a * b;
    std::snprintf(name, sizeof name, "%s/host_maps/ta=%d", tname, ta);
    fails += report(name, C == want);
    // device Maps
    DeviceBuffer<T> dA(A.size()), dB(B.size()), dC(C0.size());
    dA.upload(A.data());
    dB.upload(B.data());
    dC.upload(C0.data());
    CMap<T> da(dA.ptr, A.rows(), A.cols(), OuterStride<>(A.rows())), db(dB.ptr, k, n, OuterStride<>(k)),
        dc(dC.ptr, m, n, OuterStride<>(m));
    if (ta)
      dc.noalias() += T(2) * da.transpose() * db;
    else
      dc.noalias() += T(2) * da * db;
    Mat<T> got(m, n);
    dC.download(got.data());
    std::snprintf(name, sizeof name, "%s/device_maps/ta=%d", tname, ta);
    fails += report(name, got == want);
  }
  return fails;
}

int main() {
  int fails = run_all<float>("float") + run_all<double>("double");
  std::printf("%s (%d failures)\n", fails ? "SOME FAILED" : "ALL PASSED", fails);
  return fails ? 1 : 0;
}
