// Products on device Maps through Eigen's own device seam:
//   C.device(dev).noalias() {=, +=, -=} expr
// Each case runs the same expression on host matrices with Eigen's CPU path
// and compares bit for bit (small-integer operands).
#include "../../include/eigen_amd_eigen.h"
#include "common.hpp"

using Eigen::Dynamic;
using Eigen::Index;
using Eigen::OuterStride;

template <typename T, int Order = Eigen::ColMajor>
using Mat = Eigen::Matrix<T, Dynamic, Dynamic, Order>;
template <typename T, int Order = Eigen::ColMajor>
using DMap = Eigen::Map<Mat<T, Order>, 0, OuterStride<>>;

template <typename T, int Order>
struct Dev {
  Mat<T, Order> host;
  Index ld;
  DeviceBuffer<T> buf;
  Dev(Index r, Index c, unsigned seed, Index pad = 0)
      : host(r, c), ld((Order == Eigen::ColMajor ? r : c) + pad), buf((size_t)ld * (Order == Eigen::ColMajor ? c : r)) {
    fill_int(host, seed);
    std::vector<T> h(buf.n, T(-99));
    for (Index j = 0; j < c; ++j)
      for (Index i = 0; i < r; ++i) h[Order == Eigen::ColMajor ? i + j * ld : j + i * ld] = host(i, j);
    buf.upload(h.data());
  }
  DMap<T, Order> map() { return DMap<T, Order>(buf.ptr, host.rows(), host.cols(), OuterStride<>(ld)); }
  Mat<T, Order> fetch() const {
    std::vector<T> h(buf.n);
    buf.download(h.data());
    Mat<T, Order> out(host.rows(), host.cols());
    for (Index j = 0; j < host.cols(); ++j)
      for (Index i = 0; i < host.rows(); ++i) out(i, j) = h[Order == Eigen::ColMajor ? i + j * ld : j + i * ld];
    return out;
  }
};

template <typename T>
int run_all(const char* tname) {
  int fails = 0;
  eigen_amd::HipDevice dev;
  char name[128];
  const Index m = 70, n = 45, k = 33;
  auto tag = [&](const char* what) {
    std::snprintf(name, sizeof name, "%s/%s", tname, what);
    return name;
  };
  {  // =
    Dev<T, 0> A(m, k, 1), B(k, n, 2), C(m, n, 3);
    auto c = C.map();
    c.device(dev).noalias() = A.map() * B.map();
    Mat<T> want = A.host * B.host;
    fails += report(tag("assign"), C.fetch() == want && dev.status == EH_SUCCESS);
  }
  {  // += alpha * A * B
    Dev<T, 0> A(m, k, 4), B(k, n, 5), C(m, n, 6);
    auto c = C.map();
    c.device(dev).noalias() += T(2) * A.map() * B.map();
    Mat<T> want = C.host + T(2) * A.host * B.host;
    fails += report(tag("add_scaled"), C.fetch() == want);
  }
  {  // -= A^T * B with A stored k x m, padded strides
    Dev<T, 0> A(k, m, 7, 5), B(k, n, 8, 3), C(m, n, 9, 11);
    auto c = C.map();
    c.device(dev).noalias() -= A.map().transpose() * B.map();
    Mat<T> want = C.host - A.host.transpose() * B.host;
    fails += report(tag("sub_transposed_padded"), C.fetch() == want);
  }
  {  // = (A*3)^T * B^T * -1 style factors on both sides
    Dev<T, 0> A(k, m, 10), B(n, k, 11), C(m, n, 12);
    auto c = C.map();
    c.device(dev).noalias() = (A.map() * T(3)).transpose() * (-B.map().transpose());
    Mat<T> want = (A.host * T(3)).transpose() * (-B.host.transpose());
    fails += report(tag("factors_and_negation"), C.fetch() == want);
  }
  {  // row-major operands and destination
    Dev<T, Eigen::RowMajor> A(m, k, 13), C(m, n, 15);
    Dev<T, 0> B(k, n, 14);
    auto c = C.map();
    c.device(dev).noalias() += A.map() * B.map();
    Mat<T, Eigen::RowMajor> want = C.host + A.host * B.host;
    fails += report(tag("row_major"), C.fetch() == want);
  }
  {  // blocks of device maps
    Dev<T, 0> A(m, k, 16), B(k, n, 17), C(m, n, 18);
    auto c = C.map();
    c.block(3, 4, 40, 30).device(dev).noalias() = A.map().block(5, 2, 40, 20) * B.map().block(1, 6, 20, 30);
    Mat<T> want = C.host;
    want.block(3, 4, 40, 30) = A.host.block(5, 2, 40, 20) * B.host.block(1, 6, 20, 30);
    fails += report(tag("blocks"), C.fetch() == want);
  }
  {  // matrix * vector and tiny products go to the device as well
    Dev<T, 0> A(m, k, 19), x(k, 1, 20), y(m, 1, 21);
    auto yv = y.map();
    yv.device(dev).noalias() += A.map() * x.map();
    Mat<T> want = y.host + A.host * x.host;
    fails += report(tag("gemv_shape"), y.fetch() == want);
    Dev<T, 0> a(2, 3, 22), b(3, 1, 23), cc(2, 1, 24);
    auto cm = cc.map();
    cm.device(dev).noalias() = a.map() * b.map();
    Mat<T> want2 = a.host * b.host;
    fails += report(tag("tiny"), cc.fetch() == want2);
  }
  {  // destination with an inner stride (Stride<Dynamic, 2>), after test/product.h:345-357:
     // row-major square map over every other element of a 2r x 2r buffer
    const Index r = 37;
    using RowSq = Mat<T, Eigen::RowMajor>;
    using SMap = Eigen::Map<RowSq, 0, Eigen::Stride<Dynamic, 2>>;
    Mat<T> m1(r, r), m2(r, r), init(r, r);
    fill_int(m1, 28);
    fill_int(m2, 29);
    fill_int(init, 30);
    DeviceBuffer<T> d1(r * r), d2(r * r), buf(4 * r * r);
    d1.upload(m1.data());
    d2.upload(m2.data());
    const Eigen::Map<const Mat<T>> M1(d1.ptr, r, r), M2(d2.ptr, r, r);
    std::vector<T> hb(4 * r * r);
    auto put = [&](const Mat<T>& v) {  // the strided elements get v, the gaps -99
      std::fill(hb.begin(), hb.end(), T(-99));
      for (Index i = 0; i < r; ++i)
        for (Index j = 0; j < r; ++j) hb[i * 2 * r + j * 2] = v(i, j);
      buf.upload(hb.data());
    };
    auto get = [&](bool& gaps_ok) {
      buf.download(hb.data());
      Mat<T> out(r, r);
      gaps_ok = true;
      for (size_t q = 0; q < hb.size(); ++q) {
        const Index i = (Index)q / (2 * r), j = ((Index)q % (2 * r)) / 2;
        if ((Index)q % (2 * r) % 2 == 0 && i < r && j < r)
          out(i, j) = hb[q];
        else if (hb[q] != T(-99))
          gaps_ok = false;
      }
      return out;
    };
    SMap map1(buf.ptr, r, r, Eigen::Stride<Dynamic, 2>(2 * r, 2));
    const Mat<T> prod = m1 * m2.transpose();
    bool gaps = false;
    put(init);
    map1.device(dev).noalias() = M1 * M2.transpose();
    fails += report(tag("inner_stride_assign"), get(gaps) == prod && gaps && dev.status == EH_SUCCESS);
    put(init);
    map1.device(dev).noalias() += M1 * M2.transpose();
    fails += report(tag("inner_stride_add"), get(gaps) == Mat<T>(init + prod) && gaps);
    put(init);
    map1.device(dev).noalias() -= T(2) * M1 * M2.transpose();
    fails += report(tag("inner_stride_sub_scaled"), get(gaps) == Mat<T>(init - T(2) * prod) && gaps);
  }
  {  // split-BF16 precision keeps integer products exact
    if (std::is_same<T, float>::value) {
      eigen_amd::HipDevice d3;
      d3.precision = EH_PREC_SPLIT3;
      Dev<T, 0> A(m, k, 25), B(k, n, 26), C(m, n, 27);
      auto c = C.map();
      c.device(d3).noalias() = A.map() * B.map();
      Mat<T> want = A.host * B.host;
      fails += report(tag("split3"), C.fetch() == want);
    }
  }
  {  // self-adjoint products through the seam (eh_symm): S * B, B * S, row-major
    Dev<T, 0> S(m, m, 31, 2), B(m, n, 32), C(m, n, 33);
    auto c = C.map();
    c.device(dev).noalias() = S.map().template selfadjointView<Eigen::Lower>() * B.map();
    Mat<T> want = S.host.template selfadjointView<Eigen::Lower>() * B.host;
    fails += report(tag("symm_lower_left"), C.fetch() == want && dev.status == EH_SUCCESS);
    Dev<T, 0> S2(n, n, 34), B2(m, n, 35), C2(m, n, 36);
    auto c2 = C2.map();
    c2.device(dev).noalias() -= B2.map() * T(2) * S2.map().template selfadjointView<Eigen::Upper>();
    Mat<T> want2 = C2.host - B2.host * T(2) * S2.host.template selfadjointView<Eigen::Upper>();
    fails += report(tag("symm_upper_right_scaled_sub"), C2.fetch() == want2);
    Dev<T, Eigen::RowMajor> S3(m, m, 37, 1), B3(m, n, 38, 3), C3(m, n, 39, 2);
    auto c3 = C3.map();
    c3.device(dev).noalias() += S3.map().template selfadjointView<Eigen::Upper>() * B3.map();
    Mat<T, Eigen::RowMajor> want3 = C3.host + S3.host.template selfadjointView<Eigen::Upper>() * B3.host;
    fails += report(tag("symm_row_major_add"), C3.fetch() == want3);
  }
  return fails;
}

int main() {
  int fails = run_all<float>("float") + run_all<double>("double");
  std::printf("%s (%d failures)\n", fails ? "SOME FAILED" : "ALL PASSED", fails);
  return fails ? 1 : 0;
}
