// eigen_amd_eigen.h -- Eigen drop-in: products on device-resident matrices.
//
// Two ways in, both ending in the same device call (device_product below):
//
// 1. Eigen's own device seam, on ordinary Maps of device memory:
//
//   eigen_amd::HipDevice dev;                       // stream, precision
//   Eigen::Map<Eigen::MatrixXf, 0, Eigen::OuterStride<>> A(dA, m, k, OuterStride<>(lda)), ...;
//   C.device(dev).noalias() += alpha * A * B;       // runs eh_gemm on dev.stream
//   C.device(dev).noalias() = A.transpose() * B;
//   C.device(dev).noalias() -= A * B;
//
//    The seam is Eigen::internal::AssignmentWithDevice<Dst, Product<Lhs,Rhs,
//    Opt>, Func, Device, Dense2Dense> (Eigen/src/Core/DeviceWrapper.h:59-70),
//    reached from DeviceWrapper::noalias() (DeviceWrapper.h:46-48) through
//    call_assignment_no_alias (DeviceWrapper.h:94-117); this header
//    specialises it for eigen_amd::HipDevice.
//
// 2. A device map type, with no `.device()` at the call site:
//
//   eigen_amd::DeviceMap<Eigen::MatrixXf> A(dA, m, k), B(dB, k, n), C(dC, m, n);
//   C = A * B;  C.noalias() = A * B;  C += 2 * A.transpose() * B;  C -= A * B;
//   eigen_amd::DeviceMap<Eigen::VectorXf> x(dx, k), y(dy, m);
//   y = A * x;                                       // vector and tiny shapes too
//
//    DeviceMap<Plain> is an Eigen::Map<Plain, 0, OuterStride<>> whose
//    assignment operators take products only: `=`, `+=` and `-=` of a product
//    (with or without noalias()) run on the device on the stream of
//    eigen_amd::default_device(), for every shape.  This is what Eigen's
//    host evaluation of `dst = lhs * rhs` cannot do on device memory: its
//    `=` clears dst on the host first (GeneralMatrixMatrix.h:450-456), a
//    plain `=` evaluates into a host temporary (evaluator_assume_aliasing,
//    ProductEvaluators.h:88, AssignEvaluator.h:981-984), and small or
//    vector shapes take the coefficient-based or GEMV host kernels
//    (GeneralProduct.h:26-150).  Assigning anything but a product to a
//    DeviceMap is a compile-time error (it would read device memory on the
//    host); operands may be DeviceMaps, Maps of device memory, or
//    transposes / scaled / negated / conjugated forms of them.
//
// Operands are taken apart exactly as Eigen's own GEMM driver does it
// (GeneralMatrixMatrix.h:476-514): blas_traits
// (Eigen/src/Core/util/BlasUtil.h:459-567) strips scalar factors, negation,
// conjugation and transposes and yields the underlying data pointer, stride
// and storage order, and combine_scalar_factors (BlasUtil.h:593-621) gives
// alpha.  No device data is touched on the host.
//
// Assignment kinds: `=` is C = alpha*op(A)*op(B) (beta = 0, C not read),
// `+=` beta = 1, `-=` beta = 1 with alpha negated (GeneralMatrixMatrix.h:
// 454, 465, 472).  A row-major destination is computed as the transposed
// product (GeneralMatrixMatrix.h:98-115); a real destination with an inner
// stride (Eigen's resIncr, GeneralMatrixMatrix.h:117-137) goes through
// eh_gemm_strided.  Scalars: float, double (eh_gemm) and std::complex<float>,
// std::complex<double> (eh_gemm_complex).  A conjugated operand must be
// stored so that it reaches the device as a conjugate transpose (op C), e.g.
// A.adjoint() of a column-major A; conj(A) without a transpose has no BLAS
// op and is a compile-time error.  Errors from the device call are stored in
// HipDevice::status (and reported through eigen_assert).
#ifndef EIGEN_AMD_EIGEN_H
#define EIGEN_AMD_EIGEN_H

#include <Eigen/Core>
#include <complex>
#include <type_traits>

#include "eigen_amd.h"

namespace eigen_amd {

struct HipDevice {
  void* stream = nullptr;                  // hipStream_t the products run on (NULL = default stream)
  eh_precision precision = EH_PREC_AUTO;   // FP32 arithmetic (plain / split-BF16)
  bool synchronize = false;                // wait for the stream after each product
  eh_status status = EH_SUCCESS;           // status of the last product
};

// The device DeviceMap assignments run on: one per thread, so that threads
// can pick their own streams and read their own status.
inline HipDevice& default_device() {
  static thread_local HipDevice dev;
  return dev;
}

namespace detail {

enum class Kind { Assign, Add, Sub };

template <typename T>
struct scalar_info {
  static constexpr bool supported = false;
};
template <>
struct scalar_info<float> {
  static constexpr bool supported = true, complex = false;
  static constexpr eh_dtype dtype = EH_F32;
};
template <>
struct scalar_info<double> {
  static constexpr bool supported = true, complex = false;
  static constexpr eh_dtype dtype = EH_F64;
};
template <>
struct scalar_info<std::complex<float>> {
  static constexpr bool supported = true, complex = true;
  static constexpr eh_dtype dtype = EH_F32;
};
template <>
struct scalar_info<std::complex<double>> {
  static constexpr bool supported = true, complex = true;
  static constexpr eh_dtype dtype = EH_F64;
};

template <typename X>
constexpr bool row_major() {
  return (int(X::Flags) & Eigen::RowMajorBit) != 0;
}

// BLAS op of an operand stored as X (column-major view) that enters the
// product as X, X^T, or their conjugates, after the destination's transpose.
// `flip`: the product is computed transposed (row-major destination).
template <bool RowMajorOperand, bool Conj, bool Flip>
struct operand_op {
  static constexpr bool transposed = RowMajorOperand != Flip;
  // a conjugate is only expressible together with a transpose (op C)
  static constexpr bool valid = !Conj || transposed;
  static constexpr eh_op op = !transposed ? EH_OP_N : (Conj ? EH_OP_C : EH_OP_T);
};

// dst (=, +=, -=) lhs * rhs on the device; Lhs / Rhs are the product's
// operand expressions (possibly transposed, scaled, negated, conjugated).
template <typename Dst, typename Lhs, typename Rhs>
void device_product(Dst& dst, const Lhs& a_lhs, const Rhs& a_rhs, Kind kind, HipDevice& dev) {
  using Scalar = typename Dst::Scalar;
  using Info = scalar_info<Scalar>;
  using LhsBlasTraits = Eigen::internal::blas_traits<Lhs>;
  using RhsBlasTraits = Eigen::internal::blas_traits<Rhs>;
  using ActualLhs = Eigen::internal::remove_all_t<typename LhsBlasTraits::DirectLinearAccessType>;
  using ActualRhs = Eigen::internal::remove_all_t<typename RhsBlasTraits::DirectLinearAccessType>;
  static_assert(Info::supported,
                "eigen_amd device products support float, double, std::complex<float> and std::complex<double>");
  static_assert(std::is_same<typename Lhs::Scalar, Scalar>::value && std::is_same<typename Rhs::Scalar, Scalar>::value,
                "mixed scalar types are not supported");
  // Eigen would evaluate an operand without unit inner stride into a host
  // temporary (blas_traits::DirectLinearAccessType), i.e. read device memory
  // on the host: such operands must be copied to a contiguous device buffer
  // by the caller.
  static_assert(bool(LhsBlasTraits::HasUsableDirectAccess) && bool(RhsBlasTraits::HasUsableDirectAccess),
                "device product operands need direct access with unit inner stride");
  constexpr bool flip = row_major<Dst>();
  using OpA = operand_op<row_major<ActualLhs>(), bool(LhsBlasTraits::NeedToConjugate), flip>;
  using OpB = operand_op<row_major<ActualRhs>(), bool(RhsBlasTraits::NeedToConjugate), flip>;
  static_assert(OpA::valid && OpB::valid,
                "a conjugated device operand must enter the product transposed (e.g. A.adjoint()); "
                "copy conj(A) to a device buffer first");

  eigen_assert(dst.rows() == a_lhs.rows() && dst.cols() == a_rhs.cols());
  const ActualLhs& lhs = LhsBlasTraits::extract(a_lhs);
  const ActualRhs& rhs = RhsBlasTraits::extract(a_rhs);
  Scalar alpha = Eigen::internal::combine_scalar_factors(Scalar(1), a_lhs, a_rhs);
  if (kind == Kind::Sub) alpha = -alpha;
  const Scalar beta = kind == Kind::Assign ? Scalar(0) : Scalar(1);
  eigen_assert(lhs.innerStride() == 1 && rhs.innerStride() == 1);

  // Column-major view of each operand: a row-major m x k matrix with outer
  // stride s is the column-major k x m matrix X^T with leading dimension s.
  int64_t m = dst.rows(), n = dst.cols(), k = a_lhs.cols();
  const void* A = lhs.data();
  const void* B = rhs.data();
  int64_t lda = lhs.outerStride(), ldb = rhs.outerStride();
  eh_op opa = OpA::op, opb = OpB::op;
  if (flip) {  // C^T = op(B)^T op(A)^T, computed into the column-major view of C
    std::swap(m, n);
    std::swap(A, B);
    std::swap(lda, ldb);
    std::swap(opa, opb);
  }
  // leading dimensions of empty / degenerate operands
  if (lda < 1) lda = 1;
  if (ldb < 1) ldb = 1;
  const int64_t ldc = dst.outerStride() < 1 ? 1 : dst.outerStride();
  if (m == 0 || n == 0) return;
  const eh_dtype dt = Info::dtype;
  const eh_precision prec = dt == EH_F32 ? dev.precision : EH_PREC_AUTO;
  // a destination with an inner stride keeps it in either storage order:
  // element (i, j) of the column-major view is at i*inner + j*outer
  const int64_t incc = dst.innerStride();
  if (Info::complex) {
    dev.status = incc == 1 ? eh_gemm_complex(dt, opa, opb, m, n, k, &alpha, A, lda, B, ldb, &beta, dst.data(), ldc,
                                             prec, dev.stream)
                           : EH_ERR_UNSUPPORTED;  // no strided complex destination
  } else if (incc == 1) {
    dev.status = eh_gemm(dt, opa, opb, m, n, k, &alpha, A, lda, B, ldb, &beta, dst.data(), ldc, prec, dev.stream);
  } else {
    dev.status = eh_gemm_strided(dt, opa, opb, m, n, k, &alpha, A, lda, B, ldb, &beta, dst.data(), incc, ldc, prec,
                                 dev.stream);
  }
  if (dev.status == EH_SUCCESS && dev.synchronize) dev.status = eh_synchronize(dev.stream);
  eigen_assert(dev.status == EH_SUCCESS && "eigen_amd device product failed (see eh_last_error)");
}

template <typename T>
struct always_false : std::false_type {};

}  // namespace detail

// An Eigen::Map of device memory whose assignments are device products.
template <typename Plain>
class DeviceMap : public Eigen::Map<Plain, 0, Eigen::OuterStride<>> {
 public:
  using Base = Eigen::Map<Plain, 0, Eigen::OuterStride<>>;
  using Scalar = typename Plain::Scalar;
  using Index = Eigen::Index;

  // rows x cols with leading dimension ld (the outer stride: between columns
  // of a column-major Plain, between rows of a row-major one)
  DeviceMap(Scalar* data, Index rows, Index cols, Index ld) : Base(data, rows, cols, Eigen::OuterStride<>(ld)) {}
  DeviceMap(Scalar* data, Index rows, Index cols)
      : Base(data, rows, cols, Eigen::OuterStride<>(Plain::IsRowMajor ? cols : rows)) {}
  // a vector of `size` elements
  DeviceMap(Scalar* data, Index size)
      : Base(data, Plain::RowsAtCompileTime == 1 ? 1 : size, Plain::RowsAtCompileTime == 1 ? size : 1,
             Eigen::OuterStride<>(size)) {}
  DeviceMap(const DeviceMap&) = default;

  template <typename L, typename R, int O>
  DeviceMap& operator=(const Eigen::Product<L, R, O>& p) {
    return run(p, detail::Kind::Assign);
  }
  template <typename L, typename R, int O>
  DeviceMap& operator+=(const Eigen::Product<L, R, O>& p) {
    return run(p, detail::Kind::Add);
  }
  template <typename L, typename R, int O>
  DeviceMap& operator-=(const Eigen::Product<L, R, O>& p) {
    return run(p, detail::Kind::Sub);
  }
  // Anything else would be evaluated coefficient by coefficient on the host.
  template <typename Other>
  DeviceMap& operator=(const Eigen::DenseBase<Other>&) {
    static_assert(detail::always_false<Other>::value, "a DeviceMap can only be assigned a product (A * B)");
    return *this;
  }
  template <typename Other>
  DeviceMap& operator+=(const Eigen::DenseBase<Other>&) {
    static_assert(detail::always_false<Other>::value, "a DeviceMap can only be updated by a product (A * B)");
    return *this;
  }
  template <typename Other>
  DeviceMap& operator-=(const Eigen::DenseBase<Other>&) {
    static_assert(detail::always_false<Other>::value, "a DeviceMap can only be updated by a product (A * B)");
    return *this;
  }
  DeviceMap& operator=(const DeviceMap&) = delete;  // an element copy on the host

  // A sub-block stays a DeviceMap (same storage order and leading dimension),
  // so products assigned into it run on the device too.  (Eigen's own
  // block(), col() and the like return plain Eigen blocks, whose assignments
  // would run on the host.)
  DeviceMap block(Index i, Index j, Index rows, Index cols) {
    eigen_assert(i >= 0 && j >= 0 && rows >= 0 && cols >= 0 && i + rows <= this->rows() && j + cols <= this->cols());
    Scalar* p = this->data() + (Plain::IsRowMajor ? i * this->outerStride() + j : i + j * this->outerStride());
    return DeviceMap(p, rows, cols, this->outerStride());
  }

  // C.noalias() = / += / -= A * B: the same device products (they never
  // evaluate into a temporary)
  struct NoAlias {
    DeviceMap& m;
    template <typename L, typename R, int O>
    DeviceMap& operator=(const Eigen::Product<L, R, O>& p) {
      return m.run(p, detail::Kind::Assign);
    }
    template <typename L, typename R, int O>
    DeviceMap& operator+=(const Eigen::Product<L, R, O>& p) {
      return m.run(p, detail::Kind::Add);
    }
    template <typename L, typename R, int O>
    DeviceMap& operator-=(const Eigen::Product<L, R, O>& p) {
      return m.run(p, detail::Kind::Sub);
    }
  };
  NoAlias noalias() { return NoAlias{*this}; }

 private:
  template <typename L, typename R, int O>
  DeviceMap& run(const Eigen::Product<L, R, O>& p, detail::Kind kind) {
    Base& self = *this;
    detail::device_product(self, p.lhs(), p.rhs(), kind, default_device());
    return *this;
  }
};

}  // namespace eigen_amd

namespace Eigen {
namespace internal {

template <typename DstXprType, typename Lhs, typename Rhs, int Options, typename Functor, typename Weak>
struct AssignmentWithDevice<DstXprType, Product<Lhs, Rhs, Options>, Functor, eigen_amd::HipDevice, Dense2Dense, Weak> {
  using SrcXprType = Product<Lhs, Rhs, Options>;
  using Scalar = typename SrcXprType::Scalar;
  static_assert(std::is_same<typename DstXprType::Scalar, Scalar>::value, "mixed scalar types are not supported");
  static constexpr bool kAssign = std::is_same<Functor, assign_op<Scalar, Scalar>>::value;
  static constexpr bool kAdd = std::is_same<Functor, add_assign_op<Scalar, Scalar>>::value;
  static constexpr bool kSub = std::is_same<Functor, sub_assign_op<Scalar, Scalar>>::value;
  static_assert(kAssign || kAdd || kSub, "only =, += and -= are supported on the device");

  static void run(DstXprType& dst, const SrcXprType& src, const Functor&, eigen_amd::HipDevice& dev) {
    using eigen_amd::detail::Kind;
    eigen_amd::detail::device_product(dst, src.lhs(), src.rhs(), kAssign ? Kind::Assign : (kAdd ? Kind::Add : Kind::Sub),
                                      dev);
  }
};

}  // namespace internal
}  // namespace Eigen

#endif  // EIGEN_AMD_EIGEN_H
