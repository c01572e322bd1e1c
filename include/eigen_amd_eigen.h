// eigen_amd_eigen.h -- Eigen drop-in: products on device-resident Maps.
//
//   #include <Eigen/Core>
//   #include "eigen_amd_eigen.h"
//   eigen_amd::HipDevice dev;                       // stream, precision
//   Eigen::Map<Eigen::MatrixXf, 0, Eigen::OuterStride<>> A(dA, m, k, OuterStride<>(lda)), ...;
//   C.device(dev).noalias() += alpha * A * B;       // runs eh_gemm on dev.stream
//   C.device(dev).noalias() = A.transpose() * B;
//   C.device(dev).noalias() -= A * B;
//
// The seam is the reference's own device dispatch:
// Eigen::internal::AssignmentWithDevice<Dst, Product<Lhs,Rhs,Opt>, Func,
// Device, Dense2Dense> (Eigen/src/Core/DeviceWrapper.h:59-70), reached from
// DeviceWrapper::noalias() (DeviceWrapper.h:46-48) through
// call_assignment_no_alias (DeviceWrapper.h:94-117).  This header specialises
// it for eigen_amd::HipDevice.  Operands are taken apart exactly as Eigen's
// own GEMM driver does it (GeneralMatrixMatrix.h:476-514): blas_traits
// (Eigen/src/Core/util/BlasUtil.h:459-567) strips scalar factors, negation
// and transposes and yields the underlying data pointer, stride and storage
// order, and combine_scalar_factors (BlasUtil.h:593-621) gives alpha.  No
// device data is touched on the host: unlike the host path
// (GeneralMatrixMatrix.h:445-457, 480-490) there is no coefficient-based or
// GEMV fallback, every shape goes to the device kernels.
//
// Assignment kinds: `=` is C = alpha*op(A)*op(B) (beta = 0, C not read),
// `+=` beta = 1, `-=` beta = 1 with alpha negated (GeneralMatrixMatrix.h:
// 454, 465, 472).  A row-major destination is computed as the transposed
// product (GeneralMatrixMatrix.h:98-115); a destination with an inner stride
// (Eigen's resIncr, GeneralMatrixMatrix.h:117-137) goes through
// eh_gemm_strided.  Only float and double, and direct-access operands with
// unit inner stride (Matrix / Map / Block / Ref and their transposes,
// possibly scaled or negated) are accepted; anything else is a compile-time
// error.  Errors from the device call are stored in
// HipDevice::status (and reported through eigen_assert).
#ifndef EIGEN_AMD_EIGEN_H
#define EIGEN_AMD_EIGEN_H

#include <Eigen/Core>
#include <type_traits>

#include "eigen_amd.h"

namespace eigen_amd {

struct HipDevice {
  void* stream = nullptr;                  // hipStream_t the products run on (NULL = default stream)
  eh_precision precision = EH_PREC_AUTO;   // FP32 arithmetic (plain / split-BF16)
  bool synchronize = false;                // wait for the stream after each product
  eh_status status = EH_SUCCESS;           // status of the last product
};

}  // namespace eigen_amd

namespace Eigen {
namespace internal {

template <typename DstXprType, typename Lhs, typename Rhs, int Options, typename Functor, typename Weak>
struct AssignmentWithDevice<DstXprType, Product<Lhs, Rhs, Options>, Functor, eigen_amd::HipDevice, Dense2Dense, Weak> {
  using SrcXprType = Product<Lhs, Rhs, Options>;
  using Scalar = typename SrcXprType::Scalar;
  using LhsBlasTraits = blas_traits<Lhs>;
  using RhsBlasTraits = blas_traits<Rhs>;
  using ActualLhs = remove_all_t<typename LhsBlasTraits::DirectLinearAccessType>;
  using ActualRhs = remove_all_t<typename RhsBlasTraits::DirectLinearAccessType>;

  static_assert(std::is_same<Scalar, float>::value || std::is_same<Scalar, double>::value,
                "eigen_amd device products support float and double");
  static_assert(std::is_same<typename DstXprType::Scalar, Scalar>::value, "mixed scalar types are not supported");
  static_assert(!LhsBlasTraits::NeedToConjugate && !RhsBlasTraits::NeedToConjugate, "real scalars only");
  // Eigen would evaluate an operand without unit inner stride into a host
  // temporary (blas_traits::DirectLinearAccessType), i.e. read device memory
  // on the host: such operands must be copied to a contiguous device buffer
  // by the caller.
  static_assert(bool(LhsBlasTraits::HasUsableDirectAccess) && bool(RhsBlasTraits::HasUsableDirectAccess),
                "device product operands need direct access with unit inner stride");

  static constexpr bool kAssign = std::is_same<Functor, assign_op<Scalar, Scalar>>::value;
  static constexpr bool kAdd = std::is_same<Functor, add_assign_op<Scalar, Scalar>>::value;
  static constexpr bool kSub = std::is_same<Functor, sub_assign_op<Scalar, Scalar>>::value;
  static_assert(kAssign || kAdd || kSub, "only =, += and -= are supported on the device");

  template <typename X>
  static bool row_major(const X&) {
    return (int(X::Flags) & RowMajorBit) != 0;
  }

  static void run(DstXprType& dst, const SrcXprType& src, const Functor&, eigen_amd::HipDevice& dev) {
    const auto& a_lhs = src.lhs();
    const auto& a_rhs = src.rhs();
    eigen_assert(dst.rows() == a_lhs.rows() && dst.cols() == a_rhs.cols());
    const ActualLhs& lhs = LhsBlasTraits::extract(a_lhs);
    const ActualRhs& rhs = RhsBlasTraits::extract(a_rhs);
    Scalar alpha = combine_scalar_factors(Scalar(1), a_lhs, a_rhs);
    if (kSub) alpha = -alpha;
    const Scalar beta = kAssign ? Scalar(0) : Scalar(1);
    eigen_assert(lhs.innerStride() == 1 && rhs.innerStride() == 1);

    // Column-major view of each operand: a row-major m x k matrix with outer
    // stride s is the column-major k x m matrix X^T with leading dimension s.
    const bool lt = row_major(lhs), rt = row_major(rhs);
    int64_t m = dst.rows(), n = dst.cols(), k = a_lhs.cols();
    const void* A = lhs.data();
    const void* B = rhs.data();
    int64_t lda = lhs.outerStride(), ldb = rhs.outerStride();
    eh_op opa = lt ? EH_OP_T : EH_OP_N, opb = rt ? EH_OP_T : EH_OP_N;
    if (row_major(dst)) {
      // C^T = op(B)^T op(A)^T, computed into the column-major view of C
      std::swap(m, n);
      std::swap(A, B);
      std::swap(lda, ldb);
      const eh_op oa = opa;
      opa = opb == EH_OP_N ? EH_OP_T : EH_OP_N;
      opb = oa == EH_OP_N ? EH_OP_T : EH_OP_N;
    }
    // leading dimensions of empty / degenerate operands
    if (lda < 1) lda = 1;
    if (ldb < 1) ldb = 1;
    int64_t ldc = dst.outerStride() < 1 ? 1 : dst.outerStride();
    if (m == 0 || n == 0) return;
    const eh_dtype dt = std::is_same<Scalar, float>::value ? EH_F32 : EH_F64;
    const eh_precision prec = dt == EH_F32 ? dev.precision : EH_PREC_AUTO;
    // a destination with an inner stride keeps it in either storage order:
    // element (i, j) of the column-major view is at i*inner + j*outer
    const int64_t incc = dst.innerStride();
    if (incc == 1)
      dev.status = eh_gemm(dt, opa, opb, m, n, k, &alpha, A, lda, B, ldb, &beta, dst.data(), ldc, prec, dev.stream);
    else
      dev.status = eh_gemm_strided(dt, opa, opb, m, n, k, &alpha, A, lda, B, ldb, &beta, dst.data(), incc, ldc, prec,
                                   dev.stream);
    if (dev.status == EH_SUCCESS && dev.synchronize) dev.status = eh_synchronize(dev.stream);
    eigen_assert(dev.status == EH_SUCCESS && "eigen_amd device product failed (see eh_last_error)");
  }
};

}  // namespace internal
}  // namespace Eigen

#endif  // EIGEN_AMD_EIGEN_H
