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
//    specialises it for eigen_amd::HipDevice.  A product with a
//    selfAdjointView<>() factor (S * B or B * S) runs as eh_symm / eh_hemm.
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

// The arguments of one column-major BLAS product for dst (=, +=, -=)
// lhs * rhs; Lhs / Rhs are the product's operand expressions (possibly
// transposed, scaled, negated, conjugated).  A row-major destination is
// computed as the transposed product into its column-major view (`flip`).
template <typename Scalar>
struct GemmArgs {
  int64_t m, n, k;
  const void* A;
  int64_t lda;
  eh_op opa;
  const void* B;
  int64_t ldb;
  eh_op opb;
  Scalar alpha, beta;
  bool flip;
};

template <typename Dst, typename Lhs, typename Rhs>
GemmArgs<typename Dst::Scalar> gemm_args(const Dst& dst, const Lhs& a_lhs, const Rhs& a_rhs, Kind kind) {
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
  GemmArgs<Scalar> g;
  g.alpha = Eigen::internal::combine_scalar_factors(Scalar(1), a_lhs, a_rhs);
  if (kind == Kind::Sub) g.alpha = -g.alpha;
  g.beta = kind == Kind::Assign ? Scalar(0) : Scalar(1);
  eigen_assert(lhs.innerStride() == 1 && rhs.innerStride() == 1);

  // Column-major view of each operand: a row-major m x k matrix with outer
  // stride s is the column-major k x m matrix X^T with leading dimension s.
  g.m = dst.rows();
  g.n = dst.cols();
  g.k = a_lhs.cols();
  g.A = lhs.data();
  g.B = rhs.data();
  g.lda = lhs.outerStride();
  g.ldb = rhs.outerStride();
  g.opa = OpA::op;
  g.opb = OpB::op;
  g.flip = flip;
  if (flip) {  // C^T = op(B)^T op(A)^T, computed into the column-major view of C
    std::swap(g.m, g.n);
    std::swap(g.A, g.B);
    std::swap(g.lda, g.ldb);
    std::swap(g.opa, g.opb);
  }
  // leading dimensions of empty / degenerate operands
  if (g.lda < 1) g.lda = 1;
  if (g.ldb < 1) g.ldb = 1;
  return g;
}

// dst (=, +=, -=) lhs * rhs on the device.
template <typename Dst, typename Lhs, typename Rhs>
void device_product(Dst& dst, const Lhs& a_lhs, const Rhs& a_rhs, Kind kind, HipDevice& dev) {
  using Scalar = typename Dst::Scalar;
  using Info = scalar_info<Scalar>;
  const GemmArgs<Scalar> g = gemm_args(dst, a_lhs, a_rhs, kind);
  const int64_t ldc = dst.outerStride() < 1 ? 1 : dst.outerStride();
  if (g.m == 0 || g.n == 0) return;
  const eh_dtype dt = Info::dtype;
  const eh_precision prec = dt == EH_F32 ? dev.precision : EH_PREC_AUTO;
  // a destination with an inner stride keeps it in either storage order:
  // element (i, j) of the column-major view is at i*inner + j*outer
  const int64_t incc = dst.innerStride();
  if (Info::complex) {
    dev.status = incc == 1 ? eh_gemm_complex(dt, g.opa, g.opb, g.m, g.n, g.k, &g.alpha, g.A, g.lda, g.B, g.ldb,
                                             &g.beta, dst.data(), ldc, prec, dev.stream)
                           : EH_ERR_UNSUPPORTED;  // no strided complex destination
  } else if (incc == 1) {
    dev.status = eh_gemm(dt, g.opa, g.opb, g.m, g.n, g.k, &g.alpha, g.A, g.lda, g.B, g.ldb, &g.beta, dst.data(), ldc,
                         prec, dev.stream);
  } else {
    dev.status = eh_gemm_strided(dt, g.opa, g.opb, g.m, g.n, g.k, &g.alpha, g.A, g.lda, g.B, g.ldb, &g.beta,
                                 dst.data(), incc, ldc, prec, dev.stream);
  }
  if (dev.status == EH_SUCCESS && dev.synchronize) dev.status = eh_synchronize(dev.stream);
  eigen_assert(dev.status == EH_SUCCESS && "eigen_amd device product failed (see eh_last_error)");
}

// BLAS op codes of the level-3 views: 0 = N, 1 = T, 2 = C (conjugate
// transpose), 3 = the conjugate alone (no BLAS op; rejected where used).
constexpr int op_transpose(int o) { return o == 0 ? 1 : o == 1 ? 0 : o == 2 ? 3 : 2; }
constexpr int op_adjoint(int o, bool cplx) {
  return !cplx ? op_transpose(o) : (o == 0 ? 2 : o == 2 ? 0 : o == 1 ? 3 : 1);
}
constexpr eh_op eh_op_of(int o) { return o == 0 ? EH_OP_N : o == 1 ? EH_OP_T : EH_OP_C; }
// BLAS triangle of a Mode (Lower / Upper) view, flipped for a row-major
// matrix's column-major view
constexpr eh_uplo uplo_of(int mode, bool flip) {
  return ((mode & Eigen::Lower) != 0) != flip ? EH_LOWER : EH_UPPER;
}

// Finish a level-3 call: optional synchronisation, status kept on the device.
inline void finish(eh_status st, HipDevice& dev) {
  dev.status = st;
  if (dev.status == EH_SUCCESS && dev.synchronize) dev.status = eh_synchronize(dev.stream);
  eigen_assert(dev.status == EH_SUCCESS && "eigen_amd device level-3 call failed (see eh_last_error)");
}

// Eigen::SelfAdjointView<M, Mode> (any constness of M)
template <typename X>
struct sa_view : std::false_type {};
template <typename M, unsigned int Mode>
struct sa_view<Eigen::SelfAdjointView<M, Mode>> : std::true_type {
  static constexpr int mode = int(Mode);
};

// dst (=, +=, -=) factor * S * other (left) or factor * other * S (right),
// S the self-adjoint matrix (symmetric real / Hermitian complex) stored in
// the Mode triangle of `a` (eh_symm / eh_hemm).  `a`, `other` (after its
// scalar factors) and dst are stored matrices of one storage order: a
// row-major triple is the transposed product of their column-major views,
// C^T = B^T S^T, where S^T is the column-major view of S's storage (itself
// self-adjoint, stored in the other triangle).
template <int Mode, typename Dst, typename AX, typename Other>
void device_symm(Dst& dst, const AX& a, const Other& a_other, bool left, typename Dst::Scalar factor, Kind kind,
                 HipDevice& dev) {
  using Scalar = typename Dst::Scalar;
  using Info = scalar_info<Scalar>;
  using Traits = Eigen::internal::blas_traits<Other>;
  using ActualOther = Eigen::internal::remove_all_t<typename Traits::DirectLinearAccessType>;
  static_assert(Info::supported,
                "eigen_amd device products support float, double, std::complex<float> and std::complex<double>");
  static_assert(Mode == Eigen::Lower || Mode == Eigen::Upper, "a self-adjoint view is Lower or Upper");
  static_assert(std::is_same<typename AX::Scalar, Scalar>::value && std::is_same<typename Other::Scalar, Scalar>::value,
                "mixed scalar types are not supported");
  static_assert((int(Eigen::internal::traits<AX>::Flags) & Eigen::DirectAccessBit) != 0 &&
                    bool(Traits::HasUsableDirectAccess) && !bool(Traits::NeedToConjugate),
                "a device self-adjoint product needs stored operands (no conjugate)");
  static_assert(row_major<Dst>() == row_major<AX>() && row_major<Dst>() == row_major<ActualOther>(),
                "the self-adjoint matrix, B and C must share one storage order");
  constexpr bool flip = row_major<Dst>();
  const ActualOther& b = Traits::extract(a_other);
  const Scalar f = factor * Traits::extractScalarFactor(a_other);
  eigen_assert(a.rows() == a.cols() && a.innerStride() == 1 && b.innerStride() == 1 && dst.innerStride() == 1);
  eigen_assert(dst.rows() == (left ? a.rows() : b.rows()) && dst.cols() == (left ? b.cols() : a.cols()));
  const eh_side side = left != flip ? EH_LEFT : EH_RIGHT;
  const eh_uplo uplo = uplo_of(Mode, flip);
  const int64_t m = flip ? dst.cols() : dst.rows(), n = flip ? dst.rows() : dst.cols();
  const Scalar alpha = kind == Kind::Sub ? -f : f;
  const Scalar beta = kind == Kind::Assign ? Scalar(0) : Scalar(1);
  if (m == 0 || n == 0) return finish(EH_SUCCESS, dev);
  const int64_t lda = a.outerStride() < 1 ? 1 : a.outerStride();
  const int64_t ldb = b.outerStride() < 1 ? 1 : b.outerStride();
  const int64_t ldc = dst.outerStride() < 1 ? 1 : dst.outerStride();
  constexpr eh_dtype dt = Info::dtype;
  finish(Info::complex ? eh_hemm(dt, side, uplo, m, n, &alpha, a.data(), lda, b.data(), ldb, &beta, dst.data(), ldc,
                                 dev.stream)
                       : eh_symm(dt, side, uplo, m, n, &alpha, a.data(), lda, b.data(), ldb, &beta, dst.data(), ldc,
                                 dev.stream),
         dev);
}

template <typename T>
struct always_false : std::false_type {};

}  // namespace detail

template <typename Plain>
class DeviceMap;
template <typename Plain, int Mode, int Op = 0>
class DeviceTriangularView;
template <typename Plain, int Mode>
class DeviceSelfAdjointView;
template <typename View, typename P2, int Side>
struct DeviceTriangularProduct;
template <typename View, typename P2, int Side>
struct DeviceSelfAdjointProduct;

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

  // Level-3 views (see DeviceTriangularView / DeviceSelfAdjointView):
  //   A.triangularView<Lower>().solveInPlace(X)       -> eh_trsm
  //   X = A.triangularView<Lower>() * X                -> eh_trmm (in place)
  //   C.triangularView<Lower>() = / += / -= A * B      -> eh_gemmtr
  //   C = / += / -= A.selfadjointView<Lower>() * B     -> eh_symm / eh_hemm
  //   C.selfadjointView<Lower>().rankUpdate(U, a)      -> eh_syrk / eh_herk
  //   C.selfadjointView<Lower>().rankUpdate(U, V, a)   -> eh_syr2k / eh_her2k
  // They hide Map's own triangularView() / selfadjointView(), whose
  // evaluation would run on the host.
  template <unsigned int Mode>
  DeviceTriangularView<Plain, int(Mode)> triangularView() const {
    return DeviceTriangularView<Plain, int(Mode)>(*this);
  }
  template <unsigned int Mode>
  DeviceSelfAdjointView<Plain, int(Mode)> selfadjointView() const {
    return DeviceSelfAdjointView<Plain, int(Mode)>(*this);
  }
  // X = T * X, X = X * T: triangular product in place (BLAS trmm semantics:
  // the right-hand operand must be this same matrix)
  template <typename View, typename P2, int Side>
  DeviceMap& operator=(const DeviceTriangularProduct<View, P2, Side>& p) {
    p.run_into(*this);
    return *this;
  }
  // C = / += / -= S * B, B * S with S self-adjoint
  template <typename View, typename P2, int Side>
  DeviceMap& operator=(const DeviceSelfAdjointProduct<View, P2, Side>& p) {
    p.run_into(*this, detail::Kind::Assign);
    return *this;
  }
  template <typename View, typename P2, int Side>
  DeviceMap& operator+=(const DeviceSelfAdjointProduct<View, P2, Side>& p) {
    p.run_into(*this, detail::Kind::Add);
    return *this;
  }
  template <typename View, typename P2, int Side>
  DeviceMap& operator-=(const DeviceSelfAdjointProduct<View, P2, Side>& p) {
    p.run_into(*this, detail::Kind::Sub);
    return *this;
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

// ---- level-3 views on device maps -------------------------------------------
// Each view is a handle (a DeviceMap, a scalar factor, an op) whose
// operations are one call of the C-ABI level-3 routines, on the stream of
// eigen_amd::default_device(); the status lands in its `status`.  They stand
// for Eigen's TriangularView and SelfAdjointView on the same expressions,
// whose evaluation reaches triangular_solve_matrix
// (TriangularSolverMatrix.h:20-330), product_triangular_matrix_matrix
// (TriangularMatrixMatrix.h:84-420), general_matrix_matrix_triangular_product
// (GeneralMatrixMatrixTriangular.h:38-330) and product_selfadjoint_matrix
// (SelfadjointMatrixMatrix.h:323-520) -- the routines that
// TriangularSolverMatrix_BLAS.h, TriangularMatrixMatrix_BLAS.h,
// GeneralMatrixMatrixTriangular_BLAS.h and SelfadjointMatrixMatrix_BLAS.h
// redirect to ?trsm_, ?trmm_, ?syrk_ / ?gemmtr_ and ?symm_ / ?hemm_ under
// EIGEN_USE_BLAS.  Here no device data is touched on the host.
//
// A row-major matrix is handled through its column-major view (the
// transpose): its stored triangle flips and its op transposes.

// op(tri(A)) scaled by a factor: Mode Lower / Upper / UnitLower / UnitUpper
// of the DeviceMap's matrix, Op 0 = as is, 1 = transposed, 2 = adjoint.
template <typename Plain, int Mode, int Op>
class DeviceTriangularView {
 public:
  using Scalar = typename Plain::Scalar;
  using Index = Eigen::Index;
  static constexpr bool kComplex = Eigen::NumTraits<Scalar>::IsComplex;
  static_assert(detail::scalar_info<Scalar>::supported,
                "device views support float, double, std::complex<float> and std::complex<double>");
  static_assert(Mode == Eigen::Lower || Mode == Eigen::Upper || Mode == Eigen::UnitLower || Mode == Eigen::UnitUpper,
                "device triangular views: Lower, Upper, UnitLower or UnitUpper");
  // the triangle as the BLAS routines see it: the column-major matrix in
  // memory, its stored triangle and the op on it
  static constexpr bool kRowMajor = Plain::IsRowMajor;
  static constexpr int kStoredOp = kRowMajor ? detail::op_transpose(Op) : Op;
  static constexpr eh_uplo kStoredUplo = detail::uplo_of(Mode, kRowMajor);
  static constexpr eh_diag kDiag = (Mode & Eigen::UnitDiag) ? EH_UNIT : EH_NON_UNIT;

  explicit DeviceTriangularView(const DeviceMap<Plain>& m, Scalar factor = Scalar(1)) : m_(m), factor_(factor) {}
  const DeviceMap<Plain>& nestedExpression() const { return m_; }
  Scalar factor() const { return factor_; }
  Index rows() const { return m_.rows(); }
  Index cols() const { return m_.cols(); }

  DeviceTriangularView<Plain, Mode, detail::op_transpose(Op)> transpose() const {
    return DeviceTriangularView<Plain, Mode, detail::op_transpose(Op)>(m_, factor_);
  }
  DeviceTriangularView<Plain, Mode, detail::op_adjoint(Op, kComplex)> adjoint() const {
    return DeviceTriangularView<Plain, Mode, detail::op_adjoint(Op, kComplex)>(m_, Eigen::numext::conj(factor_));
  }

  // op(T) X = X_old (OnTheLeft) or X op(T) = X_old (OnTheRight); X
  // overwrites the right-hand sides (eh_trsm).  A factor s solves with s*T.
  template <int Side = Eigen::OnTheLeft, typename P2>
  void solveInPlace(const DeviceMap<P2>& X) const {
    static_assert(std::is_same<typename P2::Scalar, Scalar>::value, "mixed scalar types are not supported");
    static_assert(Side == Eigen::OnTheLeft || Side == Eigen::OnTheRight, "OnTheLeft or OnTheRight");
    // a row-major X is solved through its column-major view X^T:
    // op(T) X = B  <=>  X^T op(T)^T = B^T
    constexpr bool xrow = P2::IsRowMajor;
    constexpr int op = xrow ? detail::op_transpose(kStoredOp) : kStoredOp;
    static_assert(op != 3, "this combination of storage orders and adjoint needs a conjugate without transpose, "
                           "which no BLAS op expresses");
    eigen_assert(rows() == cols());
    eigen_assert((Side == Eigen::OnTheLeft ? X.rows() : X.cols()) == rows());
    const eh_side side = (Side == Eigen::OnTheLeft) != xrow ? EH_LEFT : EH_RIGHT;
    const int64_t m = xrow ? X.cols() : X.rows(), n = xrow ? X.rows() : X.cols();
    const Scalar alpha = Scalar(1) / factor_;
    HipDevice& dev = default_device();
    if (m == 0 || n == 0) return detail::finish(EH_SUCCESS, dev);
    const int64_t lda = m_.outerStride() < 1 ? 1 : m_.outerStride();
    const int64_t ldx = X.outerStride() < 1 ? 1 : X.outerStride();
    void* x = const_cast<Scalar*>(X.data());
    constexpr eh_dtype dt = detail::scalar_info<Scalar>::dtype;
    detail::finish(kComplex ? eh_trsm_complex(dt, side, kStoredUplo, detail::eh_op_of(op), kDiag, m, n, &alpha,
                                              m_.data(), lda, x, ldx, dev.stream)
                            : eh_trsm(dt, side, kStoredUplo, detail::eh_op_of(op), kDiag, m, n, &alpha, m_.data(),
                                      lda, x, ldx, dev.stream),
                   dev);
  }

  // C.triangularView<Mode>() = / += / -= A * B: the product on the Mode
  // triangle of C only (eh_gemmtr); the rest of C is not touched.
  template <typename L, typename R, int O>
  DeviceTriangularView& operator=(const Eigen::Product<L, R, O>& p) {
    return assign_product(p, detail::Kind::Assign);
  }
  template <typename L, typename R, int O>
  DeviceTriangularView& operator+=(const Eigen::Product<L, R, O>& p) {
    return assign_product(p, detail::Kind::Add);
  }
  template <typename L, typename R, int O>
  DeviceTriangularView& operator-=(const Eigen::Product<L, R, O>& p) {
    return assign_product(p, detail::Kind::Sub);
  }

 private:
  template <typename L, typename R, int O>
  DeviceTriangularView& assign_product(const Eigen::Product<L, R, O>& p, detail::Kind kind) {
    static_assert(Op == 0 && (Mode == Eigen::Lower || Mode == Eigen::Upper),
                  "assign to the Lower or Upper triangle of the matrix itself");
    eigen_assert(rows() == cols());
    eigen_assert(factor_ == Scalar(1));
    const typename DeviceMap<Plain>::Base& dst = m_;
    const detail::GemmArgs<Scalar> g = detail::gemm_args(dst, p.lhs(), p.rhs(), kind);
    HipDevice& dev = default_device();
    if (g.m == 0) return detail::finish(EH_SUCCESS, dev), *this;
    const int64_t ldc = m_.outerStride() < 1 ? 1 : m_.outerStride();
    const eh_uplo uplo = detail::uplo_of(Mode, g.flip);
    void* c = const_cast<Scalar*>(m_.data());
    constexpr eh_dtype dt = detail::scalar_info<Scalar>::dtype;
    detail::finish(kComplex ? eh_gemmtr_complex(dt, uplo, g.opa, g.opb, g.m, g.k, &g.alpha, g.A, g.lda, g.B, g.ldb,
                                                &g.beta, c, ldc, dev.stream)
                            : eh_gemmtr(dt, uplo, g.opa, g.opb, g.m, g.k, &g.alpha, g.A, g.lda, g.B, g.ldb, &g.beta,
                                        c, ldc, dev.stream),
                   dev);
    return *this;
  }

  DeviceMap<Plain> m_;
  Scalar factor_;
};

// s * T: the view scaled (for products; a solve divides by it)
template <typename Plain, int Mode, int Op>
DeviceTriangularView<Plain, Mode, Op> operator*(const typename Plain::Scalar& s,
                                                const DeviceTriangularView<Plain, Mode, Op>& t) {
  return DeviceTriangularView<Plain, Mode, Op>(t.nestedExpression(), s * t.factor());
}

// T * X (Side OnTheLeft) or X * T (OnTheRight), to be assigned to X itself.
template <typename View, typename P2, int Side>
struct DeviceTriangularProduct {
  View tri;
  DeviceMap<P2> x;
  template <typename Dst>
  void run_into(Dst& dst) const {
    using Scalar = typename View::Scalar;
    static_assert(std::is_same<typename P2::Scalar, Scalar>::value, "mixed scalar types are not supported");
    static_assert(std::is_same<typename Dst::Scalar, Scalar>::value && bool(Dst::IsRowMajor) == bool(P2::IsRowMajor),
                  "the destination must be the product's other operand");
    // in place, as BLAS trmm: the destination is the other operand
    eigen_assert(dst.data() == x.data() && dst.rows() == x.rows() && dst.cols() == x.cols() &&
                 dst.outerStride() == x.outerStride() &&
                 "X = T * X needs the same X on both sides (copy B into C first for C = T * B)");
    constexpr bool xrow = P2::IsRowMajor;
    constexpr int op = xrow ? detail::op_transpose(View::kStoredOp) : View::kStoredOp;
    static_assert(op != 3, "this combination of storage orders and adjoint needs a conjugate without transpose, "
                           "which no BLAS op expresses");
    eigen_assert(tri.rows() == tri.cols());
    eigen_assert((Side == Eigen::OnTheLeft ? x.rows() : x.cols()) == tri.rows());
    const eh_side side = (Side == Eigen::OnTheLeft) != xrow ? EH_LEFT : EH_RIGHT;
    const int64_t m = xrow ? x.cols() : x.rows(), n = xrow ? x.rows() : x.cols();
    const Scalar alpha = tri.factor();
    HipDevice& dev = default_device();
    if (m == 0 || n == 0) return detail::finish(EH_SUCCESS, dev);
    const auto& A = tri.nestedExpression();
    const int64_t lda = A.outerStride() < 1 ? 1 : A.outerStride();
    const int64_t ldx = x.outerStride() < 1 ? 1 : x.outerStride();
    void* xp = const_cast<Scalar*>(x.data());
    constexpr eh_dtype dt = detail::scalar_info<Scalar>::dtype;
    detail::finish(View::kComplex ? eh_trmm_complex(dt, side, View::kStoredUplo, detail::eh_op_of(op), View::kDiag, m,
                                                    n, &alpha, A.data(), lda, xp, ldx, dev.stream)
                                  : eh_trmm(dt, side, View::kStoredUplo, detail::eh_op_of(op), View::kDiag, m, n,
                                            &alpha, A.data(), lda, xp, ldx, dev.stream),
                   dev);
  }
};

template <typename Plain, int Mode, int Op, typename P2>
DeviceTriangularProduct<DeviceTriangularView<Plain, Mode, Op>, P2, Eigen::OnTheLeft> operator*(
    const DeviceTriangularView<Plain, Mode, Op>& t, const DeviceMap<P2>& x) {
  return {t, x};
}
template <typename Plain, int Mode, int Op, typename P2>
DeviceTriangularProduct<DeviceTriangularView<Plain, Mode, Op>, P2, Eigen::OnTheRight> operator*(
    const DeviceMap<P2>& x, const DeviceTriangularView<Plain, Mode, Op>& t) {
  return {t, x};
}

// The self-adjoint matrix (symmetric for real scalars, Hermitian for
// complex ones) stored in the Mode triangle of the DeviceMap's matrix,
// scaled by a factor (products), or the destination of rank updates.
template <typename Plain, int Mode>
class DeviceSelfAdjointView {
 public:
  using Scalar = typename Plain::Scalar;
  using RealScalar = typename Eigen::NumTraits<Scalar>::Real;
  using Index = Eigen::Index;
  static constexpr bool kComplex = Eigen::NumTraits<Scalar>::IsComplex;
  static_assert(detail::scalar_info<Scalar>::supported,
                "device views support float, double, std::complex<float> and std::complex<double>");
  static_assert(Mode == Eigen::Lower || Mode == Eigen::Upper, "device self-adjoint views: Lower or Upper");
  // A row-major matrix's column-major view is its transpose, self-adjoint
  // too (the conjugate, for a Hermitian one), stored in the other triangle.
  static constexpr bool kRowMajor = Plain::IsRowMajor;
  static constexpr eh_uplo kStoredUplo = detail::uplo_of(Mode, kRowMajor);
  static constexpr int kMode = Mode;

  explicit DeviceSelfAdjointView(const DeviceMap<Plain>& m, Scalar factor = Scalar(1)) : m_(m), factor_(factor) {}
  const DeviceMap<Plain>& nestedExpression() const { return m_; }
  Scalar factor() const { return factor_; }
  Index rows() const { return m_.rows(); }
  Index cols() const { return m_.cols(); }

  // C = C + alpha * U * U^*  on the stored triangle (eh_syrk / eh_herk)
  template <typename U>
  DeviceSelfAdjointView& rankUpdate(const Eigen::MatrixBase<U>& u, const RealScalar& alpha = RealScalar(1)) {
    using Traits = Eigen::internal::blas_traits<U>;
    using Actual = Eigen::internal::remove_all_t<typename Traits::DirectLinearAccessType>;
    static_assert(std::is_same<typename U::Scalar, Scalar>::value, "mixed scalar types are not supported");
    static_assert(bool(Traits::HasUsableDirectAccess), "rank-update operands need direct access with unit inner stride");
    static_assert(!kComplex || !kRowMajor, "a Hermitian rank update of a row-major matrix is not supported");
    using Op = detail::operand_op<detail::row_major<Actual>(), bool(Traits::NeedToConjugate), false>;
    static_assert(Op::valid && (!kComplex || Op::op != EH_OP_T),
                  "the rank-update operand must be U, U^T (real) or U^H (complex) of a stored matrix");
    eigen_assert(rows() == cols() && u.rows() == rows());
    const Actual& a = Traits::extract(u.derived());
    // a scalar factor f of U enters as |f|^2
    const RealScalar f2 = Eigen::numext::abs2(Traits::extractScalarFactor(u.derived()));
    const int64_t n = rows(), k = u.cols();
    const int64_t lda = a.outerStride() < 1 ? 1 : a.outerStride();
    const int64_t ldc = m_.outerStride() < 1 ? 1 : m_.outerStride();
    void* c = const_cast<Scalar*>(m_.data());
    constexpr eh_dtype dt = detail::scalar_info<Scalar>::dtype;
    HipDevice& dev = default_device();
    if (n == 0) return detail::finish(EH_SUCCESS, dev), *this;
    if (kComplex) {
      const RealScalar ar = alpha * f2, br = RealScalar(1);
      detail::finish(eh_herk(dt, kStoredUplo, Op::op, n, k, &ar, a.data(), lda, &br, c, ldc, dev.stream), dev);
    } else {
      const Scalar as = Scalar(alpha * f2), bs = Scalar(1);
      detail::finish(eh_syrk(dt, kStoredUplo, Op::op, n, k, &as, a.data(), lda, &bs, c, ldc, dev.stream), dev);
    }
    return *this;
  }

  // C = C + alpha * U * V^* + conj(alpha) * V * U^*  (eh_syr2k / eh_her2k).
  // For vectors U, V this is Eigen's rank-2 update; n x k factors give the
  // rank-2k update (Eigen's own rankUpdate(u, v) takes vectors only).
  template <typename U, typename V>
  DeviceSelfAdjointView& rankUpdate(const Eigen::MatrixBase<U>& u, const Eigen::MatrixBase<V>& v,
                                    const Scalar& alpha = Scalar(1)) {
    using TU = Eigen::internal::blas_traits<U>;
    using TV = Eigen::internal::blas_traits<V>;
    using AU = Eigen::internal::remove_all_t<typename TU::DirectLinearAccessType>;
    using AV = Eigen::internal::remove_all_t<typename TV::DirectLinearAccessType>;
    static_assert(std::is_same<typename U::Scalar, Scalar>::value && std::is_same<typename V::Scalar, Scalar>::value,
                  "mixed scalar types are not supported");
    static_assert(bool(TU::HasUsableDirectAccess) && bool(TV::HasUsableDirectAccess),
                  "rank-update operands need direct access with unit inner stride");
    static_assert(!kComplex || !kRowMajor, "a Hermitian rank update of a row-major matrix is not supported");
    using OpU = detail::operand_op<detail::row_major<AU>(), bool(TU::NeedToConjugate), false>;
    using OpV = detail::operand_op<detail::row_major<AV>(), bool(TV::NeedToConjugate), false>;
    static_assert(OpU::valid && OpV::valid && OpU::op == OpV::op && (!kComplex || OpU::op != EH_OP_T),
                  "U and V must both be stored matrices, both transposed (real) or both adjoints (complex)");
    eigen_assert(rows() == cols() && u.rows() == rows() && v.rows() == rows() && u.cols() == v.cols());
    const AU& a = TU::extract(u.derived());
    const AV& b = TV::extract(v.derived());
    // scalar factors of U and V fold into alpha: (fu U)(fv V)^* = fu conj(fv) U V^*
    const Scalar al = alpha * TU::extractScalarFactor(u.derived()) *
                      Eigen::numext::conj(TV::extractScalarFactor(v.derived()));
    const int64_t n = rows(), k = u.cols();
    const int64_t lda = a.outerStride() < 1 ? 1 : a.outerStride();
    const int64_t ldb = b.outerStride() < 1 ? 1 : b.outerStride();
    const int64_t ldc = m_.outerStride() < 1 ? 1 : m_.outerStride();
    void* c = const_cast<Scalar*>(m_.data());
    constexpr eh_dtype dt = detail::scalar_info<Scalar>::dtype;
    HipDevice& dev = default_device();
    if (n == 0) return detail::finish(EH_SUCCESS, dev), *this;
    if (kComplex) {
      const RealScalar br = RealScalar(1);
      detail::finish(
          eh_her2k(dt, kStoredUplo, OpU::op, n, k, &al, a.data(), lda, b.data(), ldb, &br, c, ldc, dev.stream), dev);
    } else {
      const Scalar bs = Scalar(1);
      detail::finish(
          eh_syr2k(dt, kStoredUplo, OpU::op, n, k, &al, a.data(), lda, b.data(), ldb, &bs, c, ldc, dev.stream), dev);
    }
    return *this;
  }

 private:
  DeviceMap<Plain> m_;
  Scalar factor_;
};

template <typename Plain, int Mode>
DeviceSelfAdjointView<Plain, Mode> operator*(const typename Plain::Scalar& s, const DeviceSelfAdjointView<Plain, Mode>& a) {
  return DeviceSelfAdjointView<Plain, Mode>(a.nestedExpression(), s * a.factor());
}

// S * B (Side OnTheLeft) or B * S (OnTheRight), assigned (=, +=, -=) to a
// DeviceMap C of the same storage order as B (eh_symm / eh_hemm).
template <typename View, typename P2, int Side>
struct DeviceSelfAdjointProduct {
  View sa;
  DeviceMap<P2> b;
  template <typename Dst>
  void run_into(Dst& dst, detail::Kind kind) const {
    // as plain Maps (Eigen's traits are defined for Map, not for DeviceMap)
    using AMap = typename std::decay_t<decltype(sa.nestedExpression())>::Base;
    const AMap& a = sa.nestedExpression();
    const typename DeviceMap<P2>::Base& bm = b;
    typename Dst::Base& d = dst;
    detail::device_symm<View::kMode>(d, a, bm, Side == Eigen::OnTheLeft, sa.factor(), kind, default_device());
  }
};

template <typename Plain, int Mode, typename P2>
DeviceSelfAdjointProduct<DeviceSelfAdjointView<Plain, Mode>, P2, Eigen::OnTheLeft> operator*(
    const DeviceSelfAdjointView<Plain, Mode>& a, const DeviceMap<P2>& b) {
  return {a, b};
}
template <typename Plain, int Mode, typename P2>
DeviceSelfAdjointProduct<DeviceSelfAdjointView<Plain, Mode>, P2, Eigen::OnTheRight> operator*(
    const DeviceMap<P2>& b, const DeviceSelfAdjointView<Plain, Mode>& a) {
  return {a, b};
}

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
    using LV = eigen_amd::detail::sa_view<remove_all_t<Lhs>>;
    using RV = eigen_amd::detail::sa_view<remove_all_t<Rhs>>;
    const Kind kind = kAssign ? Kind::Assign : (kAdd ? Kind::Add : Kind::Sub);
    if constexpr (LV::value)  // S * B: eh_symm / eh_hemm, left
      eigen_amd::detail::device_symm<LV::mode>(dst, src.lhs().nestedExpression(), src.rhs(), true, Scalar(1), kind,
                                               dev);
    else if constexpr (RV::value)  // B * S: right
      eigen_amd::detail::device_symm<RV::mode>(dst, src.rhs().nestedExpression(), src.lhs(), false, Scalar(1), kind,
                                               dev);
    else
      eigen_amd::detail::device_product(dst, src.lhs(), src.rhs(), kind, dev);
  }
};

}  // namespace internal
}  // namespace Eigen

#endif  // EIGEN_AMD_EIGEN_H
