"""Compile-time guards of the Eigen drop-in (include/eigen_amd_eigen.h).

Expressions that Eigen would evaluate on the host over device memory must not
compile: assigning a non-product expression to an eigen_amd::DeviceMap, a
conjugated device operand that has no BLAS op (conj(A) without a transpose),
and level-3 view forms no BLAS routine expresses (a conjugate alone, a
Hermitian rank update of a row-major matrix, mixed storage orders in a
self-adjoint product, a triangle of a transposed view as destination).
Each snippet is compiled with g++ -fsyntax-only against the Eigen headers of
the reference tree; the failing ones must fail with the header's own message,
and a control snippet with the supported forms must compile.  Skipped when
the Eigen headers are absent."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EIGEN = "/root/reference"

HEAD = """
#include <complex>
#include "eigen_amd_eigen.h"
using M = Eigen::MatrixXf;
using CM = Eigen::MatrixXcf;
using RM = Eigen::Matrix<float, Eigen::Dynamic, Eigen::Dynamic, Eigen::RowMajor>;
using RCM = Eigen::Matrix<std::complex<float>, Eigen::Dynamic, Eigen::Dynamic, Eigen::RowMajor>;
void f(float* a, float* b, float* c, std::complex<float>* x, std::complex<float>* y, std::complex<float>* z,
       float* a_, float* b_, float* c_) {
  eigen_amd::DeviceMap<M> A(a, 4, 4), B(b, 4, 4), C(c, 4, 4);
  eigen_amd::DeviceMap<CM> X(x, 4, 4), Y(y, 4, 4), Z(z, 4, 4);
  eigen_amd::DeviceMap<RM> R(c, 4, 4);
  eigen_amd::DeviceMap<RCM> RX(x, 4, 4);
  (void)A; (void)B; (void)C; (void)X; (void)Y; (void)Z; (void)R; (void)RX;
"""


def _compile(body: str):
    src = HEAD + body + "\n}\n"
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + EIGEN, "-I" + os.path.join(ROOT, "include"),
                        "-x", "c++", "-"], input=src, capture_output=True, text=True, timeout=300)
    return r.returncode, r.stderr


pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(EIGEN, "Eigen")), reason="Eigen headers absent")


def test_supported_forms_compile():
    rc, err = _compile("""
  C = A * B; C.noalias() = A * B; C += 2.0f * A.transpose() * B; C -= A * B.transpose();
  Z = X.adjoint() * Y; Z.noalias() -= std::complex<float>(1, 2) * X * Y.adjoint();
  A.triangularView<Eigen::Lower>().solveInPlace(B); A.triangularView<Eigen::UnitUpper>().solveInPlace<Eigen::OnTheRight>(R);
  B = A.triangularView<Eigen::Lower>().transpose() * B; B = B * (2.0f * A.triangularView<Eigen::Upper>());
  C.triangularView<Eigen::Lower>() += A * B.transpose();
  C = A.selfadjointView<Eigen::Lower>() * B; C -= B * A.selfadjointView<Eigen::Upper>();
  C.selfadjointView<Eigen::Lower>().rankUpdate(A, 2.0f); C.selfadjointView<Eigen::Upper>().rankUpdate(A.transpose(), B.transpose());
  R.selfadjointView<Eigen::Lower>().rankUpdate(A);
  X.triangularView<Eigen::Upper>().adjoint().solveInPlace(Y); Z.selfadjointView<Eigen::Lower>().rankUpdate(X.adjoint(), 1.0f);
  Z.selfadjointView<Eigen::Upper>().rankUpdate(X, Y, std::complex<float>(0, 1)); Z = X.selfadjointView<Eigen::Lower>() * Y;
  eigen_amd::HipDevice dev; Eigen::Map<M> ma(a_, 4, 4), mb(b_, 4, 4), mc(c_, 4, 4);
  mc.device(dev).noalias() += ma.selfadjointView<Eigen::Lower>() * mb; mc.device(dev).noalias() = 2.0f * mb * ma.selfadjointView<Eigen::Upper>();
""")
    assert rc == 0, err[-3000:]


@pytest.mark.parametrize("body,message", [
    ("C = A + B;", "a DeviceMap can only be assigned a product"),
    ("C += A.cwiseProduct(B);", "a DeviceMap can only be updated by a product"),
    ("Z = X.conjugate() * Y;", "a conjugated device operand must enter the product transposed"),
    ("C = A;", "deleted function"),
    ("X.triangularView<Eigen::Lower>().adjoint().solveInPlace(RX);", "conjugate without transpose"),
    ("RX.selfadjointView<Eigen::Lower>().rankUpdate(X);", "Hermitian rank update of a row-major matrix"),
    ("Z.selfadjointView<Eigen::Lower>().rankUpdate(X.transpose());", "U^H (complex)"),
    ("C = R.selfadjointView<Eigen::Lower>() * B;", "must share one storage order"),
    ("C.triangularView<Eigen::Lower>().transpose() = A * B;", "assign to the Lower or Upper triangle"),
    ("eigen_amd::HipDevice d; Eigen::Map<M> ma(a_, 4, 4), mb(b_, 4, 4), mc(c_, 4, 4);"
     " mc.device(d).noalias() = ma.selfadjointView<Eigen::Lower>() * mb.transpose();",
     "must share one storage order"),
])
def test_unsafe_forms_rejected(body, message):
    rc, err = _compile(body)
    assert rc != 0
    assert message in err, err[-3000:]
