"""TRSM on ill-conditioned triangles (runs on an MI355X).

The device solve multiplies each block of right-hand sides by the explicit
inverse of its 64 x 64 diagonal block (computed by substitution in one
batched launch); Eigen's dtrsm_ solves by substitution
(TriangularSolverMatrix.h:67).  On unit lower triangles of order 512 with
off-diagonal entries U[-s, s], s chosen for condition numbers 1e6 to 1e10 --
where an inverse-based solve could lose accuracy -- both solvers run on the
same inputs (Eigen's from oracle/_ref), and their errors are measured against
a long-double substitution:
* backward error ||T X - B||_F / (||T||_F ||X||_F + ||B||_F) of the device
  within 10x Eigen's (or below the unit roundoff: backward stable);
* forward error ||X - X_true||_F / ||X_true||_F within 10x Eigen's.
(A CPU emulation of the same blocking measured the inverse-based solve within
5x in backward and 2x in forward error of substitution up to condition 1e18.)
All four side / transpose arrangements of the lower triangle are covered."""
import ctypes as C

import numpy as np
import pytest

import eigen_amd as E
import oracle as O
from helpers import from_dev, to_dev

pytestmark = pytest.mark.gpu

U64 = 2.0 ** -53


def _ld_solve(T, B, left, trans):
    """X with op(T) X = B (left) or X op(T) = B, by substitution in long double."""
    Op = np.asarray(T, np.longdouble)
    Op = Op.T if trans == "T" else Op
    Bt = np.asarray(B, np.longdouble)
    if not left:  # X op(T) = B  <=>  op(T)^T X^T = B^T
        Op, Bt = Op.T, Bt.T
    n = Op.shape[0]
    lower = np.allclose(np.triu(np.nan_to_num(np.asarray(Op, np.float64)), 1), 0)
    X = np.zeros_like(Bt)
    for i in (range(n) if lower else range(n - 1, -1, -1)):
        acc = Bt[i] - (Op[i, :i] @ X[:i] if lower else Op[i, i + 1:] @ X[i + 1:])
        X[i] = acc / Op[i, i]
    return X if left else X.T


def _errors(T, X, B, Xt, left, trans):
    Op = np.asarray(T, np.longdouble)
    Op = Op.T if trans == "T" else Op
    Xl = np.asarray(X, np.longdouble)
    R = (Op @ Xl if left else Xl @ Op) - np.asarray(B, np.longdouble)
    bwd = float(np.linalg.norm(np.asarray(R, np.float64))) / (
        np.linalg.norm(T) * np.linalg.norm(X) + np.linalg.norm(B))
    fwd = float(np.linalg.norm(np.asarray(Xl - Xt, np.float64))) / float(np.linalg.norm(np.asarray(Xt, np.float64)))
    return bwd, fwd


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref (Eigen built from the reference tree) absent")
@pytest.mark.parametrize("s,lo,hi", [(0.36, 1e5, 1e7), (0.45, 1e7, 1e9), (0.5, 1e9, 1e11)])
@pytest.mark.parametrize("side,trans", [("L", "N"), ("L", "T"), ("R", "N"), ("R", "T")])
def test_trsm_ill_conditioned_vs_eigen_dtrsm(gpu, s, lo, hi, side, trans):
    rng = np.random.default_rng(5)
    n, r = 512, 32
    T = np.tril(rng.uniform(-1, 1, (n, n)) * s, -1)
    np.fill_diagonal(T, 1.0)
    T = np.asfortranarray(T)
    cond = np.linalg.cond(T)
    assert lo <= cond <= hi, cond
    left = side == "L"
    m, nn = (n, r) if left else (r, n)
    B = np.asfortranarray(rng.uniform(-1, 1, (m, nn)))
    Xt = _ld_solve(T, B, left, trans)
    # Eigen's dtrsm_ (substitution) on the host
    L = O.blas_ref()
    Xe = B.copy(order="F")
    i_ = lambda v: C.byref(C.c_int(v))
    L.dtrsm_(side.encode(), b"L", trans.encode(), b"U", i_(m), i_(nn), C.byref(C.c_double(1.0)),
             T.ctypes.data_as(C.c_void_p), i_(n), Xe.ctypes.data_as(C.c_void_p), i_(m))
    # the device solve
    dT, dB = to_dev(T), to_dev(B)
    E.trsm(E.F64, E.LEFT if left else E.RIGHT, E.LOWER, E.OP_T if trans == "T" else E.OP_N, E.UNIT, m, nn, 1.0, dT,
           n, dB, m)
    Xg = from_dev(dB, (m, nn))
    bwd_e, fwd_e = _errors(T, Xe, B, Xt, left, trans)
    bwd_g, fwd_g = _errors(T, Xg, B, Xt, left, trans)
    print(f"cond {cond:.1e} {side}{trans}: device bwd {bwd_g:.2e} fwd {fwd_g:.2e}; Eigen bwd {bwd_e:.2e} fwd {fwd_e:.2e}")
    assert bwd_g <= max(10 * bwd_e, U64), (bwd_g, bwd_e)
    assert fwd_g <= 10 * fwd_e, (fwd_g, fwd_e)
