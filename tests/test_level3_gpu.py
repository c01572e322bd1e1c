"""Level-3 routines on the device (eh_gemmtr, eh_syr2k, eh_symm, eh_trmm,
eh_trsm) against the oracle's restatements, which the CPU tests pin to
Eigen's BLAS.

Bounds (u = unit roundoff of the dtype, written in each test):
* products (GEMMTR, SYR2K, SYMM, TRMM): normwise ||X - X_ref||_F <=
  c * ka * u * || |op(A)| |B| ||_F, the GEMM bound with ka the inner order;
* TRSM: the backward error ||op(T) X - alpha B||_F <= c * ka * u *
  ||T||_F ||X||_F (independent of conditioning), and the forward error
  against the oracle on well-conditioned triangles (unit-scaled diagonal,
  off-diagonal entries O(1/ka)).
Inputs carry NaN wherever the routine must not read: the other triangle of
A, its diagonal when diag = UNIT, the padding rows of every leading
dimension, the other triangle of C for GEMMTR / SYR2K (which must also stay
NaN)."""
import itertools

import numpy as np
import pytest

import eigen_amd as E
import oracle as O
from helpers import empty_dev, from_dev, to_dev

pytestmark = pytest.mark.gpu

DT = [np.float32, np.float64]
U = {np.float32: 2.0 ** -24, np.float64: 2.0 ** -53}
ED = {np.float32: E.F32, np.float64: E.F64}
SIDE = {"L": E.LEFT, "R": E.RIGHT}
UPLO = {"U": E.UPPER, "L": E.LOWER}
OP = {"N": E.OP_N, "T": E.OP_T}
DIAG = {"N": E.NON_UNIT, "U": E.UNIT}


def _padded(rng, rows, cols, dt, ld, lo=-1.0, hi=1.0):
    a = np.full((ld, max(cols, 1)), np.nan, dtype=dt, order="F")
    a[:rows, :cols] = rng.uniform(lo, hi, (rows, cols))
    return a


def _triangle(rng, ka, dt, uplo, diag, ld):
    """Well-conditioned uplo triangle (diagonal 1..2 in magnitude, off-diagonal
    O(1/ka)); NaN in the other triangle, and on the diagonal when unit."""
    A = _padded(rng, ka, ka, dt, ld, -1.0 / ka, 1.0 / ka)
    i, j = np.indices((ka, ka))
    other = (i < j) if uplo == "L" else (i > j)
    A[:ka, :ka][other] = np.nan
    d = rng.uniform(1.0, 2.0, ka) * np.where(np.arange(ka) % 2, 1.0, -1.0)
    A[np.arange(ka), np.arange(ka)] = np.nan if diag == "U" else d
    return A


def _dense_tri(A, ka, uplo, diag):
    T = np.nan_to_num(np.asarray(A[:ka, :ka], np.float64))
    T = np.tril(T) if uplo == "L" else np.triu(T)
    if diag == "U":
        np.fill_diagonal(T, 1.0)
    return T


def _dev(a):
    return to_dev(a)


def _back(t, a):
    return from_dev(t, a.shape)


# ---- TRMM / TRSM ---------------------------------------------------------------
TRI = list(itertools.product("LR", "UL", "NT", "NU"))
TRI_SIZES = [(1, 1), (31, 7), (33, 40), (70, 129), (300, 65)]


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("side,uplo,trans,diag", TRI)
def test_trmm_matches_oracle(gpu, dt, side, uplo, trans, diag):
    rng = np.random.default_rng(11)
    for m, n in TRI_SIZES + ([(65, 290)] if side == "R" else []):
        ka = m if side == "L" else n
        A = _triangle(rng, ka, dt, uplo, diag, ka + 3)
        B = _padded(rng, m, n, dt, m + 2)
        alpha = -0.75
        dB = _dev(B)
        E.trmm(ED[dt], SIDE[side], UPLO[uplo], OP[trans], DIAG[diag], m, n, alpha, _dev(A), A.shape[0], dB,
               B.shape[0])
        got = _back(dB, B)
        assert np.isnan(got[m:]).all(), "padding rows of B written"
        want = O.trmm_ref(side, uplo, trans, diag, alpha, np.nan_to_num(A), B[:m])
        T = _dense_tri(A, ka, uplo, diag)
        T = T.T if trans == "T" else T
        absprod = np.abs(alpha) * (np.abs(T) @ np.abs(B[:m].astype(np.float64)) if side == "L"
                                   else np.abs(B[:m].astype(np.float64)) @ np.abs(T))
        bound = 4 * ka * U[dt] * np.linalg.norm(absprod)  # c = 4
        err = np.linalg.norm(got[:m].astype(np.float64) - want)
        assert err <= bound, (m, n, err, bound)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("side,uplo,trans,diag", TRI)
def test_trsm_matches_oracle(gpu, dt, side, uplo, trans, diag):
    rng = np.random.default_rng(12)
    for m, n in TRI_SIZES + ([(40, 300)] if side == "R" else []):
        ka = m if side == "L" else n
        A = _triangle(rng, ka, dt, uplo, diag, ka + 1)
        B = _padded(rng, m, n, dt, m + 3)
        alpha = 1.5
        dB = _dev(B)
        E.trsm(ED[dt], SIDE[side], UPLO[uplo], OP[trans], DIAG[diag], m, n, alpha, _dev(A), A.shape[0], dB,
               B.shape[0])
        got = _back(dB, B)
        assert np.isnan(got[m:]).all(), "padding rows of B written"
        X = got[:m].astype(np.float64)
        assert np.isfinite(X).all()
        T = _dense_tri(A, ka, uplo, diag)
        T = T.T if trans == "T" else T
        # backward error: c = 8
        resid = (T @ X if side == "L" else X @ T) - alpha * B[:m].astype(np.float64)
        assert np.linalg.norm(resid) <= 8 * ka * U[dt] * np.linalg.norm(T) * np.linalg.norm(X), (m, n)
        # forward error on these well-conditioned triangles: c = 8
        want = O.trsm_ref(side, uplo, trans, diag, alpha, np.nan_to_num(A), B[:m])
        assert np.linalg.norm(X - want) <= 8 * ka * U[dt] * np.linalg.norm(want), (m, n)


@pytest.mark.parametrize("dt", DT)
def test_alpha_zero_clears_b_without_reading_a(gpu, dt):
    rng = np.random.default_rng(13)
    m, n = 40, 9
    B = _padded(rng, m, n, dt, m)
    A = np.full((m, m), np.nan, dtype=dt, order="F")
    for f in (E.trmm, E.trsm):
        dB = _dev(B)
        f(ED[dt], E.LEFT, E.LOWER, E.OP_N, E.NON_UNIT, m, n, 0.0, _dev(A), m, dB, m)
        assert (_back(dB, B) == 0).all()


@pytest.mark.parametrize("side", "LR")
def test_trsm_inverts_trmm_large(gpu, side):
    """4096-order FP64 triangle: TRMM then TRSM returns B (size-independent
    round trip; both recursions run many levels of MFMA products)."""
    rng = np.random.default_rng(14)
    ka, other = 4096, 384
    m, n = (ka, other) if side == "L" else (other, ka)
    A = _triangle(rng, ka, np.float64, "L", "N", ka)
    B = rng.uniform(-1, 1, (m, n))
    dA, dB = _dev(A), _dev(np.asfortranarray(B))
    E.trmm(E.F64, SIDE[side], E.LOWER, E.OP_T, E.NON_UNIT, m, n, 2.0, dA, ka, dB, m)
    E.trsm(E.F64, SIDE[side], E.LOWER, E.OP_T, E.NON_UNIT, m, n, 0.5, dA, ka, dB, m)
    got = from_dev(dB, (m, n))
    assert np.linalg.norm(got - B) <= 64 * ka * U[np.float64] * np.linalg.norm(B)


# ---- SYMM ------------------------------------------------------------------------
@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("side,uplo", list(itertools.product("LR", "UL")))
def test_symm_matches_oracle(gpu, dt, side, uplo):
    rng = np.random.default_rng(15)
    for m, n, beta in ((1, 1, 0.5), (37, 65, 0.0), (130, 70, 1.25), (70, 300, -1.0)):
        ka = m if side == "L" else n
        A = _padded(rng, ka, ka, dt, ka + 2)
        i, j = np.indices((ka, ka))
        A[:ka, :ka][(i < j) if uplo == "L" else (i > j)] = np.nan
        B = _padded(rng, m, n, dt, m + 1)
        C0 = _padded(rng, m, n, dt, m + 2)
        if beta == 0.0:
            C0[:m, :n] = np.nan  # beta == 0: C is not read
        alpha = 0.625
        dC = _dev(C0)
        E.symm(ED[dt], SIDE[side], UPLO[uplo], m, n, alpha, _dev(A), A.shape[0], _dev(B), B.shape[0], beta, dC,
               C0.shape[0])
        got = _back(dC, C0)
        assert np.isnan(got[m:]).all()
        Cin = np.nan_to_num(C0[:m])
        want = O.symm_ref(side, uplo, alpha, np.nan_to_num(A), B[:m], beta, Cin)
        S = O._sym_dense(np.nan_to_num(A), ka, uplo == "L")
        absprod = abs(alpha) * (np.abs(S) @ np.abs(B[:m].astype(np.float64)) if side == "L"
                                else np.abs(B[:m].astype(np.float64)) @ np.abs(S)) + abs(beta) * np.abs(Cin)
        err = np.linalg.norm(got[:m].astype(np.float64) - want)
        assert err <= 4 * ka * U[dt] * np.linalg.norm(absprod), (m, n, err)


# ---- GEMMTR / SYR2K --------------------------------------------------------------
def _c_with_nan_outside(rng, n, dt, uplo, ld):
    Cm = _padded(rng, n, n, dt, ld)
    i, j = np.indices((n, n))
    Cm[:n, :n][(i < j) if uplo == "L" else (i > j)] = np.nan
    return Cm


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,ta,tb", list(itertools.product("UL", "NT", "NT")))
def test_gemmtr_matches_oracle(gpu, dt, uplo, ta, tb):
    rng = np.random.default_rng(16)
    for n, k, beta in ((1, 1, 1.0), (45, 17, 0.0), (200, 64, 0.75), (300, 300, -1.0)):
        sa = (n, k) if ta == "N" else (k, n)
        sb = (k, n) if tb == "N" else (n, k)
        A = _padded(rng, *sa, dt, sa[0] + 1)
        B = _padded(rng, *sb, dt, sb[0] + 2)
        C0 = _c_with_nan_outside(rng, n, dt, uplo, n + 1)
        alpha = 1.375
        dC = _dev(C0)
        E.gemmtr(ED[dt], UPLO[uplo], OP[ta], OP[tb], n, k, alpha, _dev(A), A.shape[0], _dev(B), B.shape[0], beta, dC,
                 C0.shape[0])
        got = _back(dC, C0)[:n].astype(np.float64)
        mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
        assert np.isnan(got[~mask]).all(), "other triangle written"
        want = O.gemmtr_ref(uplo, ta, tb, alpha, A[: sa[0], : sa[1]], B[: sb[0], : sb[1]], beta, C0[:n])
        opa = A[: sa[0], : sa[1]].astype(np.float64)
        opb = B[: sb[0], : sb[1]].astype(np.float64)
        opa, opb = (opa.T if ta == "T" else opa), (opb.T if tb == "T" else opb)
        absprod = abs(alpha) * np.abs(opa) @ np.abs(opb) + abs(beta) * np.nan_to_num(np.abs(C0[:n]))
        err = np.linalg.norm(got[mask] - want[mask])
        assert err <= 4 * max(k, 1) * U[dt] * np.linalg.norm(absprod[mask]), (n, k, err)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,trans", list(itertools.product("UL", "NT")))
def test_syr2k_matches_oracle(gpu, dt, uplo, trans):
    rng = np.random.default_rng(17)
    for n, k, beta in ((1, 3, 0.5), (50, 20, 0.0), (257, 129, 1.0)):
        shape = (n, k) if trans == "N" else (k, n)
        A = _padded(rng, *shape, dt, shape[0] + 1)
        B = _padded(rng, *shape, dt, shape[0] + 3)
        C0 = _c_with_nan_outside(rng, n, dt, uplo, n)
        alpha = -0.5
        dC = _dev(C0)
        E.syr2k(ED[dt], UPLO[uplo], OP[trans], n, k, alpha, _dev(A), A.shape[0], _dev(B), B.shape[0], beta, dC,
                C0.shape[0])
        got = _back(dC, C0)[:n].astype(np.float64)
        mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
        assert np.isnan(got[~mask]).all(), "other triangle written"
        want = O.syr2k_ref(uplo, trans, alpha, A[: shape[0]], B[: shape[0]], beta, C0[:n])
        a64 = np.abs(A[: shape[0], : shape[1]].astype(np.float64))
        b64 = np.abs(B[: shape[0], : shape[1]].astype(np.float64))
        absprod = (abs(alpha) * (a64 @ b64.T + b64 @ a64.T) if trans == "N" else abs(alpha) * (a64.T @ b64 + b64.T @ a64))
        absprod += abs(beta) * np.nan_to_num(np.abs(C0[:n]))
        err = np.linalg.norm(got[mask] - want[mask])
        assert err <= 8 * k * U[dt] * np.linalg.norm(absprod[mask]), (n, k, err)
