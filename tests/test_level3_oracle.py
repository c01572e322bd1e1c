"""The oracle's level-3 restatements (gemmtr_ref, syr2k_ref, symm_ref, trmm_ref,
trsm_ref) pinned two ways, on the CPU:

* against Eigen's own BLAS compiled from the reference tree (oracle/_ref,
  blas/level3_impl.h:98-758) on the same inputs, every flag combination,
  padded leading dimensions and NaN in the triangle the routine must not read;
* against identities that do not come from either restatement: TRSM undoes
  TRMM, SYMM with a symmetric matrix is a plain product, the two triangles of
  GEMMTR make the full product, SYR2K is GEMMTR applied twice.
"""
import ctypes as C
import itertools

import numpy as np
import pytest

import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref (Eigen built from the reference tree) absent")

DT = [np.float32, np.float64]


def _tol(dt):
    return 2e-5 if dt == np.float32 else 1e-13


def _rand(rng, shape, dt, ld=None):
    """Column-major array with leading dimension ld (rows past shape[0] hold NaN)."""
    ld = ld or shape[0]
    buf = np.full((ld, max(shape[1], 1)), np.nan, dtype=dt, order="F")
    buf[: shape[0], : shape[1]] = rng.uniform(-1, 1, size=shape)
    return buf


def _tri_input(rng, ka, dt, uplo, ld):
    """A with a well-conditioned uplo triangle and NaN in the other one."""
    A = _rand(rng, (ka, ka), dt, ld)
    lower = uplo == "L"
    for j in range(ka):
        for i in range(ka):
            if (lower and i < j) or (not lower and i > j):
                A[i, j] = np.nan
        A[j, j] = (2.0 + rng.uniform(0, 1)) * (1 if j % 2 else -1)
    return A


def _i(v):
    return C.byref(C.c_int(v))


def _s(v, dt):
    return C.byref((C.c_float if dt == np.float32 else C.c_double)(v))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _rel(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    den = max(np.linalg.norm(ref), 1e-300)
    return np.linalg.norm(x - ref) / den


def _pre(dt):
    return "s" if dt == np.float32 else "d"


@needs_ref
@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,ta,tb", list(itertools.product("UL", "NT", "NT")))
def test_gemmtr_vs_eigen_blas(dt, uplo, ta, tb):
    L = O.blas_ref()
    rng = np.random.default_rng(1)
    for n, k, alpha, beta in ((7, 5, 0.7, 1.3), (13, 1, 1.0, 0.0), (4, 9, -0.5, 1.0), (9, 0, 0.7, 1.3)):
        A = _rand(rng, (n, k) if ta == "N" else (k, n), dt, (n if ta == "N" else k) + 2)
        B = _rand(rng, (k, n) if tb == "N" else (n, k), dt, (k if tb == "N" else n) + 1)
        Cm = _rand(rng, (n, n), dt, n + 3)
        want = Cm.copy(order="F")
        getattr(L, _pre(dt) + "gemmtr_")(uplo.encode(), ta.encode(), tb.encode(), _i(n), _i(k), _s(alpha, dt), _p(A),
                                         _i(A.shape[0]), _p(B), _i(B.shape[0]), _s(beta, dt), _p(want),
                                         _i(want.shape[0]))
        sa = (n, k) if ta == "N" else (k, n)
        sb = (k, n) if tb == "N" else (n, k)
        got = O.gemmtr_ref(uplo, ta, tb, alpha, A[: sa[0], : sa[1]], B[: sb[0], : sb[1]], beta, Cm[:n])
        assert _rel(got, want[:n]) < _tol(dt)


@needs_ref
@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,trans", list(itertools.product("UL", "NT")))
def test_syr2k_vs_eigen_blas(dt, uplo, trans):
    L = O.blas_ref()
    rng = np.random.default_rng(2)
    for n, k, alpha, beta in ((7, 5, 0.7, 1.3), (12, 3, 1.0, 0.0), (5, 11, -0.5, 1.0)):
        shape = (n, k) if trans == "N" else (k, n)
        A = _rand(rng, shape, dt, shape[0] + 2)
        B = _rand(rng, shape, dt, shape[0] + 1)
        Cm = _rand(rng, (n, n), dt, n + 2)
        # the other triangle of C is never read: NaN there must survive untouched
        for j in range(n):
            for i in range(n):
                if (uplo == "L" and i < j) or (uplo == "U" and i > j):
                    Cm[i, j] = np.nan
        want = Cm.copy(order="F")
        getattr(L, _pre(dt) + "syr2k_")(uplo.encode(), trans.encode(), _i(n), _i(k), _s(alpha, dt), _p(A),
                                        _i(A.shape[0]), _p(B), _i(B.shape[0]), _s(beta, dt), _p(want),
                                        _i(want.shape[0]))
        got = O.syr2k_ref(uplo, trans, alpha, A[: shape[0]], B[: shape[0]], beta, Cm[:n])
        mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
        assert _rel(got[mask], want[:n][mask]) < _tol(dt)
        assert np.isnan(want[:n][~mask]).all() and np.isnan(got[~mask]).all()


@needs_ref
@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("side,uplo", list(itertools.product("LR", "UL")))
def test_symm_vs_eigen_blas(dt, side, uplo):
    L = O.blas_ref()
    rng = np.random.default_rng(3)
    for m, n, alpha, beta in ((7, 5, 0.7, 1.3), (3, 12, 1.0, 0.0), (10, 10, -0.5, 1.0)):
        ka = m if side == "L" else n
        A = _tri_input(rng, ka, dt, uplo, ka + 2)
        B = _rand(rng, (m, n), dt, m + 1)
        Cm = _rand(rng, (m, n), dt, m + 3)
        want = Cm.copy(order="F")
        getattr(L, _pre(dt) + "symm_")(side.encode(), uplo.encode(), _i(m), _i(n), _s(alpha, dt), _p(A),
                                       _i(A.shape[0]), _p(B), _i(B.shape[0]), _s(beta, dt), _p(want),
                                       _i(want.shape[0]))
        got = O.symm_ref(side, uplo, alpha, A, B, beta, Cm[:m])
        assert _rel(got, want[:m]) < _tol(dt)


TRI_CASES = list(itertools.product("LR", "UL", "NT", "NU"))


@needs_ref
@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("side,uplo,trans,diag", TRI_CASES)
def test_trmm_trsm_vs_eigen_blas(dt, side, uplo, trans, diag):
    L = O.blas_ref()
    rng = np.random.default_rng(4)
    for m, n, alpha in ((7, 5, 0.7), (1, 9, 1.0), (12, 3, -1.5)):
        ka = m if side == "L" else n
        A = _tri_input(rng, ka, dt, uplo, ka + 2)
        B = _rand(rng, (m, n), dt, m + 2)
        for name, ref in (("trmm_", O.trmm_ref), ("trsm_", O.trsm_ref)):
            want = B.copy(order="F")
            getattr(L, _pre(dt) + name)(side.encode(), uplo.encode(), trans.encode(), diag.encode(), _i(m), _i(n),
                                        _s(alpha, dt), _p(A), _i(A.shape[0]), _p(want), _i(want.shape[0]))
            got = ref(side, uplo, trans, diag, alpha, A, B[:m])
            assert _rel(got, want[:m]) < 10 * _tol(dt), name
            assert np.isnan(want[m:]).all()  # rows past m (the ld padding) untouched


# ---- identities (no reference code involved) --------------------------------
@pytest.mark.parametrize("side,uplo,trans,diag", TRI_CASES)
def test_trsm_undoes_trmm(side, uplo, trans, diag):
    rng = np.random.default_rng(5)
    m, n = 9, 6
    ka = m if side == "L" else n
    A = _tri_input(rng, ka, np.float64, uplo, ka)
    B = rng.uniform(-1, 1, (m, n))
    Y = O.trmm_ref(side, uplo, trans, diag, 2.0, A, B)
    X = O.trsm_ref(side, uplo, trans, diag, 0.5, A, Y)
    assert np.abs(X - B).max() < 1e-12
    # and the triangle really is used: a different triangle gives a different product
    other = "U" if uplo == "L" else "L"
    A2 = _tri_input(rng, ka, np.float64, other, ka)
    if ka > 1:
        assert np.abs(O.trmm_ref(side, other, trans, diag, 2.0, A2, B) - Y).max() > 1e-3


@pytest.mark.parametrize("side", "LR")
def test_symm_is_plain_product_for_symmetric_input(side):
    rng = np.random.default_rng(6)
    m, n = 8, 5
    ka = m if side == "L" else n
    S = rng.uniform(-1, 1, (ka, ka))
    S = S + S.T
    B = rng.uniform(-1, 1, (m, n))
    C0 = rng.uniform(-1, 1, (m, n))
    want = 0.5 * (S @ B if side == "L" else B @ S) + 2.0 * C0
    for uplo in "UL":
        A = np.asfortranarray(S.copy())
        A[np.triu_indices(ka, 1) if uplo == "L" else np.tril_indices(ka, -1)] = np.nan
        assert np.abs(O.symm_ref(side, uplo, 0.5, A, B, 2.0, C0) - want).max() < 1e-12


def test_gemmtr_triangles_make_the_product_and_syr2k_is_two_of_them():
    rng = np.random.default_rng(7)
    n, k = 7, 4
    A = rng.uniform(-1, 1, (n, k))
    B = rng.uniform(-1, 1, (n, k))
    C0 = rng.uniform(-1, 1, (n, n))
    full = 0.7 * A @ B.T + 1.3 * C0
    lo = O.gemmtr_ref("L", "N", "T", 0.7, A, B, 1.3, C0)
    up = O.gemmtr_ref("U", "N", "T", 0.7, A, B, 1.3, C0)
    assert np.abs(np.tril(lo) + np.triu(up, 1) - full).max() < 1e-12
    assert np.abs(np.triu(lo, 1) - np.triu(C0, 1)).max() == 0  # other triangle unchanged
    for uplo in "UL":
        two = O.gemmtr_ref(uplo, "N", "T", 0.7, B, A, 1.0, O.gemmtr_ref(uplo, "N", "T", 0.7, A, B, 1.3, C0))
        assert np.abs(O.syr2k_ref(uplo, "N", 0.7, A, B, 1.3, C0) - two).max() < 1e-12
    # beta == 0 ignores NaN in C (BLAS: C is not read)
    Cn = np.full((n, n), np.nan)
    assert np.isfinite(np.tril(O.gemmtr_ref("L", "N", "T", 1.0, A, B, 0.0, Cn))[np.tril_indices(n)]).all()


@needs_ref
@pytest.mark.parametrize("s", [0.36, 0.45, 0.5])
@pytest.mark.parametrize("side,trans", [("L", "N"), ("R", "T")])
def test_trsm_ref_ill_conditioned_matches_eigen_dtrsm(s, side, trans):
    """Unit lower triangles of order 300 with off-diagonal U[-s, s] (condition
    up to ~1e5-1e7 at this order): the oracle's substitution and Eigen's dtrsm_
    agree far below the conditioning, as two backward-stable substitutions do."""
    rng = np.random.default_rng(9)
    n, r = 300, 8
    T = np.tril(rng.uniform(-1, 1, (n, n)) * s, -1)
    np.fill_diagonal(T, 1.0)
    T = np.asfortranarray(T)
    m, nn = (n, r) if side == "L" else (r, n)
    B = np.asfortranarray(rng.uniform(-1, 1, (m, nn)))
    want = B.copy(order="F")
    O.blas_ref().dtrsm_(side.encode(), b"L", trans.encode(), b"U", _i(m), _i(nn), _s(1.0, np.float64), _p(T), _i(n),
                        _p(want), _i(m))
    got = O.trsm_ref(side, "L", trans, "U", 1.0, T, B)
    assert _rel(got, want) < 1e-12 * np.linalg.cond(T) ** 0.5
