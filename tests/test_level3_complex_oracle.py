"""The oracle's complex level-3 restatements (trsm_ref, trmm_ref, symm_ref,
hemm_ref, syrk_ref, herk_ref, syr2k_ref, her2k_ref, gemmtr_ref on complex
inputs) pinned on the CPU:

* against Eigen's own complex BLAS compiled from the reference tree
  (oracle/_ref: blas/complex_single.cpp, complex_double.cpp instantiating
  blas/level3_impl.h) on the same inputs, every flag combination, complex
  alpha and beta (real where the routine takes real ones), padded leading
  dimensions and NaN in the triangle the routine must not read;
* against identities that come from neither restatement: TRSM undoes TRMM
  with conjugate transposes, HEMM of a Hermitian matrix is the plain
  product, HERK equals GEMMTR with B = A^H, HER2K is two GEMMTRs, and the
  imaginary part of a Hermitian diagonal is never read.
"""
import ctypes as C
import itertools

import numpy as np
import pytest

import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref (Eigen built from the reference tree) absent")

DT = [np.complex64, np.complex128]


def _tol(dt):
    return 3e-5 if dt == np.complex64 else 1e-13


def _rand(rng, shape, dt, ld=None):
    """Column-major complex array with leading dimension ld (rows past shape[0] hold NaN)."""
    ld = ld or shape[0]
    buf = np.full((ld, max(shape[1], 1)), np.nan + 1j * np.nan, dtype=dt, order="F")
    buf[: shape[0], : shape[1]] = rng.uniform(-1, 1, size=shape) + 1j * rng.uniform(-1, 1, size=shape)
    return buf


def _tri_input(rng, ka, dt, uplo, ld):
    """A with a well-conditioned uplo triangle and NaN in the other one."""
    A = _rand(rng, (ka, ka), dt, ld)
    lower = uplo == "L"
    for j in range(ka):
        for i in range(ka):
            if (lower and i < j) or (not lower and i > j):
                A[i, j] = np.nan
        A[j, j] = (2.0 + rng.uniform(0, 1)) * (1 if j % 2 else -1) + 0.5j
    return A


def _i(v):
    return C.byref(C.c_int(v))


def _c(v, dt):
    """Pointer to one complex scalar (two reals)."""
    a = np.array([v], dtype=dt)
    _c.keep.append(a)
    return a.ctypes.data_as(C.c_void_p)


_c.keep = []


def _r(v, dt):
    """Pointer to one real scalar of the complex type's precision."""
    return C.byref((C.c_float if dt == np.complex64 else C.c_double)(v))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _pre(dt):
    return "c" if dt == np.complex64 else "z"


def _rel(x, ref):
    x = np.asarray(x, np.complex128)
    ref = np.asarray(ref, np.complex128)
    den = max(np.linalg.norm(ref), 1e-300)
    return np.linalg.norm(x - ref) / den


ALPHAS = [(0.7 - 0.9j, 1.3 - 1.1j), (1.0, 0.0), (-0.5 + 0.25j, 1.0), (0.0, 0.4 + 0.2j)]


@needs_ref
@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("side,uplo,trans,diag", list(itertools.product("LR", "UL", "NTC", "NU")))
def test_trmm_trsm_vs_eigen_blas(dt, side, uplo, trans, diag):
    L = O.blas_ref()
    rng = np.random.default_rng(11)
    for m, n, alpha in ((7, 5, 0.7 - 0.9j), (1, 9, 1.0), (12, 3, -1.5 + 0.5j)):
        ka = m if side == "L" else n
        A = _tri_input(rng, ka, dt, uplo, ka + 2)
        B = _rand(rng, (m, n), dt, m + 2)
        for name, ref in (("trmm_", O.trmm_ref), ("trsm_", O.trsm_ref)):
            want = B.copy(order="F")
            getattr(L, _pre(dt) + name)(side.encode(), uplo.encode(), trans.encode(), diag.encode(), _i(m), _i(n),
                                        _c(alpha, dt), _p(A), _i(A.shape[0]), _p(want), _i(want.shape[0]))
            got = ref(side, uplo, trans, diag, alpha, A, B[:m])
            assert _rel(got, want[:m]) < 10 * _tol(dt), name
            assert np.isnan(want[m:]).all()


@needs_ref
@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("herm", [False, True])
@pytest.mark.parametrize("side,uplo", list(itertools.product("LR", "UL")))
def test_symm_hemm_vs_eigen_blas(dt, herm, side, uplo):
    L = O.blas_ref()
    rng = np.random.default_rng(12)
    for (m, n), (alpha, beta) in zip(((7, 5), (3, 12), (10, 10), (4, 6)), ALPHAS):
        ka = m if side == "L" else n
        A = _tri_input(rng, ka, dt, uplo, ka + 2)
        B = _rand(rng, (m, n), dt, m + 1)
        Cm = _rand(rng, (m, n), dt, m + 3)
        want = Cm.copy(order="F")
        getattr(L, _pre(dt) + ("hemm_" if herm else "symm_"))(
            side.encode(), uplo.encode(), _i(m), _i(n), _c(alpha, dt), _p(A), _i(A.shape[0]), _p(B), _i(B.shape[0]),
            _c(beta, dt), _p(want), _i(want.shape[0]))
        got = (O.hemm_ref if herm else O.symm_ref)(side, uplo, alpha, A, B, beta, Cm[:m])
        assert _rel(got, want[:m]) < _tol(dt)


def _nan_other_triangle(Cm, n, uplo):
    for j in range(n):
        for i in range(n):
            if (uplo == "L" and i < j) or (uplo == "U" and i > j):
                Cm[i, j] = np.nan


@needs_ref
@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,trans", list(itertools.product("UL", "NT")))
def test_syrk_syr2k_vs_eigen_blas(dt, uplo, trans):
    L = O.blas_ref()
    rng = np.random.default_rng(13)
    for (n, k), (alpha, beta) in zip(((7, 5), (12, 3), (5, 11), (6, 0)), ALPHAS):
        shape = (n, k) if trans == "N" else (k, n)
        A = _rand(rng, shape, dt, shape[0] + 2)
        B = _rand(rng, shape, dt, shape[0] + 1)
        Cm = _rand(rng, (n, n), dt, n + 2)
        _nan_other_triangle(Cm, n, uplo)
        mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
        want = Cm.copy(order="F")
        getattr(L, _pre(dt) + "syrk_")(uplo.encode(), trans.encode(), _i(n), _i(k), _c(alpha, dt), _p(A),
                                       _i(A.shape[0]), _c(beta, dt), _p(want), _i(want.shape[0]))
        got = O.syrk_ref(uplo, trans, alpha, A[: shape[0], : shape[1]], beta, Cm[:n])
        assert _rel(got[mask], want[:n][mask]) < _tol(dt)
        assert np.isnan(want[:n][~mask]).all() and np.isnan(got[~mask]).all()
        want = Cm.copy(order="F")
        getattr(L, _pre(dt) + "syr2k_")(uplo.encode(), trans.encode(), _i(n), _i(k), _c(alpha, dt), _p(A),
                                        _i(A.shape[0]), _p(B), _i(B.shape[0]), _c(beta, dt), _p(want),
                                        _i(want.shape[0]))
        got = O.syr2k_ref(uplo, trans, alpha, A[: shape[0], : shape[1]], B[: shape[0], : shape[1]], beta, Cm[:n])
        assert _rel(got[mask], want[:n][mask]) < _tol(dt)


@needs_ref
@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,trans", list(itertools.product("UL", "NC")))
def test_herk_her2k_vs_eigen_blas(dt, uplo, trans):
    L = O.blas_ref()
    rng = np.random.default_rng(14)
    cases = (((7, 5), 0.7, 1.3, 0.7 - 0.9j), ((12, 3), 1.0, 0.0, 1.0), ((5, 11), -0.5, 1.0, -0.5 + 0.25j),
             ((6, 4), 0.0, 1.0, 0.0), ((6, 4), 0.0, 0.4, 0.0), ((6, 0), 1.0, 1.0, 1.0 + 1j), ((4, 3), 2.0, 1.0, 0.3j))
    for (n, k), ra, rb, ca in cases:
        shape = (n, k) if trans == "N" else (k, n)
        A = _rand(rng, shape, dt, shape[0] + 2)
        B = _rand(rng, shape, dt, shape[0] + 1)
        Cm = _rand(rng, (n, n), dt, n + 2)
        _nan_other_triangle(Cm, n, uplo)
        mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
        want = Cm.copy(order="F")
        getattr(L, _pre(dt) + "herk_")(uplo.encode(), trans.encode(), _i(n), _i(k), _r(ra, dt), _p(A),
                                       _i(A.shape[0]), _r(rb, dt), _p(want), _i(want.shape[0]))
        got = O.herk_ref(uplo, trans, ra, A[: shape[0], : shape[1]], rb, Cm[:n])
        assert _rel(got[mask], want[:n][mask]) < _tol(dt), ("herk", n, k, ra, rb)
        want = Cm.copy(order="F")
        getattr(L, _pre(dt) + "her2k_")(uplo.encode(), trans.encode(), _i(n), _i(k), _c(ca, dt), _p(A),
                                        _i(A.shape[0]), _p(B), _i(B.shape[0]), _r(rb, dt), _p(want),
                                        _i(want.shape[0]))
        got = O.her2k_ref(uplo, trans, ca, A[: shape[0], : shape[1]], B[: shape[0], : shape[1]], rb, Cm[:n])
        assert _rel(got[mask], want[:n][mask]) < _tol(dt), ("her2k", n, k, ca, rb)
        assert np.isnan(want[:n][~mask]).all()


@needs_ref
@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,ta,tb", list(itertools.product("UL", "NTC", "NTC")))
def test_gemmtr_vs_eigen_blas(dt, uplo, ta, tb):
    L = O.blas_ref()
    rng = np.random.default_rng(15)
    for (n, k), (alpha, beta) in zip(((7, 5), (13, 1), (4, 9), (9, 0)), ALPHAS):
        A = _rand(rng, (n, k) if ta == "N" else (k, n), dt, (n if ta == "N" else k) + 2)
        B = _rand(rng, (k, n) if tb == "N" else (n, k), dt, (k if tb == "N" else n) + 1)
        Cm = _rand(rng, (n, n), dt, n + 3)
        want = Cm.copy(order="F")
        getattr(L, _pre(dt) + "gemmtr_")(uplo.encode(), ta.encode(), tb.encode(), _i(n), _i(k), _c(alpha, dt), _p(A),
                                         _i(A.shape[0]), _p(B), _i(B.shape[0]), _c(beta, dt), _p(want),
                                         _i(want.shape[0]))
        sa = (n, k) if ta == "N" else (k, n)
        sb = (k, n) if tb == "N" else (n, k)
        got = O.gemmtr_ref(uplo, ta, tb, alpha, A[: sa[0], : sa[1]], B[: sb[0], : sb[1]], beta, Cm[:n])
        assert _rel(got, want[:n]) < _tol(dt)


# ---- identities (no reference code involved) --------------------------------
@pytest.mark.parametrize("side,uplo,trans,diag", list(itertools.product("LR", "UL", "NTC", "NU")))
def test_complex_trsm_undoes_trmm(side, uplo, trans, diag):
    rng = np.random.default_rng(16)
    m, n = 9, 6
    ka = m if side == "L" else n
    A = _tri_input(rng, ka, np.complex128, uplo, ka)
    B = rng.uniform(-1, 1, (m, n)) + 1j * rng.uniform(-1, 1, (m, n))
    Y = O.trmm_ref(side, uplo, trans, diag, 2.0 - 1j, A, B)
    X = O.trsm_ref(side, uplo, trans, diag, 1 / (2.0 - 1j), A, Y)
    assert np.abs(X - B).max() < 1e-12
    # N, T and C really differ
    if trans != "N" and ka > 1:
        other = "T" if trans == "C" else "C"
        assert np.abs(O.trmm_ref(side, uplo, other, diag, 2.0 - 1j, A, B) - Y).max() > 1e-3


@pytest.mark.parametrize("side", "LR")
def test_hemm_is_plain_product_and_ignores_imaginary_diagonal(side):
    rng = np.random.default_rng(17)
    m, n = 8, 5
    ka = m if side == "L" else n
    H = rng.uniform(-1, 1, (ka, ka)) + 1j * rng.uniform(-1, 1, (ka, ka))
    H = H + H.conj().T  # Hermitian: real diagonal
    B = rng.uniform(-1, 1, (m, n)) + 1j * rng.uniform(-1, 1, (m, n))
    C0 = rng.uniform(-1, 1, (m, n)) + 1j * rng.uniform(-1, 1, (m, n))
    want = (0.5 + 0.5j) * (H @ B if side == "L" else B @ H) + 2.0 * C0
    for uplo in "UL":
        A = np.asfortranarray(H.copy())
        A[np.triu_indices(ka, 1) if uplo == "L" else np.tril_indices(ka, -1)] = np.nan
        A[np.diag_indices(ka)] += 7j  # the imaginary part of the diagonal is not read
        assert np.abs(O.hemm_ref(side, uplo, 0.5 + 0.5j, A, B, 2.0, C0) - want).max() < 1e-12
        # symm_ref mirrors without conjugation: different unless H is real
        assert np.abs(O.symm_ref(side, uplo, 0.5 + 0.5j, A, B, 2.0, C0) - want).max() > 1e-3


def test_herk_and_her2k_are_gemmtrs():
    rng = np.random.default_rng(18)
    n, k = 7, 4
    A = rng.uniform(-1, 1, (n, k)) + 1j * rng.uniform(-1, 1, (n, k))
    B = rng.uniform(-1, 1, (n, k)) + 1j * rng.uniform(-1, 1, (n, k))
    C0 = rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n))
    d = np.arange(n)
    for uplo in "UL":
        mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
        g = O.gemmtr_ref(uplo, "N", "C", 0.7, A, A, 1.3, C0)
        g[d, d] = g[d, d].real
        h = O.herk_ref(uplo, "N", 0.7, A, 1.3, C0)
        assert np.abs((h - g)[mask]).max() < 1e-12
        two = O.gemmtr_ref(uplo, "N", "C", np.conj(0.7 - 0.2j), B, A, 1.0,
                           O.gemmtr_ref(uplo, "N", "C", 0.7 - 0.2j, A, B, 1.3, C0))
        two[d, d] = two[d, d].real
        assert np.abs((O.her2k_ref(uplo, "N", 0.7 - 0.2j, A, B, 1.3, C0) - two)[mask]).max() < 1e-12
        # the Hermitian result's triangle is that of a Hermitian matrix: diagonal real
        assert np.abs(h[d, d].imag).max() == 0
