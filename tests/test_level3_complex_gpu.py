"""Complex level-3 routines on the device (eh_trsm_complex, eh_trmm_complex,
eh_symm_complex, eh_hemm, eh_syrk_complex, eh_herk, eh_syr2k_complex,
eh_her2k, eh_gemmtr_complex) against the oracle's restatements, which
tests/test_level3_complex_oracle.py pins to Eigen's complex BLAS.

Bounds (u = unit roundoff of the real type, written in each test):
* products: ||X - X_ref||_F <= 8 * ka * u * || |alpha| |op(A)| |B| + |beta| |C| ||_F
  (a complex product of order ka is 4 real products of order ka summed in
  pairs: the real-GEMM bound with a factor 2 to spare);
* TRSM: the backward error ||op(T) X - alpha B||_F <= 8 * ka * u * ||T||_F
  ||X||_F, and the forward error against the oracle on well-conditioned
  triangles.
NaN sits wherever a routine must not read: the other triangle of A (and its
diagonal when diag = UNIT, the imaginary part of a Hermitian diagonal), the
padding rows of every leading dimension, the other triangle of C for the rank
updates (which must stay NaN there)."""
import ctypes as C
import itertools

import numpy as np
import pytest

import eigen_amd as E
import oracle as O

pytestmark = pytest.mark.gpu

DT = [np.complex64, np.complex128]
U = {np.complex64: 2.0 ** -24, np.complex128: 2.0 ** -53}
ED = {np.complex64: E.F32, np.complex128: E.F64}
RD = {np.complex64: np.float32, np.complex128: np.float64}
SIDE = {"L": E.LEFT, "R": E.RIGHT}
UPLO = {"U": E.UPPER, "L": E.LOWER}
OP = {"N": E.OP_N, "T": E.OP_T, "C": E.OP_C}
DIAG = {"N": E.NON_UNIT, "U": E.UNIT}
NANC = np.nan + 1j * np.nan


def _padded(rng, rows, cols, dt, ld, scale=1.0):
    a = np.full((ld, max(cols, 1)), NANC, dtype=dt, order="F")
    a[:rows, :cols] = scale * (rng.uniform(-1, 1, (rows, cols)) + 1j * rng.uniform(-1, 1, (rows, cols)))
    return a


def _triangle(rng, ka, dt, uplo, diag, ld):
    """Well-conditioned uplo triangle (diagonal of magnitude 1..2, off-diagonal
    O(1/ka)); NaN in the other triangle, and on the diagonal when unit."""
    A = _padded(rng, ka, ka, dt, ld, 1.0 / ka)
    i, j = np.indices((ka, ka))
    other = (i < j) if uplo == "L" else (i > j)
    A[:ka, :ka][other] = NANC
    d = rng.uniform(1.0, 2.0, ka) * np.exp(1j * rng.uniform(0, 2 * np.pi, ka))
    A[np.arange(ka), np.arange(ka)] = NANC if diag == "U" else d
    return A


def _dev(a):
    import torch
    return torch.from_numpy(np.asfortranarray(a).reshape(-1, order="F").copy()).to("cuda:0")


def _back(t, a):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().reshape(a.shape, order="F")


def _cs(v, dt):
    """Host pointer to one complex scalar (kept alive on the function)."""
    a = np.array([v], dtype=dt)
    _cs.keep.append(a)
    return a.ctypes.data_as(C.c_void_p)


_cs.keep = []


def _rs(v, dt):
    a = np.array([v], dtype=RD[dt])
    _cs.keep.append(a)
    return a.ctypes.data_as(C.c_void_p)


def _call(name, *args):
    E.check(getattr(E.lib(), name)(*args), name)


def _vp(t):
    return C.c_void_p(t.data_ptr())


def _abs_dense_tri(A, ka, uplo, diag):
    T = np.nan_to_num(np.asarray(A[:ka, :ka], np.complex128))
    T = np.tril(T) if uplo == "L" else np.triu(T)
    if diag == "U":
        np.fill_diagonal(T, 1.0)
    return T


TRI = list(itertools.product("LR", "UL", "NTC", "NU"))
TRI_SIZES = [(1, 1), (31, 7), (33, 40), (70, 129)]


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("side,uplo,trans,diag", TRI)
def test_trmm_complex_matches_oracle(gpu, dt, side, uplo, trans, diag):
    rng = np.random.default_rng(21)
    for m, n in TRI_SIZES:
        ka = m if side == "L" else n
        A = _triangle(rng, ka, dt, uplo, diag, ka + 2)
        B = _padded(rng, m, n, dt, m + 1)
        alpha = 0.75 - 0.5j
        dA, dB = _dev(A), _dev(B)
        _call("eh_trmm_complex", ED[dt], SIDE[side], UPLO[uplo], OP[trans], DIAG[diag], m, n, _cs(alpha, dt), _vp(dA),
              A.shape[0], _vp(dB), B.shape[0], None)
        got = _back(dB, B)
        want = O.trmm_ref(side, uplo, trans, diag, alpha, A, B[:m, :n])
        T = np.abs(O._op(_abs_dense_tri(A, ka, uplo, diag), trans))
        absprod = abs(alpha) * (T @ np.abs(B[:m, :n]) if side == "L" else np.abs(B[:m, :n]) @ T)
        err = np.linalg.norm(got[:m, :n] - want)
        assert err <= 8 * ka * U[dt] * np.linalg.norm(absprod), (m, n, err)
        assert np.isnan(got[m:]).all()  # padding rows untouched


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("side,uplo,trans,diag", TRI)
def test_trsm_complex_backward_and_forward_error(gpu, dt, side, uplo, trans, diag):
    rng = np.random.default_rng(22)
    for m, n in TRI_SIZES:
        ka = m if side == "L" else n
        A = _triangle(rng, ka, dt, uplo, diag, ka + 1)
        B = _padded(rng, m, n, dt, m + 2)
        alpha = -1.25 + 0.5j
        dA, dB = _dev(A), _dev(B)
        _call("eh_trsm_complex", ED[dt], SIDE[side], UPLO[uplo], OP[trans], DIAG[diag], m, n, _cs(alpha, dt), _vp(dA),
              A.shape[0], _vp(dB), B.shape[0], None)
        X = _back(dB, B)[:m, :n].astype(np.complex128)
        T = O._op(_abs_dense_tri(A, ka, uplo, diag), trans)
        resid = (T @ X if side == "L" else X @ T) - alpha * B[:m, :n]
        assert np.linalg.norm(resid) <= 8 * ka * U[dt] * np.linalg.norm(T) * np.linalg.norm(X), (m, n)
        want = O.trsm_ref(side, uplo, trans, diag, alpha, A, B[:m, :n])
        assert np.linalg.norm(X - want) <= 8 * ka * U[dt] * np.linalg.norm(np.asarray(want, np.complex128)) * 4


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("herm", [False, True])
@pytest.mark.parametrize("side,uplo", list(itertools.product("LR", "UL")))
def test_symm_hemm_complex_matches_oracle(gpu, dt, herm, side, uplo):
    rng = np.random.default_rng(23)
    for (m, n), (alpha, beta) in zip(((1, 1), (33, 20), (70, 129), (129, 65)),
                                     ((0.7 - 0.9j, 1.3 - 1.1j), (1.0, 0.0), (-0.5 + 0.25j, 1.0), (0.3j, 0.5))):
        ka = m if side == "L" else n
        A = _padded(rng, ka, ka, dt, ka + 2)
        i, j = np.indices((ka, ka))
        A[:ka, :ka][(i < j) if uplo == "L" else (i > j)] = NANC
        if herm:  # only the real part of a Hermitian diagonal is read
            A.imag[np.arange(ka), np.arange(ka)] = np.nan
        B = _padded(rng, m, n, dt, m + 1)
        C0 = _padded(rng, m, n, dt, m + 3)
        if beta == 0:
            C0[:m, :n] = NANC  # beta == 0: C is not read
        dA, dB, dC = _dev(A), _dev(B), _dev(C0)
        _call("eh_hemm" if herm else "eh_symm_complex", ED[dt], SIDE[side], UPLO[uplo], m, n, _cs(alpha, dt), _vp(dA),
              A.shape[0], _vp(dB), B.shape[0], _cs(beta, dt), _vp(dC), C0.shape[0], None)
        got = _back(dC, C0)[:m, :n]
        want = (O.hemm_ref if herm else O.symm_ref)(side, uplo, alpha, A, B, beta, np.nan_to_num(C0[:m, :n]))
        S = np.abs(np.nan_to_num((O._herm_dense if herm else O._sym_dense)(np.nan_to_num(A), ka, uplo == "L")))
        absprod = abs(alpha) * (S @ np.abs(B[:m, :n]) if side == "L" else np.abs(B[:m, :n]) @ S)
        absprod = absprod + abs(beta) * np.abs(np.nan_to_num(C0[:m, :n]))
        assert np.isfinite(got).all()
        assert np.linalg.norm(got - want) <= 8 * ka * U[dt] * np.linalg.norm(absprod), (m, n)


RANK = [(1, 1), (33, 20), (70, 129), (129, 4), (40, 0)]


def _rank_case(rng, dt, n, k, trans, uplo, two):
    shape = (n, k) if trans == "N" else (k, n)
    A = _padded(rng, *shape, dt, shape[0] + 2)
    B = _padded(rng, *shape, dt, shape[0] + 1) if two else None
    C0 = _padded(rng, n, n, dt, n + 1)
    i, j = np.indices((n, n))
    C0[:n, :n][(i < j) if uplo == "L" else (i > j)] = NANC  # never read, must stay NaN
    return shape, A, B, C0


def _check_rank(got, want, C0, n, uplo, k, absprod):
    i, j = np.indices((n, n))
    mask = (i >= j) if uplo == "L" else (i <= j)
    assert np.isnan(got[:n, :n][~mask]).all()
    g = got[:n, :n][mask].astype(np.complex128)
    w = np.asarray(want, np.complex128)[mask]
    assert np.isfinite(g).all()
    assert np.linalg.norm(g - w) <= 8 * max(k, 1) * absprod + 1e-300, (n, k)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("herm", [False, True])
@pytest.mark.parametrize("uplo,trans", list(itertools.product("UL", "NX")))
def test_syrk_herk_complex_matches_oracle(gpu, dt, herm, uplo, trans):
    trans = trans if trans == "N" else ("C" if herm else "T")
    rng = np.random.default_rng(24)
    for (n, k), (ca, cb) in zip(RANK, ((0.7 - 0.9j, 1.3 - 1.1j), (1.0, 0.0), (-0.5 + 0.25j, 1.0), (0.0, 0.4),
                                       (1.0, 0.5))):
        shape, A, _, C0 = _rank_case(rng, dt, n, k, trans, uplo, False)
        ra, rb = float(np.real(ca)), float(np.real(cb))
        dA, dC = _dev(A), _dev(C0)
        if herm:
            _call("eh_herk", ED[dt], UPLO[uplo], OP[trans], n, k, _rs(ra, dt), _vp(dA), A.shape[0], _rs(rb, dt),
                  _vp(dC), C0.shape[0], None)
            want = O.herk_ref(uplo, trans, ra, A[: shape[0], : shape[1]], rb, np.nan_to_num(C0[:n, :n]))
            al, be = ra, rb
        else:
            _call("eh_syrk_complex", ED[dt], UPLO[uplo], OP[trans], n, k, _cs(ca, dt), _vp(dA), A.shape[0],
                  _cs(cb, dt), _vp(dC), C0.shape[0], None)
            want = O.syrk_ref(uplo, trans, ca, A[: shape[0], : shape[1]], cb, np.nan_to_num(C0[:n, :n]))
            al, be = ca, cb
        got = _back(dC, C0)
        Aa = np.abs(A[: shape[0], : shape[1]])
        P = Aa @ Aa.T if trans == "N" else Aa.T @ Aa
        absprod = U[dt] * np.linalg.norm(abs(al) * P + abs(be) * np.abs(np.nan_to_num(C0[:n, :n])))
        _check_rank(got, want, C0, n, uplo, k, absprod)
        if herm and not (rb == 1 and (k == 0 or ra == 0)):  # unless C is left unchanged,
            assert np.all(np.diag(got[:n, :n]).imag == 0)  # a Hermitian result has a real diagonal


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("herm", [False, True])
@pytest.mark.parametrize("uplo,trans", list(itertools.product("UL", "NX")))
def test_syr2k_her2k_complex_matches_oracle(gpu, dt, herm, uplo, trans):
    trans = trans if trans == "N" else ("C" if herm else "T")
    rng = np.random.default_rng(25)
    for (n, k), (ca, cb) in zip(RANK, ((0.7 - 0.9j, 1.3 - 1.1j), (1.0, 0.0), (-0.5 + 0.25j, 1.0), (0.0, 0.4),
                                       (1.0 + 1j, 0.5))):
        shape, A, B, C0 = _rank_case(rng, dt, n, k, trans, uplo, True)
        rb = float(np.real(cb))
        dA, dB, dC = _dev(A), _dev(B), _dev(C0)
        As, Bs = A[: shape[0], : shape[1]], B[: shape[0], : shape[1]]
        if herm:
            _call("eh_her2k", ED[dt], UPLO[uplo], OP[trans], n, k, _cs(ca, dt), _vp(dA), A.shape[0], _vp(dB),
                  B.shape[0], _rs(rb, dt), _vp(dC), C0.shape[0], None)
            want = O.her2k_ref(uplo, trans, ca, As, Bs, rb, np.nan_to_num(C0[:n, :n]))
            be = rb
        else:
            _call("eh_syr2k_complex", ED[dt], UPLO[uplo], OP[trans], n, k, _cs(ca, dt), _vp(dA), A.shape[0], _vp(dB),
                  B.shape[0], _cs(cb, dt), _vp(dC), C0.shape[0], None)
            want = O.syr2k_ref(uplo, trans, ca, As, Bs, cb, np.nan_to_num(C0[:n, :n]))
            be = cb
        got = _back(dC, C0)
        Aa, Ba = np.abs(As), np.abs(Bs)
        P = Aa @ Ba.T + Ba @ Aa.T if trans == "N" else Aa.T @ Ba + Ba.T @ Aa
        absprod = U[dt] * np.linalg.norm(abs(ca) * P + abs(be) * np.abs(np.nan_to_num(C0[:n, :n])))
        _check_rank(got, want, C0, n, uplo, k, absprod)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("uplo,ta,tb", list(itertools.product("UL", "NTC", "NTC")))
def test_gemmtr_complex_matches_oracle(gpu, dt, uplo, ta, tb):
    rng = np.random.default_rng(26)
    for (n, k), (ca, cb) in zip(((1, 1), (70, 33), (129, 4)), ((0.7 - 0.9j, 1.3 - 1.1j), (1.0, 0.0), (0.3j, 1.0))):
        sa = (n, k) if ta == "N" else (k, n)
        sb = (k, n) if tb == "N" else (n, k)
        A = _padded(rng, *sa, dt, sa[0] + 1)
        B = _padded(rng, *sb, dt, sb[0] + 2)
        C0 = _padded(rng, n, n, dt, n + 1)
        i, j = np.indices((n, n))
        C0[:n, :n][(i < j) if uplo == "L" else (i > j)] = NANC
        dA, dB, dC = _dev(A), _dev(B), _dev(C0)
        _call("eh_gemmtr_complex", ED[dt], UPLO[uplo], OP[ta], OP[tb], n, k, _cs(ca, dt), _vp(dA), A.shape[0],
              _vp(dB), B.shape[0], _cs(cb, dt), _vp(dC), C0.shape[0], None)
        got = _back(dC, C0)
        As, Bs = A[: sa[0], : sa[1]], B[: sb[0], : sb[1]]
        want = O.gemmtr_ref(uplo, ta, tb, ca, As, Bs, cb, np.nan_to_num(C0[:n, :n]))
        P = np.abs(O._op(As, ta)) @ np.abs(O._op(Bs, tb))
        absprod = U[dt] * np.linalg.norm(abs(ca) * P + abs(cb) * np.abs(np.nan_to_num(C0[:n, :n])))
        _check_rank(got, want, C0, n, uplo, k, absprod)


def test_blas_complex_exports_host_and_device(gpu):
    """ztrsm_ and chemm_ through the BLAS library: host pointers (staged) and
    device pointers give the results of the C-ABI routines."""
    import torch
    L = E.blas()
    rng = np.random.default_rng(27)
    i_ = lambda v: C.byref(C.c_int(v))
    m, n = 45, 30
    A = _triangle(rng, m, np.complex128, "L", "N", m)
    B = _padded(rng, m, n, np.complex128, m)
    want = O.trsm_ref("L", "L", "C", "N", 2.0 - 1j, A, B)
    Bh = B.copy(order="F")
    L.ztrsm_(b"L", b"L", b"C", b"N", i_(m), i_(n), _cs(2.0 - 1j, np.complex128), A.ctypes.data_as(C.c_void_p), i_(m),
             Bh.ctypes.data_as(C.c_void_p), i_(m))
    assert np.linalg.norm(Bh - want) <= 1e-12 * np.linalg.norm(np.asarray(want, np.complex128))
    dA, dB = _dev(A), _dev(B)
    L.ztrsm_(b"L", b"L", b"C", b"N", i_(m), i_(n), _cs(2.0 - 1j, np.complex128), _vp(dA), i_(m), _vp(dB), i_(m))
    assert np.array_equal(_back(dB, B), Bh)
    H = _padded(rng, n, n, np.complex64, n)
    Bc = _padded(rng, m, n, np.complex64, m)
    Cc = _padded(rng, m, n, np.complex64, m)
    Ch = Cc.copy(order="F")
    L.chemm_(b"R", b"U", i_(m), i_(n), _cs(0.5 + 1j, np.complex64), H.ctypes.data_as(C.c_void_p), i_(n),
             Bc.ctypes.data_as(C.c_void_p), i_(m), _cs(1.5, np.complex64), Ch.ctypes.data_as(C.c_void_p), i_(m))
    dH, dBc, dC = _dev(H), _dev(Bc), _dev(Cc)  # held until the call has run
    _call("eh_hemm", E.F32, E.RIGHT, E.UPPER, m, n, _cs(0.5 + 1j, np.complex64), _vp(dH), n, _vp(dBc), m,
          _cs(1.5, np.complex64), _vp(dC), m, None)
    torch.cuda.synchronize()
    assert np.array_equal(_back(dC, Cc), Ch)
