"""Complex product on the device (eh_gemm_complex, cgemm_/zgemm_) against the
oracle's cgemm_ref (extended precision, pinned to Eigen's cgemm_/zgemm_ on the
CPU).  Bound, written in each test: ||C - C_ref||_F <= 4 k u ||E||_F with
E = |alpha| |op(A)| |op(B)| + |beta| |C0| (entrywise moduli; each complex
multiply-add is four real ones summed into two FP accumulators).  Padding
rows hold NaN and must not reach the result; beta == 0 must not read C."""
import ctypes as C
import itertools

import numpy as np
import pytest

import eigen_amd as E
import oracle as O
from helpers import from_dev

pytestmark = pytest.mark.gpu

U = {np.complex64: 2.0 ** -24, np.complex128: 2.0 ** -53}
ED = {np.complex64: E.F32, np.complex128: E.F64}
OPS = {"N": E.OP_N, "T": E.OP_T, "C": E.OP_C}


def _z(rng, shape, dt, ld):
    buf = np.full((ld, max(shape[1], 1)), np.nan + 1j * np.nan, dtype=dt, order="F")
    buf[: shape[0], : shape[1]] = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
    return buf


def _dev(a):
    import torch
    flat = np.asfortranarray(a).reshape(-1, order="F")
    return torch.from_numpy(flat.copy()).to("cuda:0")


def _back(t, shape):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()[: shape[0] * shape[1]].reshape(shape, order="F")


def _op_abs(x, t):
    x = np.abs(np.asarray(x, np.complex128))
    return x if t == "N" else x.T


SHAPES = [(1, 1, 1), (67, 45, 33), (300, 129, 257), (130, 513, 64)]


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
@pytest.mark.parametrize("ta,tb", list(itertools.product("NTC", "NTC")))
def test_complex_product_bound(gpu, dt, ta, tb):
    rng = np.random.default_rng(21)
    for (m, n, k), (alpha, beta) in zip(SHAPES, [(0.7 - 0.9j, 1.3 - 1.1j), (1.0, 0.0), (-0.5j, 0.75),
                                                  (2.0 + 1.0j, 1.0)]):
        sa = (m, k) if ta == "N" else (k, m)
        sb = (k, n) if tb == "N" else (n, k)
        A = _z(rng, sa, dt, sa[0] + 1)
        B = _z(rng, sb, dt, sb[0] + 2)
        C0 = _z(rng, (m, n), dt, m + 3)
        if beta == 0:
            C0[:m, :n] = np.nan  # beta == 0: C is not read
        dC = _dev(C0)
        E.gemm_complex(ED[dt], OPS[ta], OPS[tb], m, n, k, alpha, _dev(A), A.shape[0], _dev(B), B.shape[0], beta, dC,
                       C0.shape[0])
        got = _back(dC, C0.shape)
        assert np.isnan(got[m:]).all(), "padding rows of C written"
        Am, Bm = A[: sa[0], : sa[1]], B[: sb[0], : sb[1]]
        want = O.cgemm_ref(ta, tb, alpha, Am, Bm, beta, C0[:m]).astype(np.complex128)
        Eb = abs(alpha) * (_op_abs(Am, ta) @ _op_abs(Bm, tb))
        if beta != 0:
            Eb = Eb + abs(beta) * np.abs(C0[:m].astype(np.complex128))
        err = np.linalg.norm(got[:m].astype(np.complex128) - want)
        assert err <= 4 * k * U[dt] * np.linalg.norm(Eb), (m, n, k, err)


def test_complex_split3_precision(gpu):
    """complex64 through the split-BF16 product keeps the FP32 bound."""
    rng = np.random.default_rng(22)
    m, n, k = 256, 200, 512
    A = _z(rng, (m, k), np.complex64, m)
    B = _z(rng, (k, n), np.complex64, k)
    dC = _dev(np.zeros((m, n), np.complex64))
    E.gemm_complex(E.F32, E.OP_N, E.OP_N, m, n, k, 1.0 + 0.5j, _dev(A), m, _dev(B), k, 0.0, dC, m,
                   prec=E.PREC_SPLIT3)
    got = _back(dC, (m, n)).astype(np.complex128)
    want = O.cgemm_ref("N", "N", 1.0 + 0.5j, A, B, 0.0, None).astype(np.complex128)
    Eb = abs(1.0 + 0.5j) * np.abs(A.astype(np.complex128)) @ np.abs(B.astype(np.complex128))
    assert np.linalg.norm(got - want) <= 4 * k * U[np.complex64] * np.linalg.norm(Eb)


def test_complex_alpha_zero_scales_by_complex_beta(gpu):
    rng = np.random.default_rng(23)
    m, n, k = 20, 30, 10
    C0 = _z(rng, (m, n), np.complex128, m)
    nanA = np.full((m, k), np.nan + 0j, np.complex128, order="F")
    dC = _dev(C0)
    E.gemm_complex(E.F64, E.OP_N, E.OP_N, m, n, k, 0.0, _dev(nanA), m, _dev(nanA), k, 0.5 - 2.0j, dC, m)
    assert np.abs(_back(dC, (m, n)) - (0.5 - 2.0j) * C0).max() < 1e-15


def test_zgemm_host_pointers(gpu):
    """zgemm_ on host arrays (staged) against the device entry point."""
    L = E.blas()
    rng = np.random.default_rng(24)
    m, n, k = 40, 33, 25
    A = _z(rng, (k, m), np.complex128, k)
    B = _z(rng, (k, n), np.complex128, k + 1)
    Cm = _z(rng, (m, n), np.complex128, m)
    dC = _dev(Cm)
    E.gemm_complex(E.F64, E.OP_C, E.OP_N, m, n, k, 0.7 - 0.9j, _dev(A), k, _dev(B), k + 1, 1.3 - 1.1j, dC, m)
    want = _back(dC, (m, n))
    i = lambda v: C.byref(C.c_int(v))
    za = (C.c_double * 2)(0.7, -0.9)
    zb = (C.c_double * 2)(1.3, -1.1)
    L.zgemm_(b"C", b"N", i(m), i(n), i(k), za, A.ctypes.data_as(C.c_void_p), i(k), B.ctypes.data_as(C.c_void_p),
             i(k + 1), zb, Cm.ctypes.data_as(C.c_void_p), i(m))
    assert np.array_equal(Cm, want)


def _strided_z(vals, inc, dt):
    n = len(vals)
    buf = np.full(max(1, 1 + (n - 1) * abs(inc)), np.nan + 1j * np.nan, dtype=dt)
    for i, v in enumerate(vals):
        buf[(i if inc > 0 else i - (n - 1)) * inc] = v
    return buf


def _gather_z(buf, n, inc):
    return np.array([buf[(i if inc > 0 else i - (n - 1)) * inc] for i in range(n)])


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
@pytest.mark.parametrize("trans,incs", list(itertools.product("NTC", [(1, 1), (-2, 3), (2, -1)])))
def test_complex_gemv_bound(gpu, dt, trans, incs):
    """eh_gemv_complex: bound 4 k u || |alpha||op(A)||x| + |beta||y0| || against the
    extended-precision product; gaps between strided elements stay NaN."""
    rng = np.random.default_rng(25)
    incx, incy = incs
    for m, n, alpha, beta in ((1, 1, 1.0, 0.0), (300, 129, 0.7 - 0.9j, 1.3 - 1.1j), (64, 1000, -0.5j, 0.0)):
        A = _z(rng, (m, n), dt, m + 2)
        lx, ly = (n, m) if trans == "N" else (m, n)
        x = rng.uniform(-1, 1, lx) + 1j * rng.uniform(-1, 1, lx)
        y0 = rng.uniform(-1, 1, ly) + 1j * rng.uniform(-1, 1, ly)
        xb, yb = _strided_z(x, incx, dt), _strided_z(y0, incy, dt)
        if beta == 0:
            yb[~np.isnan(yb)] = np.nan  # beta == 0: y is not read
        dy = _dev(yb)
        E.gemv_complex(ED[dt], OPS[trans], m, n, alpha, _dev(A), A.shape[0], _dev(xb), incx, beta, dy, incy)
        got_buf = _back(dy, (len(yb), 1))[:, 0]
        got = _gather_z(got_buf, ly, incy).astype(np.complex128)
        gaps = np.ones(len(yb), bool)
        for i in range(ly):
            gaps[(i if incy > 0 else i - (ly - 1)) * incy] = False
        assert np.isnan(got_buf[gaps]).all()
        Am = A[:m, :n]
        want = O.cgemm_ref(trans, "N", alpha, Am, x.astype(dt).reshape(-1, 1), beta,
                           None if beta == 0 else y0.astype(dt).reshape(-1, 1))[:, 0].astype(np.complex128)
        Eb = abs(alpha) * (_op_abs(Am, trans) @ np.abs(x.astype(dt).astype(np.complex128)))
        if beta != 0:
            Eb = Eb + abs(beta) * np.abs(y0.astype(dt).astype(np.complex128))
        assert np.linalg.norm(got - want) <= 4 * lx * U[dt] * np.linalg.norm(Eb), (m, n)


def test_zgemv_host_vs_eigen_blas(gpu):
    """zgemv_ on host arrays with negative increments against Eigen's own zgemv_."""
    if not O.ref_available():
        pytest.skip("oracle/_ref absent")
    R = O.blas_ref()
    L = E.blas()
    rng = np.random.default_rng(26)
    m, n, incx, incy = 150, 90, -2, 3
    A = _z(rng, (m, n), np.complex128, m + 1)
    for trans in "NTC":
        lx, ly = (n, m) if trans == "N" else (m, n)
        xb = _strided_z(rng.uniform(-1, 1, lx) + 1j * rng.uniform(-1, 1, lx), incx, np.complex128)
        yb = _strided_z(rng.uniform(-1, 1, ly) + 1j * rng.uniform(-1, 1, ly), incy, np.complex128)
        ours, ref = yb.copy(), yb.copy()
        i = lambda v: C.byref(C.c_int(v))
        za, zb = (C.c_double * 2)(0.7, -0.9), (C.c_double * 2)(1.3, -1.1)
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        for lib, out in ((L, ours), (R, ref)):
            lib.zgemv_(trans.encode(), i(m), i(n), za, p(A), i(m + 1), p(xb), i(incx), zb, p(out), i(incy))
        ok = ~np.isnan(ref)
        assert np.array_equal(ok, ~np.isnan(ours))
        assert np.abs(ours[ok] - ref[ok]).max() < 1e-12 * lx, trans
