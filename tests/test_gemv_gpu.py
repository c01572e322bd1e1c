"""Strided matrix-vector products (eh_gemv, sgemv_/dgemv_) on the device.

Reference: the product in long double on the host (numpy), and Eigen's own
dgemv_/sgemv_ (blas/level2_impl.h:27-84, built into oracle/_ref) for the BLAS
entry points on host arrays.  Bound, written in each test:
||y - y_ref||_2 <= k u || |alpha||op(A)||x| + |beta||y0| ||_2 (k the inner
length).  Vector gaps between strided elements hold NaN and must stay
untouched; alpha == 0 must not read A or x, beta == 0 must not read y."""
import ctypes as C
import itertools

import numpy as np
import pytest

import eigen_amd as E
import oracle as O

pytestmark = pytest.mark.gpu

U = {np.float32: 2.0 ** -24, np.float64: 2.0 ** -53}
ED = {np.float32: E.F32, np.float64: E.F64}


def _strided(vals, inc, dt):
    """BLAS storage of a vector with increment inc (NaN in the gaps)."""
    n = len(vals)
    buf = np.full(max(1, 1 + (n - 1) * abs(inc)), np.nan, dtype=dt)
    for i, v in enumerate(vals):
        buf[(i if inc > 0 else i - (n - 1)) * inc] = v
    return buf


def _gather(buf, n, inc):
    return np.array([buf[(i if inc > 0 else i - (n - 1)) * inc] for i in range(n)])


def _dev(a):
    import torch
    return torch.from_numpy(np.asfortranarray(a).reshape(-1, order="F").copy()).to("cuda:0")


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


CASES = list(itertools.product([np.float32, np.float64], ["N", "T"], [(1, 1), (2, -3), (-1, 2), (-2, -1)]))


@pytest.mark.parametrize("dt,trans,incs", CASES)
def test_gemv_strided_bound(gpu, dt, trans, incs):
    rng = np.random.default_rng(31)
    incx, incy = incs
    for m, n, alpha, beta in ((1, 1, 1.0, 0.0), (300, 257, 0.75, -1.5), (4100, 70, -1.0, 1.0), (64, 5000, 2.0, 0.5)):
        lda = m + 3
        A = np.full((lda, n), np.nan, dt, order="F")
        A[:m] = rng.uniform(-1, 1, (m, n))
        lx, ly = (n, m) if trans == "N" else (m, n)
        x = rng.uniform(-1, 1, lx)
        y0 = rng.uniform(-1, 1, ly)
        xb, yb = _strided(x, incx, dt), _strided(y0, incy, dt)
        dy = _dev(yb)
        E.gemv(ED[dt], E.OP_N if trans == "N" else E.OP_T, m, n, alpha, _dev(A), lda, _dev(xb), incx, beta, dy, incy)
        got_buf = _host(dy)
        gaps = np.ones(len(yb), bool)
        for i in range(ly):
            gaps[(i if incy > 0 else i - (ly - 1)) * incy] = False
        assert np.isnan(got_buf[gaps]).all(), "gap between strided elements written"
        got = _gather(got_buf, ly, incy).astype(np.float64)
        Am = A[:m].astype(np.longdouble)
        op = Am if trans == "N" else Am.T
        xv = x.astype(dt).astype(np.longdouble)
        y0v = y0.astype(dt).astype(np.longdouble)
        want = alpha * (op @ xv) + (beta * y0v if beta != 0 else 0)
        bound = lx * U[dt] * np.linalg.norm(abs(alpha) * np.abs(op) @ np.abs(xv) + abs(beta) * np.abs(y0v))
        assert np.linalg.norm(got - want.astype(np.float64)) <= bound, (m, n)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_gemv_alpha_zero_and_beta_zero(gpu, dt):
    rng = np.random.default_rng(32)
    m, n = 50, 40
    Anan = np.full((m, n), np.nan, dt, order="F")
    y0 = rng.uniform(-1, 1, m).astype(dt)
    dy = _dev(y0)
    E.gemv(ED[dt], E.OP_N, m, n, 0.0, _dev(Anan), m, _dev(np.full(n, np.nan, dt)), 1, 2.0, dy, 1)
    assert np.array_equal(_host(dy), 2.0 * y0)
    A = rng.uniform(-1, 1, (m, n)).astype(dt, order="F")
    dy = _dev(np.full(m, np.nan, dt))
    E.gemv(ED[dt], E.OP_N, m, n, 1.0, _dev(A), m, _dev(np.ones(n, dt)), 1, 0.0, dy, 1)
    assert np.isfinite(_host(dy)).all()


@pytest.mark.parametrize("dt,trans", list(itertools.product([np.float32, np.float64], "NT")))
def test_gemv_blas_host_vs_eigen_blas(gpu, dt, trans):
    """sgemv_/dgemv_ on host arrays with negative increments, against Eigen's
    own ?gemv_ on the same arrays."""
    if not O.ref_available():
        pytest.skip("oracle/_ref absent")
    R = O.blas_ref()
    L = E.blas()
    rng = np.random.default_rng(33)
    m, n, incx, incy = 333, 130, -2, 3
    A = rng.uniform(-1, 1, (m + 1, n)).astype(dt, order="F")
    lx, ly = (n, m) if trans == "N" else (m, n)
    xb = _strided(rng.uniform(-1, 1, lx), incx, dt)
    yb = _strided(rng.uniform(-1, 1, ly), incy, dt)
    ours, ref = yb.copy(), yb.copy()
    ct = C.c_float if dt == np.float32 else C.c_double
    i = lambda v: C.byref(C.c_int(v))
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    for lib, out in ((L, ours), (R, ref)):
        f = lib.sgemv_ if dt == np.float32 else lib.dgemv_
        f(trans.encode(), i(m), i(n), C.byref(ct(0.5)), p(A), i(m + 1), p(xb), i(incx), C.byref(ct(-1.25)), p(out),
          i(incy))
    assert np.array_equal(np.isnan(ours), np.isnan(ref))
    ok = ~np.isnan(ref)
    op = A[:m].astype(np.float64) if trans == "N" else A[:m].astype(np.float64).T
    xv = np.abs(_gather(xb, lx, incx).astype(np.float64))
    bound = 2 * lx * U[dt] * np.linalg.norm(0.5 * np.abs(op) @ xv + 1.25 * np.abs(_gather(yb, ly, incy)))
    assert np.linalg.norm(ours[ok].astype(np.float64) - ref[ok]) <= bound
