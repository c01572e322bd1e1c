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
