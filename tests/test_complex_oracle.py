"""The oracle's complex GEMM (cgemm_ref) pinned on the CPU against Eigen's own
cgemm_ / zgemm_ (blas/level3_impl.h:12-96 built from the reference tree into
oracle/_ref), every op combination including conjugation, complex alpha and
beta, padded leading dimensions; and against identities that involve no
reference code: conj(A)^T B by the 'C' op equals the explicit conjugate
transpose, and the real case reduces to the real product."""
import ctypes as C
import itertools

import numpy as np
import pytest

import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref (Eigen built from the reference tree) absent")


def _z(rng, shape, dt, ld):
    buf = np.full((ld, max(shape[1], 1)), np.nan + 1j * np.nan, dtype=dt, order="F")
    buf[: shape[0], : shape[1]] = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
    return buf


@needs_ref
@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
@pytest.mark.parametrize("ta,tb", list(itertools.product("NTC", "NTC")))
def test_cgemm_ref_vs_eigen_blas(dt, ta, tb):
    L = O.blas_ref()
    rng = np.random.default_rng(3)
    real = C.c_float if dt == np.complex64 else C.c_double
    fn = L.cgemm_ if dt == np.complex64 else L.zgemm_
    tol = 3e-6 if dt == np.complex64 else 1e-14
    for m, n, k, alpha, beta in ((5, 7, 3, 0.7 - 0.9j, 1.3 - 1.1j), (9, 2, 11, 1.0, 0.0), (4, 4, 1, -0.5j, 1.0)):
        sa = (m, k) if ta == "N" else (k, m)
        sb = (k, n) if tb == "N" else (n, k)
        A = _z(rng, sa, dt, sa[0] + 2)
        B = _z(rng, sb, dt, sb[0] + 1)
        Cm = _z(rng, (m, n), dt, m + 3)
        want = Cm.copy(order="F")
        i = lambda v: C.byref(C.c_int(v))
        za = (real * 2)(complex(alpha).real, complex(alpha).imag)
        zb = (real * 2)(complex(beta).real, complex(beta).imag)
        fn(ta.encode(), tb.encode(), i(m), i(n), i(k), za, A.ctypes.data_as(C.c_void_p), i(A.shape[0]),
           B.ctypes.data_as(C.c_void_p), i(B.shape[0]), zb, want.ctypes.data_as(C.c_void_p), i(want.shape[0]))
        got = O.cgemm_ref(ta, tb, alpha, A[: sa[0], : sa[1]], B[: sb[0], : sb[1]], beta, Cm[:m])
        err = np.abs(got.astype(np.complex128) - want[:m]).max()
        assert err < tol * max(1.0, np.abs(got).max()) * (k + 2), (m, n, k, err)
        assert np.isnan(want[m:]).all()


def test_cgemm_ref_identities():
    rng = np.random.default_rng(4)
    A = rng.uniform(-1, 1, (6, 4)) + 1j * rng.uniform(-1, 1, (6, 4))
    B = rng.uniform(-1, 1, (6, 5)) + 1j * rng.uniform(-1, 1, (6, 5))
    got = O.cgemm_ref("C", "N", 1.0, A, B, 0.0, None)
    assert np.abs(got - A.conj().T @ B).max() < 1e-14
    Ar, Br, Cr = rng.uniform(-1, 1, (3, 4)), rng.uniform(-1, 1, (4, 2)), rng.uniform(-1, 1, (3, 2))
    assert np.abs(O.cgemm_ref("N", "N", 0.5, Ar, Br, 2.0, Cr) - (0.5 * Ar @ Br + 2.0 * Cr)).max() < 1e-14
    # beta == 0 ignores NaN in C
    assert np.isfinite(O.cgemm_ref("N", "N", 1.0, Ar, Br, 0.0, np.full((3, 2), np.nan))).all()
