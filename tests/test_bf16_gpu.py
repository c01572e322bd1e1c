"""bfloat16-input product with an FP32 result (eh_gemm_bf16 / sbgemm_).

Oracle: the inputs widened exactly to FP32 and multiplied in long double
(oracle.gemm_ref).  Bars, written in each test:
* exact: integer-valued bf16 inputs whose every partial sum is an integer
  below 2^24 give a bit-exact FP32 result;
* bound: ||C - C_ref||_F <= k * u32 * || |alpha| |A||B| + |beta| |C0| ||_F
  (bf16 x bf16 products are exact in FP32, only the FP32 accumulation rounds).
Padding rows of every leading dimension hold NaN bit patterns, which must
never reach the result."""
import ctypes as C
import itertools

import numpy as np
import pytest

import eigen_amd as E
import oracle as O
from helpers import from_dev

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24
NAN_BF16 = np.uint16(0x7FC0)


def _bf16(rng, rows, cols, ld, lo=-1.0, hi=1.0, ints=False):
    """(ld x cols) uint16 array of bf16 bits, column-major; NaN in the padding rows."""
    buf = np.full((ld, max(cols, 1)), NAN_BF16, dtype=np.uint16, order="F")
    vals = rng.integers(-8, 9, (rows, cols)).astype(np.float32) if ints else \
        rng.uniform(lo, hi, (rows, cols)).astype(np.float32)
    buf[:rows, :cols] = O.bf16_rtne(vals).reshape(rows, cols)
    return buf


def _widen(bits):
    return O.bf16_to_f32(bits)


def _dev16(a, offset=0):
    """Flat device copy (int16 storage) with `offset` leading elements of slack."""
    import torch
    flat = np.concatenate([np.zeros(offset, np.uint16), np.asfortranarray(a).reshape(-1, order="F")])
    t = torch.from_numpy(flat.view(np.int16).copy()).to("cuda:0")
    return t, t.data_ptr() + 2 * offset


def _dev32(a):
    import torch
    return torch.from_numpy(np.asfortranarray(a, dtype=np.float32).reshape(-1, order="F").copy()).to("cuda:0")


SHAPES = [(1, 1, 1), (67, 45, 33), (129, 300, 257), (300, 129, 1000), (1024, 768, 4096), (2048, 2048, 512)]


@pytest.mark.parametrize("ta,tb", list(itertools.product((0, 1), (0, 1))))
@pytest.mark.parametrize("m,n,k", SHAPES)
def test_bf16_product_bound(gpu, ta, tb, m, n, k):
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    sa = (k, m) if ta else (m, k)
    sb = (n, k) if tb else (k, n)
    A = _bf16(rng, *sa, sa[0] + 3)
    B = _bf16(rng, *sb, sb[0] + 1)
    C0 = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    alpha, beta = 0.75, -0.5
    tA, pA = _dev16(A, offset=1)  # 2-byte aligned start
    tB, pB = _dev16(B)
    dC = _dev32(C0)
    E.gemm_bf16(ta, tb, m, n, k, alpha, pA, A.shape[0], pB, B.shape[0], beta, dC, m)
    got = from_dev(dC, (m, n)).astype(np.float64)
    Af = _widen(A[: sa[0], : sa[1]]).reshape(sa).astype(np.float32, order="F")
    Bf = _widen(B[: sb[0], : sb[1]]).reshape(sb).astype(np.float32, order="F")
    want, absprod = O.gemm_ref(bool(ta), bool(tb), alpha, np.asfortranarray(Af), np.asfortranarray(Bf), beta,
                               np.asfortranarray(C0), with_abs=True)
    absprod = absprod + abs(beta) * np.abs(C0)
    assert np.isfinite(got).all()
    err = np.linalg.norm(got - want)
    assert err <= k * U32 * np.linalg.norm(absprod), (err, k * U32 * np.linalg.norm(absprod))


@pytest.mark.parametrize("ta,tb", list(itertools.product((0, 1), (0, 1))))
def test_bf16_integer_inputs_exact(gpu, ta, tb):
    rng = np.random.default_rng(5)
    m, n, k = 190, 70, 1000
    sa = (k, m) if ta else (m, k)
    sb = (n, k) if tb else (k, n)
    A = _bf16(rng, *sa, sa[0], ints=True)
    B = _bf16(rng, *sb, sb[0], ints=True)
    tA, pA = _dev16(A)
    tB, pB = _dev16(B)
    dC = _dev32(np.full((m, n), np.nan, np.float32))  # beta == 0: C is not read
    E.gemm_bf16(ta, tb, m, n, k, 1.0, pA, A.shape[0], pB, B.shape[0], 0.0, dC, m)
    got = from_dev(dC, (m, n))
    a = _widen(A).reshape(A.shape).astype(np.int64)
    b = _widen(B).reshape(B.shape).astype(np.int64)
    a, b = (a.T if ta else a), (b.T if tb else b)
    assert np.array_equal(got, (a @ b).astype(np.float32))


@pytest.mark.parametrize("ta,tb", list(itertools.product((0, 1), (0, 1))))
@pytest.mark.parametrize("m,n,k,pad,offset", [(300, 270, 64, 8, 0), (129, 513, 1024, 16, 8), (64, 64, 4096, 8, 0),
                                              (300, 270, 64, 8, 1), (200, 130, 96, 3, 0)])
def test_bf16_in_place_operands_exact(gpu, ta, tb, m, n, k, pad, offset):
    """Operands whose op(X) rows are k-contiguous (A^T, B) are read in place
    when k is a multiple of 32 and the rows are 16-byte aligned (leading
    dimension a multiple of 8, base aligned); the others, and these with an
    odd padding or a 2-byte shifted base, go through the packed layout.
    Integer inputs (exact result), NaN in every padding element, ragged m
    and n (rows past the matrix), k = 4096 on a 64 x 64 result (split k)."""
    rng = np.random.default_rng(m + n + k + pad + offset)
    sa = (k, m) if ta else (m, k)
    sb = (n, k) if tb else (k, n)
    A = _bf16(rng, *sa, sa[0] + pad, ints=True)
    B = _bf16(rng, *sb, sb[0] + pad, ints=True)
    tA, pA = _dev16(A, offset=offset)
    tB, pB = _dev16(B, offset=offset)
    dC = _dev32(np.full((m, n), np.nan, np.float32))
    E.gemm_bf16(ta, tb, m, n, k, 1.0, pA, A.shape[0], pB, B.shape[0], 0.0, dC, m)
    got = from_dev(dC, (m, n))
    a = _widen(A[: sa[0], : sa[1]]).reshape(sa).astype(np.int64)
    b = _widen(B[: sb[0], : sb[1]]).reshape(sb).astype(np.int64)
    a, b = (a.T if ta else a), (b.T if tb else b)
    assert np.array_equal(got, (a @ b).astype(np.float32))


def test_bf16_alpha_zero_and_empty(gpu):
    rng = np.random.default_rng(6)
    m, n, k = 33, 17, 40
    C0 = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    dC = _dev32(C0)
    A = np.full((m, k), NAN_BF16, np.uint16, order="F")
    tA, pA = _dev16(A)
    E.gemm_bf16(0, 0, m, n, k, 0.0, pA, m, pA, k, 2.0, dC, m)  # A, B not read
    assert np.array_equal(from_dev(dC, (m, n)), 2.0 * C0)
    E.gemm_bf16(0, 0, m, n, 0, 1.0, pA, m, pA, 1, 0.5, dC, m)  # k == 0
    assert np.array_equal(from_dev(dC, (m, n)), C0)


def test_sbgemm_host_pointers(gpu):
    """sbgemm_ on host arrays (staged through device buffers), against the device entry point."""
    L = E.blas()
    rng = np.random.default_rng(7)
    m, n, k = 50, 40, 70
    A = _bf16(rng, m, k, m + 2)
    B = _bf16(rng, n, k, n)
    Cm = rng.uniform(-1, 1, (m + 1, n)).astype(np.float32, order="F")
    want = Cm.copy(order="F")
    tA, pA = _dev16(A)
    tB, pB = _dev16(B)
    dC = _dev32(want)
    E.gemm_bf16(0, 1, m, n, k, 1.5, pA, m + 2, pB, n, 0.25, dC, m + 1)
    want = from_dev(dC, (m + 1, n))
    i = lambda v: C.byref(C.c_int(v))
    f = lambda v: C.byref(C.c_float(v))
    L.sbgemm_(b"N", b"T", i(m), i(n), i(k), f(1.5), A.ctypes.data_as(C.c_void_p), i(m + 2),
              B.ctypes.data_as(C.c_void_p), i(n), f(0.25), Cm.ctypes.data_as(C.c_void_p), i(m + 1))
    assert np.array_equal(Cm[:m], want[:m])


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
def test_bf16_large_tile_sampled_rows(gpu, ta, tb):
    """A product large enough for the 256x256 tile (ragged n), checked on 96
    sampled rows against the long-double oracle with the same bound."""
    rng = np.random.default_rng(8)
    m, n, k = 4096, 4160, 96
    sa = (k, m) if ta else (m, k)
    sb = (n, k) if tb else (k, n)
    A = _bf16(rng, *sa, sa[0])
    B = _bf16(rng, *sb, sb[0] + 1)
    tA, pA = _dev16(A)
    tB, pB = _dev16(B)
    dC = _dev32(np.full((m, n), np.nan, np.float32))
    E.gemm_bf16(ta, tb, m, n, k, 1.0, pA, A.shape[0], pB, B.shape[0], 0.0, dC, m)
    got = from_dev(dC, (m, n)).astype(np.float64)
    rows = np.sort(rng.choice(m, 96, replace=False))
    Af = np.asfortranarray(_widen(A[: sa[0], : sa[1]]).reshape(sa))
    Bf = np.asfortranarray(_widen(B[: sb[0], : sb[1]]).reshape(sb))
    want, absprod = O.gemm_ref(bool(ta), bool(tb), 1.0, Af, Bf, rows=rows, with_abs=True)
    assert np.isfinite(got).all()
    assert np.linalg.norm(got[rows] - want) <= k * U32 * np.linalg.norm(absprod)
