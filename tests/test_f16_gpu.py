"""IEEE half inputs with an FP32 result (eh_gemm_f16), on the FP16 matrix cores.

Oracle: the inputs widened exactly to FP32 (oracle.half_to_f32, pinned in
tests/test_oracle.py) and multiplied in long double (oracle.gemm_ref).  Bars,
written in each test:
* exact: integer-valued half inputs (and the same scaled into the subnormal
  range) whose every partial sum is exact in FP32 give a bit-exact result;
* bound: ||C - C_ref||_F <= k * u32 * || |alpha| |A||B| + |beta| |C0| ||_F
  (half x half products are exact in FP32, only the FP32 accumulation rounds);
* against Eigen's own Matrix<half> product (oracle/_ref), which rounds every
  multiply-add to half on this host: the FP32-accumulated result is at least
  as accurate, entry for entry in norm.
Padding rows of every leading dimension hold NaN bit patterns, which must
never reach the result."""
import itertools

import numpy as np
import pytest

import eigen_amd as E
import oracle as O
from helpers import from_dev

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24
NAN_F16 = np.uint16(0x7E00)


def _f16(rng, rows, cols, ld, lo=-1.0, hi=1.0, ints=False, scale=1.0):
    """(ld x cols) uint16 array of half bits, column-major; NaN in the padding rows."""
    buf = np.full((ld, max(cols, 1)), NAN_F16, dtype=np.uint16, order="F")
    vals = rng.integers(-16, 17, (rows, cols)).astype(np.float32) if ints else \
        rng.uniform(lo, hi, (rows, cols)).astype(np.float32)
    buf[:rows, :cols] = O.half_rtne(vals * np.float32(scale)).reshape(rows, cols)
    return buf


def _widen(bits):
    return O.half_to_f32(bits)


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
def test_f16_product_bound(gpu, ta, tb, m, n, k):
    rng = np.random.default_rng(m * 5 + n * 11 + k)
    sa = (k, m) if ta else (m, k)
    sb = (n, k) if tb else (k, n)
    A = _f16(rng, *sa, sa[0] + 3)
    B = _f16(rng, *sb, sb[0] + 1)
    C0 = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    alpha, beta = 0.75, -0.5
    tA, pA = _dev16(A, offset=1)  # 2-byte aligned start
    tB, pB = _dev16(B)
    dC = _dev32(C0)
    E.gemm_f16(ta, tb, m, n, k, alpha, pA, A.shape[0], pB, B.shape[0], beta, dC, m)
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
@pytest.mark.parametrize("scale", [1.0, 2.0 ** -24])
def test_f16_integer_inputs_exact(gpu, ta, tb, scale):
    """Integers in [-16, 16] (k = 1000: partial sums below 2^18, exact in FP32);
    with scale 2^-24 the A operand is made of subnormal halves (integer
    multiples of 2^-24), which must not be flushed to zero."""
    rng = np.random.default_rng(5)
    m, n, k = 190, 70, 1000
    sa = (k, m) if ta else (m, k)
    sb = (n, k) if tb else (k, n)
    A = _f16(rng, *sa, sa[0], ints=True, scale=scale)
    B = _f16(rng, *sb, sb[0], ints=True)
    tA, pA = _dev16(A)
    tB, pB = _dev16(B)
    dC = _dev32(np.full((m, n), np.nan, np.float32))  # beta == 0: C is not read
    E.gemm_f16(ta, tb, m, n, k, 1.0, pA, A.shape[0], pB, B.shape[0], 0.0, dC, m)
    got = from_dev(dC, (m, n)).astype(np.float64)
    a = np.rint(_widen(A).reshape(A.shape).astype(np.float64) / scale).astype(np.int64)
    b = _widen(B).reshape(B.shape).astype(np.int64)
    a, b = (a.T if ta else a), (b.T if tb else b)
    assert np.array_equal(got, (a @ b).astype(np.float64) * scale)


@pytest.mark.parametrize("ta,tb", list(itertools.product((0, 1), (0, 1))))
@pytest.mark.parametrize("m,n,k,pad", [(300, 270, 64, 8), (64, 64, 4096, 8), (200, 130, 96, 3)])
def test_f16_in_place_operands_exact(gpu, ta, tb, m, n, k, pad):
    """As for bf16: k-contiguous operands with 16-byte aligned rows and k a
    multiple of 32 are read in place, the rest packed; integer inputs, NaN
    padding, ragged m and n, k = 4096 on a 64 x 64 result."""
    rng = np.random.default_rng(m + n + k + pad)
    sa = (k, m) if ta else (m, k)
    sb = (n, k) if tb else (k, n)
    A = _f16(rng, *sa, sa[0] + pad, ints=True)
    B = _f16(rng, *sb, sb[0] + pad, ints=True)
    tA, pA = _dev16(A)
    tB, pB = _dev16(B)
    dC = _dev32(np.full((m, n), np.nan, np.float32))
    E.gemm_f16(ta, tb, m, n, k, 1.0, pA, A.shape[0], pB, B.shape[0], 0.0, dC, m)
    got = from_dev(dC, (m, n))
    a = _widen(A[: sa[0], : sa[1]]).reshape(sa).astype(np.int64)
    b = _widen(B[: sb[0], : sb[1]]).reshape(sb).astype(np.int64)
    a, b = (a.T if ta else a), (b.T if tb else b)
    assert np.array_equal(got, (a @ b).astype(np.float32))


def test_f16_alpha_zero_and_empty(gpu):
    rng = np.random.default_rng(6)
    m, n, k = 33, 17, 40
    C0 = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    dC = _dev32(C0)
    A = np.full((m, k), NAN_F16, np.uint16, order="F")
    tA, pA = _dev16(A)
    E.gemm_f16(0, 0, m, n, k, 0.0, pA, m, pA, k, 2.0, dC, m)  # A, B not read
    assert np.array_equal(from_dev(dC, (m, n)), 2.0 * C0)
    E.gemm_f16(0, 0, m, n, 0, 1.0, pA, m, pA, 1, 0.5, dC, m)  # k == 0
    assert np.array_equal(from_dev(dC, (m, n)), C0)


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
def test_f16_large_tile_sampled_rows(gpu, ta, tb):
    """A product large enough for the 256x256 tile (ragged n), checked on 96
    sampled rows against the long-double oracle with the same bound."""
    rng = np.random.default_rng(8)
    m, n, k = 4096, 4160, 96
    sa = (k, m) if ta else (m, k)
    sb = (n, k) if tb else (k, n)
    A = _f16(rng, *sa, sa[0])
    B = _f16(rng, *sb, sb[0] + 1)
    tA, pA = _dev16(A)
    tB, pB = _dev16(B)
    dC = _dev32(np.full((m, n), np.nan, np.float32))
    E.gemm_f16(ta, tb, m, n, k, 1.0, pA, A.shape[0], pB, B.shape[0], 0.0, dC, m)
    got = from_dev(dC, (m, n)).astype(np.float64)
    rows = np.sort(rng.choice(m, 96, replace=False))
    Af = np.asfortranarray(_widen(A[: sa[0], : sa[1]]).reshape(sa))
    Bf = np.asfortranarray(_widen(B[: sb[0], : sb[1]]).reshape(sb))
    want, absprod = O.gemm_ref(bool(ta), bool(tb), 1.0, Af, Bf, rows=rows, with_abs=True)
    assert np.isfinite(got).all()
    assert np.linalg.norm(got[rows] - want) <= k * U32 * np.linalg.norm(absprod)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref (Eigen built from the reference tree) absent")
@pytest.mark.parametrize("m,n,k", [(64, 48, 512), (33, 20, 2000)])
def test_f16_no_less_accurate_than_eigen_half_product(gpu, m, n, k):
    """Eigen's CPU product of Matrix<half> (test/product_large.cpp:366) on the same
    half inputs: its half-precision result and ours (FP32, and rounded to half)
    are compared with the exact product."""
    rng = np.random.default_rng(m + n + k)
    A = _f16(rng, m, k, m)
    B = _f16(rng, k, n, k)
    Ch = np.zeros((m, n), np.uint16, order="F")
    O.ref().ref_gemm_half(m, n, k, O._ptr(A), O._ptr(B), O._ptr(Ch))
    eig = _widen(Ch).reshape(m, n).astype(np.float64)
    tA, pA = _dev16(A)
    tB, pB = _dev16(B)
    dC = _dev32(np.zeros((m, n), np.float32))
    E.gemm_f16(0, 0, m, n, k, 1.0, pA, m, pB, k, 0.0, dC, m)
    got = from_dev(dC, (m, n))
    exact = O.gemm_ref(False, False, 1.0, np.asfortranarray(_widen(A).reshape(A.shape)),
                       np.asfortranarray(_widen(B).reshape(B.shape)))
    ours_half = _widen(O.half_rtne(got)).reshape(m, n).astype(np.float64)
    e_eig = np.linalg.norm(eig - exact)
    assert np.linalg.norm(got.astype(np.float64) - exact) <= e_eig
    assert np.linalg.norm(ours_half - exact) <= e_eig
