"""Strided batched products on the grouped kernel (eh_gemm_batched_strided).

Bars: bit-exact on small-integer inputs (every partial sum is exact) for every
op combination, fixed-size 4x4 products by the tens of thousands (the
one-thread-per-product path for shapes up to 4x4x4), ragged
shapes with padded leading dimensions and gaps between products (which must
stay untouched); alpha == 0 does not read A or B, beta == 0 does not read C;
random inputs within the FP32 product bound k u || |A||B| ||_F per product."""
import itertools

import numpy as np
import pytest

import eigen_amd as E
import oracle as O

pytestmark = pytest.mark.gpu

ED = {np.float32: E.F32, np.float64: E.F64}


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _batch(rng, count, rows, cols, ld, stride, dt, ints=True):
    """Flat buffer holding `count` column-major rows x cols matrices (leading
    dimension ld) at `stride` elements apart; NaN everywhere else."""
    buf = np.full(stride * (count - 1) + ld * (cols - 1) + rows if count else 1, np.nan, dtype=dt)
    mats = []
    for p in range(count):
        m = rng.integers(-8, 9, (rows, cols)).astype(dt) if ints else rng.uniform(-1, 1, (rows, cols)).astype(dt)
        for j in range(cols):
            buf[p * stride + j * ld: p * stride + j * ld + rows] = m[:, j]
        mats.append(m)
    return buf, mats


def _unpack(buf, p, rows, cols, ld, stride):
    return np.stack([buf[p * stride + j * ld: p * stride + j * ld + rows] for j in range(cols)], axis=1)


# every op combination on ragged shapes: through the grouped kernel (also
# products up to 16x16 deeper than 64), one thread per product (up to
# 4x4x4) and one wave per product (up to 16x16, depth up to 32 or, for
# FP64, 64); the
# 20,000-product 4x4 batch (fixed-size matrices) with op N, N
CASES = [(ta, tb, m, n, k, count) for ta, tb in itertools.product((0, 1), (0, 1))
         for (m, n, k, count) in ((33, 17, 50, 100), (1, 1, 1, 5), (70, 40, 3, 7), (3, 2, 4, 1000),
                                  (16, 16, 16, 2000), (9, 13, 40, 300), (16, 5, 64, 50), (7, 16, 17, 200),
                                  (5, 1, 3, 100), (16, 16, 256, 300), (9, 5, 130, 100))] + \
    [(0, 0, 4, 4, 4, 20000)]


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("ta,tb,m,n,k,count", CASES)
def test_batched_strided_exact(gpu, dt, ta, tb, m, n, k, count):
    rng = np.random.default_rng(m + n + k + count)
    ra, ca = (k, m) if ta else (m, k)
    rb, cb = (n, k) if tb else (k, n)
    lda, ldb, ldc = ra + 1, rb, m + 2
    sa, sb, sc = lda * ca + 3, ldb * cb, ldc * n + 5
    A, am = _batch(rng, count, ra, ca, lda, sa, dt)
    B, bm = _batch(rng, count, rb, cb, ldb, sb, dt)
    C0, cm = _batch(rng, count, m, n, ldc, sc, dt)
    dC = _dev(C0)
    E.gemm_batched_strided(ED[dt], ta, tb, m, n, k, 2.0, _dev(A), lda, sa, _dev(B), ldb, sb, -1.0, dC, ldc, sc, count)
    got = _host(dC)
    assert np.array_equal(np.isnan(got), np.isnan(C0)), "elements outside the products written"
    for p in range(0, count, max(1, count // 200)):
        a = am[p].astype(np.int64)
        b = bm[p].astype(np.int64)
        want = 2 * ((a.T if ta else a) @ (b.T if tb else b)) - cm[p].astype(np.int64)
        assert np.array_equal(_unpack(got, p, m, n, ldc, sc), want.astype(dt)), p


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("size,stride,beta", [(4, 16, 0.0), (4, 16, -1.0), (4, 20, 0.5), (4, 17, -1.0), (4, 18, 0.5),
                                              (2, 4, 0.0), (2, 4, -1.0), (2, 6, 0.5), (2, 5, -1.0),
                                              (3, 9, 0.0), (3, 9, -1.0), (3, 10, 0.5)])
def test_batched_strided_dense_exact(gpu, dt, size, stride, beta):
    """Arrays of dense 4x4, 2x2 and 3x3 matrices (leading dimension = size):
    packed (stride size^2) and spaced arrays; 4x4 and 2x2 ones whose every
    column is aligned to its byte size are read as whole columns, strides
    that leave columns misaligned (17, 18 for 4x4 FP64, 5) element by
    element; 3x3 ones as runs of nine elements at any alignment; beta == 0
    with NaN in C."""
    rng = np.random.default_rng(stride + 100 * size)
    count = 30000
    e = size * size
    A, am = _batch(rng, count, size, size, size, stride, dt)
    B, bm = _batch(rng, count, size, size, size, stride, dt)
    C0, cm = _batch(rng, count, size, size, size, stride, dt)
    if beta == 0.0:
        for p in range(count):
            C0[p * stride: p * stride + e] = np.nan
    dC = _dev(C0)
    E.gemm_batched_strided(ED[dt], 0, 0, size, size, size, 3.0, _dev(A), size, stride, _dev(B), size, stride, beta, dC,
                           size, stride, count)
    got = _host(dC)
    gaps = np.ones(len(C0), bool)
    for p in range(count):
        gaps[p * stride: p * stride + e] = False
    assert np.array_equal(got[gaps], C0[gaps], equal_nan=True), "elements between the products written"
    for p in range(0, count, 97):
        want = 3 * (am[p].astype(np.int64) @ bm[p].astype(np.int64))
        if beta != 0.0:
            want = want + beta * cm[p].astype(np.float64)
        assert np.array_equal(_unpack(got, p, size, size, size, stride), want.astype(dt)), p


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("m,n,k,count", [(20, 12, 9, 30), (4, 3, 2, 300), (12, 9, 20, 30)])  # grouped; thread; wave
def test_batched_strided_alpha_beta_zero(gpu, dt, m, n, k, count):
    rng = np.random.default_rng(3)
    C0, cm = _batch(rng, count, m, n, m, m * n, dt)
    nan = np.full(m * k * count + 8, np.nan, dt)
    dC = _dev(C0)
    E.gemm_batched_strided(ED[dt], 0, 0, m, n, k, 0.0, _dev(nan), m, m * k, _dev(nan), k, k * n, 0.5, dC, m, m * n,
                           count)
    assert np.array_equal(_host(dC), 0.5 * C0)
    A, am = _batch(rng, count, m, k, m, m * k, dt)
    B, bm = _batch(rng, count, k, n, k, k * n, dt)
    dC = _dev(np.full(m * n * count, np.nan, dt))
    E.gemm_batched_strided(ED[dt], 0, 0, m, n, k, 1.0, _dev(A), m, m * k, _dev(B), k, k * n, 0.0, dC, m, m * n, count)
    got = _host(dC)
    for p in range(count):
        want = (am[p].astype(np.int64) @ bm[p].astype(np.int64)).astype(dt)
        assert np.array_equal(_unpack(got, p, m, n, m, m * n), want)


@pytest.mark.parametrize("m,n,k,count", [(100, 90, 500, 40), (16, 16, 64, 300), (16, 12, 300, 200)])  # grouped; wave
def test_batched_strided_random_bound(gpu, m, n, k, count):
    rng = np.random.default_rng(4)
    A, am = _batch(rng, count, m, k, m, m * k, np.float32, ints=False)
    B, bm = _batch(rng, count, k, n, k, k * n, np.float32, ints=False)
    dC = _dev(np.zeros(m * n * count, np.float32))
    E.gemm_batched_strided(E.F32, 0, 0, m, n, k, 1.0, _dev(A), m, m * k, _dev(B), k, k * n, 0.0, dC, m, m * n, count)
    got = _host(dC)
    for p in range(0, count, 7):
        want, absprod = O.gemm_ref(False, False, 1.0, np.asfortranarray(am[p]), np.asfortranarray(bm[p]),
                                   with_abs=True)
        err = np.linalg.norm(_unpack(got, p, m, n, m, m * n).astype(np.float64) - want)
        assert err <= k * 2.0 ** -24 * np.linalg.norm(absprod)
