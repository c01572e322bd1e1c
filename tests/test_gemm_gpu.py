"""Parity of the HIP products with the oracle (runs on an MI355X).

Exact cases use small-integer inputs, for which every partial sum is exact
in FP32/FP64, so the GPU result must equal the long-double oracle bit for
bit whatever the summation order: this pins the MFMA operand/output
layouts, the LDS staging, transposes, strides, edges and split-K.  Random
cases are held to the normwise bound of test/product_test_helpers.h:56-74
(lambda = 5, two products compared).
"""
import ctypes as C

import numpy as np
import pytest

import eigen_amd as E
import oracle as O
from helpers import empty_dev, from_dev, op_shape, to_dev, unit_roundoff

pytestmark = pytest.mark.gpu

DT = {np.float32: E.F32, np.float64: E.F64}


def run_gemm(dtype, ta, tb, m, n, k, alpha, A, B, beta, C0, prec=E.PREC_AUTO, config=None, lda=None, ldb=None,
             ldc=None):
    dA, dB = to_dev(A), to_dev(B)
    dC = to_dev(C0) if C0 is not None else empty_dev(m * n, dtype)
    E.gemm(DT[dtype], int(ta), int(tb), m, n, k, alpha, dA, lda or A.shape[0], dB, ldb or B.shape[0], beta, dC,
           ldc or m, prec=prec, config=config)
    return from_dev(dC, (ldc or m, n))[:m]


def int_operands(dtype, ta, tb, m, n, k, seed=1):
    A = O.fill(op_shape(ta, m, k), dtype, seed, O.INT, -8, 8)
    B = O.fill(op_shape(tb, k, n), dtype, seed + 1, O.INT, -8, 8)
    C0 = O.fill((m, n), dtype, seed + 2, O.INT, -8, 8)
    return A, B, C0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("shape", [(1, 1, 1), (7, 5, 3), (33, 31, 47), (64, 64, 64), (65, 129, 100),
                                   (130, 70, 300), (257, 255, 33)])
def test_exact_integer_products(gpu, dtype, ta, tb, shape):
    m, n, k = shape
    A, B, C0 = int_operands(dtype, ta, tb, m, n, k)
    want = O.gemm_ref(ta, tb, 2.0, A, B, -1.0, C0)
    got = run_gemm(dtype, ta, tb, m, n, k, 2.0, A, B, -1.0, C0)
    assert np.array_equal(got.astype(np.float64), want)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("config", range(24))
def test_every_config_exact(gpu, dtype, config):
    m, n, k = 300, 260, 520
    for ta, tb in [(False, False), (True, True)]:
        A, B, C0 = int_operands(dtype, ta, tb, m, n, k, seed=config + 10)
        want = O.gemm_ref(ta, tb, 1.0, A, B, 1.0, C0)
        got = run_gemm(dtype, ta, tb, m, n, k, 1.0, A, B, 1.0, C0, prec=E.PREC_PLAIN, config=config)
        assert np.array_equal(got.astype(np.float64), want), (config, ta, tb)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_padded_strides_and_unaligned_pointers(gpu, dtype):
    import torch
    m, n, k = 77, 45, 61
    lda, ldb, ldc = 83, 70, 91
    A = O.fill((lda, k), dtype, 3, O.INT, -8, 8)
    B = O.fill((ldb, n), dtype, 4, O.INT, -8, 8)
    C0 = O.fill((ldc, n), dtype, 5, O.INT, -8, 8)
    want = O.gemm_ref(False, False, 1.0, np.asfortranarray(A[:m]), np.asfortranarray(B[:k]), 1.0,
                      np.asfortranarray(C0[:m]))
    got = run_gemm(dtype, False, False, m, n, k, 1.0, A, B, 1.0, C0, lda=lda, ldb=ldb, ldc=ldc)
    assert np.array_equal(got.astype(np.float64), want)
    # rows of C beyond m (inside ldc) are untouched
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C0)
    E.gemm(DT[dtype], E.OP_N, E.OP_N, m, n, k, 1.0, dA, lda, dB, ldb, 1.0, dC, ldc)
    full = from_dev(dC, (ldc, n))
    assert np.array_equal(full[m:], C0[m:])
    # pointers offset by one element (not 16-byte aligned), transposed operands
    At = O.fill((k, m), dtype, 6, O.INT, -8, 8)
    Bt = O.fill((n, k), dtype, 7, O.INT, -8, 8)
    esz = np.dtype(dtype).itemsize
    dA = torch.cat([torch.zeros(1, dtype=to_dev(At).dtype, device="cuda:0"), to_dev(At)])
    dB = torch.cat([torch.zeros(1, dtype=dA.dtype, device="cuda:0"), to_dev(Bt)])
    dC = to_dev(np.zeros((m, n), dtype, order="F"))
    E.gemm(DT[dtype], E.OP_T, E.OP_T, m, n, k, 1.0, dA.data_ptr() + esz, k, dB.data_ptr() + esz, n, 0.0, dC, m)
    assert np.array_equal(from_dev(dC, (m, n)).astype(np.float64), O.gemm_ref(True, True, 1.0, At, Bt))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_blas_semantics_alpha_beta(gpu, dtype):
    m, n, k = 40, 30, 20
    C0 = O.fill((m, n), dtype, 1)
    nanA = np.full((m, k), np.nan, dtype, order="F")
    infB = np.full((k, n), np.inf, dtype, order="F")
    # alpha == 0: A and B are not read; C = beta*C
    got = run_gemm(dtype, False, False, m, n, k, 0.0, nanA, infB, 0.5, C0)
    assert np.array_equal(got, (dtype(0.5) * C0))
    got = run_gemm(dtype, False, False, m, n, k, -0.0, nanA, infB, 1.0, C0)
    assert np.array_equal(got, C0)
    # beta == 0: NaN in C is overwritten
    A = O.fill((m, k), dtype, 2, O.INT, -8, 8)
    B = O.fill((k, n), dtype, 3, O.INT, -8, 8)
    nanC = np.full((m, n), np.nan, dtype, order="F")
    got = run_gemm(dtype, False, False, m, n, k, 1.0, A, B, 0.0, nanC)
    assert np.array_equal(got.astype(np.float64), O.gemm_ref(False, False, 1.0, A, B))
    # k == 0: C = beta*C
    got = run_gemm(dtype, False, False, m, n, 0, 1.0, A, B, 2.0, C0, lda=m, ldb=1)
    assert np.array_equal(got, dtype(2.0) * C0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_random_within_bound(gpu, dtype, ta, tb):
    m = n = k = 512  # the 512^3 parity configuration
    A = O.fill(op_shape(ta, m, k), dtype, 21)
    B = O.fill(op_shape(tb, k, n), dtype, 22)
    ref, absref = O.gemm_ref(ta, tb, 1.0, A, B, with_abs=True)
    got = run_gemm(dtype, ta, tb, m, n, k, 1.0, A, B, 0.0, None)
    u = unit_roundoff(dtype)
    err = np.linalg.norm(got.astype(np.float64) - ref)
    assert err <= O.higham_mary_bound(absref, k, u)
    nw = O.normwise_error(got, ref, k, u, np.linalg.norm(A), np.linalg.norm(B))
    assert nw <= 1.0


@pytest.mark.skipif(not O.ref_available(), reason="Eigen reference build absent")
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_matches_eigen_cpu_gebp(gpu, dtype):
    # C = A*B, 512^3, U[-1,1] from a seed, C zero-initialised, Eigen's C.noalias() = A*B on the CPU
    m = n = k = 512
    A = O.fill((m, k), dtype, 1)
    B = O.fill((k, n), dtype, 2)
    eig = np.zeros((m, n), dtype, order="F")
    O.eigen_gemm(False, False, 1.0, A, B, eig)
    got = run_gemm(dtype, False, False, m, n, k, 1.0, A, B, 0.0, None)
    ref, absref = O.gemm_ref(False, False, 1.0, A, B, with_abs=True)
    u = unit_roundoff(dtype)
    bound = O.higham_mary_bound(absref, k, u)
    assert np.linalg.norm(got.astype(np.float64) - eig) <= bound
    assert np.linalg.norm(eig.astype(np.float64) - ref) <= bound


def test_device_fill_matches_oracle(gpu):
    n = 100003
    for dtype, dist, p0, p1 in [(np.float64, O.UNIFORM, -1, 1), (np.float32, O.UNIFORM, -1, 1),
                                (np.float64, O.INT, -8, 8), (np.float32, O.INT, -8, 8)]:
        t = empty_dev(n, dtype)
        E.fill(DT[dtype], t, n, 1234, dist, p0, p1)
        got = from_dev(t, (n, 1))[:, 0]
        want = O.fill((n, 1), dtype, 1234, dist, p0, p1)[:, 0]
        assert np.array_equal(got, want)
    # normal / wide use libm-style transcendental functions: equal to a few ulp
    for dtype, dist, p0, p1 in [(np.float64, O.NORMAL, 0, 1), (np.float32, O.WIDE, -8, 8)]:
        t = empty_dev(n, dtype)
        E.fill(DT[dtype], t, n, 99, dist, p0, p1)
        got = from_dev(t, (n, 1))[:, 0].astype(np.float64)
        want = O.fill((n, 1), dtype, 99, dist, p0, p1)[:, 0].astype(np.float64)
        assert np.allclose(got, want, rtol=8 * np.finfo(dtype).eps, atol=1e-12)


def test_device_split_matches_oracle(gpu):
    import torch
    x = O.fill((300000, 1), np.float32, 5, O.WIDE, -40, 40)[:, 0]
    special = np.float32([0.0, -0.0, 1.0, 3.4e38, -3.39e38, 3.3895e38, np.inf, -np.inf, 1e-40, -1e-45,
                          0.33333, 65504.0])
    x = np.concatenate([x, special])
    n = x.size
    dx = torch.from_numpy(x).cuda()
    ts = [torch.empty(n, dtype=torch.int16, device="cuda:0") for _ in range(3)]
    for terms in (1, 2, 3):
        E.split_bf16(dx, n, terms, *ts[:terms])
        torch.cuda.synchronize()
        want = O.bf16_split(x, terms)
        for t, w in zip(ts[:terms], want):
            assert np.array_equal(t.cpu().numpy().view(np.uint16), w)
    # NaN stays NaN in the leading term
    dn = torch.tensor([float("nan")], device="cuda:0")
    E.split_bf16(dn, 1, 3, *ts)
    torch.cuda.synchronize()
    assert (int(ts[0][0].item()) & 0x7FFF) > 0x7F80


# packed configurations of the split path (tile * 4 + split index): planner
# default, 128x128 (3 stages), 256x128 with 8 waves, 256x128 with 4 waves,
# and that with split K 2, 256x256 with 16-deep stages, and that with split K 2
SPLIT_CONFIGS = [None, 0, 8, 4, 5, 12, 13]


_SPLIT_CASES = {}


def _split_case(ta, tb, shape):
    """Inputs of one split-product case with the oracle's values: the kept set
    of each split (2 and 3 terms) and every single cross product on its own.
    Cached over the tile configurations, which share inputs."""
    key = (ta, tb, shape)
    if key not in _SPLIT_CASES:
        m, n, k = shape
        A = O.fill(op_shape(ta, m, k), np.float32, 31, O.WIDE, -8, 8)
        B = O.fill(op_shape(tb, k, n), np.float32, 32, O.WIDE, -8, 8)
        C0 = O.fill((m, n), np.float32, 33)
        _, absref = O.gemm_ref(ta, tb, 1.25, A, B, with_abs=True)
        want = {t: O.gemm_split_ref(t, ta, tb, 1.25, A, B, 0.5, C0) for t in (2, 3)}
        single = {t: {b: O.gemm_split_ref(t, ta, tb, 1.25, A, B, pairs=1 << b)
                      for b in range(9) if O.split_kept_pairs(t) >> b & 1} for t in (2, 3)}
        _SPLIT_CASES[key] = (A, B, C0, absref, want, single)
    return _SPLIT_CASES[key]


@pytest.mark.parametrize("config", SPLIT_CONFIGS)
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("shape", [(1, 1, 1), (33, 65, 31), (200, 130, 257), (300, 280, 1100)])
def test_split_product_vs_split_oracle(gpu, ta, tb, shape, config):
    """The device result is the oracle's kept-set value up to FP32 accumulation.

    The tolerance is set by the wrong answers it must exclude: for each split
    the distance from the kept-set value to every wrong kept set (one kept
    cross product left out, or the other split's set) is computed by the
    oracle, and the device result must lie within a quarter of the smallest of
    them (within an eighth of the other split's value).  A kernel that kept
    the wrong cross products therefore fails.  The smallest of these distances
    is the mid*mid product (relative size ~2^-19.6 per product); an emulation
    of FP32 accumulation in 16-deep MFMA steps puts the accumulation error at
    up to 0.14 of it at k = 1100.  The normwise bound of
    test/product_test_helpers.h:56-74 is checked as well; it is 80-440 times
    the actual accumulation error here and cannot tell the kept sets apart.
    The 1x1x1 case is checked against the bound only (a single product's low
    terms may be zero)."""
    m, n, k = shape
    A, B, C0, absref, want, single = _split_case(ta, tb, shape)
    for terms, prec in ((3, E.PREC_SPLIT3), (2, E.PREC_SPLIT2)):
        got = run_gemm(np.float32, ta, tb, m, n, k, 1.25, A, B, 0.5, C0, prec=prec, config=config).astype(np.float64)
        err = np.linalg.norm(got - want[terms])
        assert err <= O.higham_mary_bound(absref, k, unit_roundoff(np.float32)) + 1e-30
        if m * n * k < 1000:
            continue
        other = 5 - terms
        d_other = np.linalg.norm(want[other] - want[terms])
        assert err <= d_other / 8, (terms, err, d_other)
        for b, v in single[terms].items():  # that pair left out: want - v
            d = np.linalg.norm(v)
            assert err <= d / 4, (terms, b, err, d)


def test_split3_error_no_worse_than_plain(gpu):
    # wide dynamic range entries sign*U(1,2)*2^U(-8,8), errors against the long-double oracle
    m = n = k = 1024
    A = O.fill((m, k), np.float32, 41, O.WIDE, -8, 8)
    B = O.fill((k, n), np.float32, 42, O.WIDE, -8, 8)
    rows = np.arange(0, m, 8)
    ref = O.gemm_ref(False, False, 1.0, A, B, rows=rows)
    u = unit_roundoff(np.float32)
    nA, nB = np.linalg.norm(A), np.linalg.norm(B)
    errs = {}
    for name, prec in (("plain", E.PREC_PLAIN), ("split3", E.PREC_SPLIT3), ("split2", E.PREC_SPLIT2)):
        got = run_gemm(np.float32, False, False, m, n, k, 1.0, A, B, 0.0, None, prec=prec)
        errs[name] = O.normwise_error(got[rows], ref, k, u, nA, nB)
    assert errs["split3"] <= errs["plain"] * 1.05, errs
    assert errs["split2"] > errs["split3"], errs


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_blas_entry_device_and_host_pointers(gpu, dtype):
    L = E.blas()
    m, n, k = 70, 50, 90
    A = O.fill((k, m), dtype, 51, O.INT, -8, 8)  # used as op(A) = A^T
    B = O.fill((k, n), dtype, 52, O.INT, -8, 8)
    C0 = O.fill((m, n), dtype, 53, O.INT, -8, 8)
    want = O.gemm_ref(True, False, 1.5, A, B, 0.5, C0)
    ct = C.c_float if dtype == np.float32 else C.c_double
    fn = L.sgemm_ if dtype == np.float32 else L.dgemm_
    i = lambda v: C.byref(C.c_int(v))
    al, be = ct(1.5), ct(0.5)
    # host pointers (staged through device buffers)
    Ch = C0.copy(order="F")
    fn(b"T", b"N", i(m), i(n), i(k), C.byref(al), A.ctypes.data_as(C.c_void_p), i(k),
       B.ctypes.data_as(C.c_void_p), i(k), C.byref(be), Ch.ctypes.data_as(C.c_void_p), i(m))
    assert np.array_equal(Ch.astype(np.float64), want)
    # device pointers (runs in place)
    import torch
    torch.cuda.synchronize()
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C0)
    torch.cuda.synchronize()
    fn(b"T", b"N", i(m), i(n), i(k), C.byref(al), C.c_void_p(dA.data_ptr()), i(k), C.c_void_p(dB.data_ptr()),
       i(k), C.byref(be), C.c_void_p(dC.data_ptr()), i(m))
    assert np.array_equal(from_dev(dC, (m, n)).astype(np.float64), want)


@pytest.mark.parametrize("dtype,prec", [(np.float32, E.PREC_PLAIN), (np.float32, E.PREC_SPLIT3),
                                        (np.float64, E.PREC_AUTO)])
@pytest.mark.parametrize("alpha,beta", [(1.5, 0.5), (2.0, 0.0), (0.0, 0.5), (0.0, 0.0)])
def test_strided_destination_matches_contiguous(gpu, dtype, prec, alpha, beta):
    """eh_gemm_strided (result increment incc) gives, bit for bit, what eh_gemm
    gives on a contiguous C, touches only the strided elements, and with
    beta == 0 never reads C."""
    import torch
    m, n, k, incc = 130, 70, 200, 3
    ldc = m * incc + 5
    A = O.fill((m, k), dtype, 61, O.WIDE, -4, 4)
    B = O.fill((k, n), dtype, 62, O.WIDE, -4, 4)
    C0 = O.fill((m, n), dtype, 63)
    if beta == 0.0:
        C0 = np.full((m, n), np.nan, dtype=dtype)  # must not be read
    want = run_gemm(dtype, False, False, m, n, k, alpha, A, B, beta, C0, prec=prec)
    sentinel = dtype(-99.0)
    big = np.full(ldc * n, sentinel, dtype=dtype)
    rows, cols = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    big[(rows * incc + cols * ldc).ravel()] = C0.ravel()
    dbig = torch.from_numpy(big).cuda()
    dA, dB = to_dev(A), to_dev(B)
    E.gemm_strided(DT[dtype], 0, 0, m, n, k, alpha, dA, m, dB, k, beta, dbig, incc, ldc, prec=prec)
    torch.cuda.synchronize()
    out = dbig.cpu().numpy()
    got = out[(rows * incc + cols * ldc).ravel()].reshape(m, n)
    assert np.array_equal(got, want, equal_nan=False)
    mask = np.ones(ldc * n, dtype=bool)
    mask[(rows * incc + cols * ldc).ravel()] = False
    assert np.all(out[mask] == sentinel)


def _padded(a: np.ndarray, pad: int) -> np.ndarray:
    """a stored with `pad` extra rows (leading dimension a.shape[0] + pad)."""
    full = np.zeros((a.shape[0] + pad, a.shape[1]), dtype=a.dtype, order="F")
    full[:a.shape[0]] = a
    return full


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("shape", [(300, 1, 257), (1, 300, 257), (1, 1, 1000), (5000, 1, 3000), (1, 4100, 2000),
                                   (17, 1, 20000), (2, 1, 3)])
def test_vector_products_exact(gpu, dtype, ta, tb, shape):
    """Products with one output column or row take the matrix-vector kernels
    (split K included); small-integer operands make every order exact."""
    m, n, k = shape
    A, B, C0 = int_operands(dtype, ta, tb, m, n, k, seed=70)
    Ap, Bp, Cp = _padded(A, 3), _padded(B, 5), _padded(C0, 2)
    for alpha, beta in ((1.5, 0.5), (-2.0, 0.0)):
        want = O.gemm_ref(ta, tb, alpha, A, B, beta, C0)
        got = run_gemm(dtype, ta, tb, m, n, k, alpha, Ap, Bp, beta, Cp, lda=Ap.shape[0], ldb=Bp.shape[0],
                       ldc=Cp.shape[0])
        assert np.array_equal(got.astype(np.float64), want), (alpha, beta)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("ta", [False, True])
def test_vector_product_random_within_bound(gpu, dtype, ta):
    m, k = 16384, 16384
    A = O.fill((k, m) if ta else (m, k), dtype, 81, O.UNIFORM, -1, 1)
    x = O.fill((k, 1), dtype, 82, O.UNIFORM, -1, 1)
    got = run_gemm(dtype, ta, False, m, 1, k, 1.0, A, x, 0.0, None, prec=E.PREC_PLAIN)
    rows = np.arange(0, m, 64)
    ref, absref = O.gemm_ref(ta, False, 1.0, A, x, rows=rows, with_abs=True)
    err = np.linalg.norm(got[rows].astype(np.float64) - ref)
    assert err <= O.higham_mary_bound(absref, k, unit_roundoff(dtype)) + 1e-300


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("uplo", [E.UPPER, E.LOWER])
@pytest.mark.parametrize("trans", [False, True])
@pytest.mark.parametrize("n,k", [(1, 1), (65, 47), (130, 300), (300, 1), (257, 2100), (256, 20000)])
def test_syrk_triangle_exact(gpu, dtype, uplo, trans, n, k):
    """eh_syrk updates exactly the uplo triangle with alpha*op(A)*op(A)^T + beta*C
    (small-integer operands: exact) and leaves the other triangle untouched."""
    A = O.fill((k, n) if trans else (n, k), dtype, 90 + n, O.INT, -8, 8)
    Ap = _padded(A, 3)
    C0 = O.fill((n, n), dtype, 91, O.INT, -8, 8)
    iu = np.triu_indices(n) if uplo == E.UPPER else np.tril_indices(n)
    other = np.ones((n, n), dtype=bool)
    other[iu] = False
    for alpha, beta in ((1.5, 0.5), (2.0, 0.0), (0.0, -1.0)):
        Cin = C0.copy(order="F")
        if beta == 0.0:
            Cin[iu] = np.nan  # the updated triangle must not be read
        Cp = _padded(Cin, 2)
        dA, dC = to_dev(Ap), to_dev(Cp)
        E.syrk(DT[dtype], uplo, int(trans), n, k, alpha, dA, Ap.shape[0], beta, dC, Cp.shape[0])
        got = from_dev(dC, (n + 2, n))[:n]
        full = O.gemm_ref(trans, not trans, alpha, A, A, beta, Cin) if beta != 0.0 else \
            O.gemm_ref(trans, not trans, alpha, A, A)
        assert np.array_equal(got[iu].astype(np.float64), full[iu]), (alpha, beta)
        assert np.array_equal(got[other], Cin[other], equal_nan=True), (alpha, beta)
