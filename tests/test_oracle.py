"""Pins the CPU oracle before anything is checked against it (CPU only).

Each oracle function is checked against something it does not compute
itself: golden vectors from the reference's own tests, Eigen compiled from
the reference tree (oracle/_ref, when present), closed forms and exact
small-integer arithmetic.
"""
import numpy as np
import pytest

import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref (Eigen built from the reference tree) absent")


# ---------------------------------------------------------------- generator
def test_splitmix64_known_answer():
    # splitmix64 seeded with 0: first output is the published 0xE220A8397B1DCDAF,
    # second 0x6E789E6AA1B965F4 (state advanced by the golden gamma each step).
    L = O.lib()
    assert L.orc_splitmix64_at(0, 0) == 0xE220A8397B1DCDAF
    assert L.orc_splitmix64_at(0, 1) == 0x6E789E6AA1B965F4


def test_fill_distributions():
    u = O.fill((1000, 100), np.float64, 7, O.UNIFORM, -1, 1)
    assert u.min() >= -1 and u.max() <= 1 and abs(u.mean()) < 0.01
    g = O.fill((1000, 100), np.float64, 8, O.NORMAL, 0, 1)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1) < 0.01
    w = O.fill((1000, 100), np.float32, 9, O.WIDE, -8, 8)
    a = np.abs(w.astype(np.float64))
    assert a.min() >= 2.0 ** -8 and a.max() <= 2.0 ** 9
    lg = np.log2(a)
    assert abs(lg.mean() - (2 - 1 / np.log(2))) < 0.05  # E[log2 U(1,2)] = 2 - 1/ln2, E[U(-8,8)] = 0
    i = O.fill((100, 100), np.float64, 10, O.INT, -8, 8)
    assert np.all(i == np.round(i)) and i.min() == -8 and i.max() == 8


# ---------------------------------------------------------------- bfloat16
def test_bf16_golden_vectors():
    # test/bfloat16_float.cpp:59-84
    cases = {1.0: 0x3F80, 0.5: 0x3F00, 0.33333: 0x3EAB, 3.38e38: 0x7F7E, 3.40e38: 0x7F80,
             -1.0: 0xBF80, 0.0: 0x0000, 2.0: 0x4000, 3.0: 0x4040, 12.0: 0x4140}
    got = O.bf16_rtne(np.array(list(cases), np.float32))
    assert [int(g) for g in got] == list(cases.values())
    # round-half-to-even: midpoints of 0x3c00/0x3c01 and 0x3c01/0x3c02
    v = O.bf16_to_f32(np.array([0x3C00, 0x3C01, 0x3C02], np.uint16))
    mids = np.array([0.5 * (v[0] + v[1]), 0.5 * (v[1] + v[2])], np.float32)
    assert [int(g) for g in O.bf16_rtne(mids)] == [0x3C00, 0x3C02]
    assert int(O.bf16_rtne(np.array([np.nan], np.float32))[0]) == 0x7FC0
    assert int(O.bf16_rtne(np.array([-np.nan], np.float32))[0]) == 0xFFC0


@needs_ref
def test_bf16_matches_eigen_sweep():
    # strided sweep over all 2^32 float bit patterns (every 4099th) vs Eigen's bfloat16(float)
    bits = (np.arange(0, 2 ** 32, 4099, dtype=np.uint64)).astype(np.uint32)
    x = bits.view(np.float32)
    ours = O.bf16_rtne(x)
    ref = np.empty_like(ours)
    O.ref().ref_bf16_rtne_n(O._ptr(np.ascontiguousarray(x)), x.size, O._ptr(ref))
    assert np.array_equal(ours, ref)


def test_bf16_split_is_exact():
    x = O.fill((200000, 1), np.float32, 11, O.WIDE, -30, 30).ravel()
    x = np.concatenate([x, np.float32([3.4e38, -3.39e38, 1e-30, 0.0, -0.0, 1.0])])
    t0, t1, t2 = O.bf16_split(x, 3)
    s = (O.bf16_to_f32(t0).astype(np.float64) + O.bf16_to_f32(t1).astype(np.float64)
         + O.bf16_to_f32(t2).astype(np.float64))
    assert np.array_equal(s, x.astype(np.float64))
    # two terms leave a residual below 2^-16 relative
    h0, h1 = O.bf16_split(x, 2)
    s2 = O.bf16_to_f32(h0).astype(np.float64) + O.bf16_to_f32(h1).astype(np.float64)
    nz = x != 0
    rel = np.abs(s2[nz] - x[nz]) / np.abs(x[nz].astype(np.float64))
    assert rel.max() <= 2.0 ** -16
    # leading term is plain RTNE except where RTNE overflows a finite value
    ok = np.isfinite(O.bf16_to_f32(O.bf16_rtne(x)))
    assert np.array_equal(t0[ok], O.bf16_rtne(x)[ok])
    # non-finite inputs: t0 carries the value, residuals are zero
    inf3 = O.bf16_split(np.float32([np.inf, -np.inf]), 3)
    assert [int(v) for v in inf3[0]] == [0x7F80, 0xFF80] and not inf3[1].any() and not inf3[2].any()


# ---------------------------------------------------------------- IEEE half
def test_half_golden_vectors():
    # test/half_float.cpp:33-52 (conversion from float, denormals, ties to even)
    cases = [(1.0, 0x3C00), (0.5, 0x3800), (0.33333, 0x3555), (0.0, 0x0000), (-0.0, 0x8000), (65504.0, 0x7BFF),
             (65536.0, 0x7C00), (-5.96046e-08, 0x8001), (5.96046e-08, 0x0001), (1.19209e-07, 0x0002),
             (-1.0, 0xBC00), (2.0, 0x4000), (3.0, 0x4200)]
    got = O.half_rtne(np.array([c[0] for c in cases], np.float32))
    assert [int(g) for g in got] == [c[1] for c in cases]
    v = O.half_to_f32(np.array([0x3C00, 0x3C01, 0x3C02], np.uint16))
    mids = np.array([0.5 * (v[0] + v[1]), 0.5 * (v[1] + v[2])], np.float32)
    assert [int(g) for g in O.half_rtne(mids)] == [0x3C00, 0x3C02]
    # widening: test/half_float.cpp:67-73
    w = O.half_to_f32(np.array([0x0000, 0x3C00, 0x8001, 0x0001, 0x0002, 0x7BFF, 0x7C00, 0xFC00], np.uint16))
    assert list(w[:2]) == [0.0, 1.0] and w[2] == -(2.0 ** -24) and w[3] == 2.0 ** -24 and w[4] == 2.0 ** -23
    assert w[5] == 65504.0 and np.isposinf(w[6]) and np.isneginf(w[7])


def test_half_matches_ieee_conversion_sweep():
    # every 4099th float bit pattern and all half values' neighbourhoods, against
    # numpy's IEEE float32 -> float16 conversion (round to nearest even)
    bits = np.arange(0, 2 ** 32, 4099, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    with np.errstate(over="ignore", invalid="ignore"):
        ref = x.astype(np.float16).view(np.uint16)
    ours = O.half_rtne(x)
    fin = ~np.isnan(x)
    assert np.array_equal(ours[fin], ref[fin])
    assert np.all((ours[~fin] & 0x7E00) == 0x7E00)  # NaN stays NaN (quiet)
    # widening is exact and inverts the rounding on every half value
    h = np.arange(0, 2 ** 16, dtype=np.uint32).astype(np.uint16)
    f = O.half_to_f32(h)
    fin = ~np.isnan(f)
    assert np.array_equal(f[fin], h[fin].view(np.float16).astype(np.float32))
    assert np.array_equal(O.half_rtne(f[fin]), h[fin])


@needs_ref
def test_half_matches_eigen_sweep():
    # Eigen's own Eigen::half(float) on a strided sweep of all float bit patterns
    bits = (np.arange(0, 2 ** 32, 4099, dtype=np.uint64)).astype(np.uint32)
    x = np.ascontiguousarray(bits.view(np.float32))
    ref = np.empty(x.shape, np.uint16)
    O.ref().ref_half_rtne_n(O._ptr(x), x.size, O._ptr(ref))
    ours = O.half_rtne(x)
    fin = ~np.isnan(x)
    assert np.array_equal(ours[fin], ref[fin])


# ---------------------------------------------------------------- split product value
def test_split_kept_pairs_are_the_published_sets():
    # bf16x3 keeps hi*hi, hi*mid, mid*hi; bf16x6 adds hi*lo, lo*hi, mid*mid
    # (Henry, Tang & Heinecke 2019): the pairs p + q < terms, bit 3p + q
    bits = lambda pairs: {(p, q) for p in range(3) for q in range(3) if O.split_kept_pairs(pairs) >> (3 * p + q) & 1}
    assert bits(1) == {(0, 0)}
    assert bits(2) == {(0, 0), (0, 1), (1, 0)}
    assert bits(3) == {(0, 0), (0, 1), (1, 0), (0, 2), (2, 0), (1, 1)}


@pytest.mark.parametrize("terms", [1, 2, 3])
def test_split_ref_bf16_exact_inputs_equal_the_product(terms):
    # integers |x| <= 256 are bfloat16 values: every lower term is zero and the
    # split product is the exact product, bit for bit
    A = O.fill((37, 53), np.float32, 71, O.INT, -256, 256)
    B = O.fill((53, 29), np.float32, 72, O.INT, -256, 256)
    C0 = O.fill((37, 29), np.float32, 73, O.INT, -8, 8)
    want = O.gemm_ref(False, False, 1.25, A, B, 0.5, C0)
    assert np.array_equal(O.gemm_split_ref(terms, False, False, 1.25, A, B, 0.5, C0), want)


def test_split_ref_all_nine_pairs_is_the_exact_product():
    # t0 + t1 + t2 == x exactly, so keeping all nine cross products gives the
    # product itself (to long-double summation rounding)
    A = O.fill((40, 300), np.float32, 74, O.WIDE, -8, 8)
    B = O.fill((300, 30), np.float32, 75, O.WIDE, -8, 8)
    ex, ab = O.gemm_ref(False, False, 1.0, A, B, with_abs=True)
    s9 = O.gemm_split_ref(3, False, False, 1.0, A, B, pairs=0x1FF)
    assert np.linalg.norm(s9 - ex) <= 2.0 ** -50 * np.linalg.norm(ab)


@pytest.mark.parametrize("shape", [(64, 64, 1), (40, 30, 200), (20, 20, 2000)])
def test_split_ref_error_bands(shape):
    # What each kept set leaves out, relative to || |A||B| ||_F, on wide-range
    # inputs.  Upper ends are deterministic bounds from the split: a term is at
    # most 2^-8 of the one before, and two terms leave at most 2^-16 of x, so
    # 1 term drops <= 2^-7, 2 terms <= 3*2^-16 < 2^-14, 3 terms (t1 t2 + t2 t1
    # + t2 t2) <= 2^-23 < 2^-22.  Lower ends hold for random inputs (a split
    # that rounded terms wrongly or kept fewer bits would leave the band); a
    # kept set with a pair missing rises out of its band (checked below).  The
    # sets themselves are pinned by test_split_kept_pairs_are_the_published_sets.
    m, n, k = shape
    A = O.fill((m, k), np.float32, 61, O.WIDE, -8, 8)
    B = O.fill((k, n), np.float32, 62, O.WIDE, -8, 8)
    ex, ab = O.gemm_ref(False, False, 1.0, A, B, with_abs=True)
    nab = np.linalg.norm(ab)
    bands = {1: (2.0 ** -13, 2.0 ** -7), 2: (2.0 ** -22, 2.0 ** -14), 3: (2.0 ** -32, 2.0 ** -22)}
    for terms, (lo, hi) in bands.items():
        rel = np.linalg.norm(O.gemm_split_ref(terms, False, False, 1.0, A, B) - ex) / nab
        assert lo <= rel <= hi, (terms, np.log2(rel))
    # each missing pair of the 3-term set lifts the error out of its band
    kept = O.split_kept_pairs(3)
    for b in range(9):
        if kept >> b & 1:
            rel = np.linalg.norm(O.gemm_split_ref(3, False, False, 1.0, A, B, pairs=kept ^ (1 << b)) - ex) / nab
            assert rel > 2.0 ** -22, (b, np.log2(rel))


def test_split_ref_single_pair_is_the_term_product():
    # k = 1: the value of pair (p, q) alone is alpha * t_p(a_i) * t_q(b_j),
    # with p indexing the terms of A and q those of B
    a = O.fill((6, 1), np.float32, 81, O.WIDE, -8, 8)
    b = O.fill((1, 5), np.float32, 82, O.WIDE, -8, 8)
    ta = [O.bf16_to_f32(t).astype(np.float64) for t in O.bf16_split(a.ravel(), 3)]
    tb = [O.bf16_to_f32(t).astype(np.float64) for t in O.bf16_split(b.ravel(), 3)]
    for p in range(3):
        for q in range(3):
            got = O.gemm_split_ref(3, False, False, 0.75, a, b, pairs=1 << (3 * p + q))
            assert np.array_equal(got, 0.75 * np.outer(ta[p], tb[q])), (p, q)


def test_split_ref_transposes_rows_and_beta():
    m, n, k = 23, 17, 41
    A = O.fill((m, k), np.float32, 91, O.WIDE, -8, 8)
    B = O.fill((k, n), np.float32, 92, O.WIDE, -8, 8)
    C0 = O.fill((m, n), np.float32, 93)
    want = O.gemm_split_ref(3, False, False, 1.5, A, B, 0.5, C0)
    At, Bt = np.asfortranarray(A.T), np.asfortranarray(B.T)
    assert np.array_equal(O.gemm_split_ref(3, True, True, 1.5, At, Bt, 0.5, C0), want)
    assert np.array_equal(O.gemm_split_ref(3, False, False, 1.5, A, B, 0.5, C0, rows=[7, 2, 22]), want[[7, 2, 22]])
    plain = O.gemm_split_ref(3, False, False, 1.5, A, B)
    assert np.allclose(want - plain, 0.5 * C0.astype(np.float64), rtol=0, atol=1e-9 * np.abs(plain).max())


# ---------------------------------------------------------------- reference product
def test_gemm_ref_brute_force_tiny():
    rng = np.random.default_rng(0)
    A = np.asfortranarray(rng.integers(-3, 4, (3, 4)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-3, 4, (4, 2)).astype(np.float64))
    C0 = np.asfortranarray(rng.integers(-3, 4, (3, 2)).astype(np.float64))
    expect = np.zeros((3, 2))
    for i in range(3):
        for j in range(2):
            expect[i, j] = 1.5 * sum(A[i, k] * B[k, j] for k in range(4)) + 0.5 * C0[i, j]
    got = O.gemm_ref(False, False, 1.5, A, B, 0.5, C0)
    assert np.array_equal(got, expect)
    # transposed storage gives the same product
    gotT = O.gemm_ref(True, True, 1.5, np.asfortranarray(A.T), np.asfortranarray(B.T), 0.5, C0)
    assert np.array_equal(gotT, expect)
    # row subset
    assert np.array_equal(O.gemm_ref(False, False, 1.5, A, B, 0.5, C0, rows=[2, 0]), expect[[2, 0]])


def test_gemm_ref_closed_forms():
    n = 37
    B = O.fill((n, 23), np.float64, 3)
    I = np.asfortranarray(np.eye(n))
    assert np.array_equal(O.gemm_ref(False, False, 1.0, I, B), B)
    ones_a = np.ones((19, 300), order="F")
    ones_b = np.ones((300, 5), order="F")
    assert np.all(O.gemm_ref(False, False, 1.0, ones_a, ones_b) == 300.0)


# ---------------------------------------------------------------- GEBP restatement
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("C", "T")])
def test_gebp_restatement_exact_integers(dtype, ta, tb):
    # small integer entries: every partial sum is exact, so any summation order is bit-equal
    m, n, k = 45, 31, 70
    A = O.fill((k, m) if ta != "N" else (m, k), dtype, 1, O.INT, -8, 8)
    B = O.fill((n, k) if tb != "N" else (k, n), dtype, 2, O.INT, -8, 8)
    C = O.fill((m, n), dtype, 3, O.INT, -8, 8)
    want = O.gemm_ref(ta != "N", tb != "N", 2.0, A, B, -1.0, C)
    got = C.copy(order="F")
    assert O.gemm_gebp(ta, tb, m, n, k, 2.0, A, A.shape[0], B, B.shape[0], -1.0, got, m, kc=16, mc=24, nc=12) == 0
    assert np.array_equal(got.astype(np.float64), want)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gebp_restatement_within_bound(dtype):
    m, n, k = 96, 80, 200
    A = O.fill((m, k), dtype, 4)
    B = O.fill((k, n), dtype, 5)
    ref, absref = O.gemm_ref(False, False, 1.0, A, B, with_abs=True)
    got = np.zeros((m, n), dtype, order="F")
    O.gemm_gebp("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, got, m, kc=48, mc=32, nc=16)
    err = np.linalg.norm(got - ref)
    u = np.finfo(dtype).eps / 2
    assert err <= O.higham_mary_bound(absref, k, u)
    assert err > 0  # the restatement really rounds in the working precision


@needs_ref
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_gebp_restatement_vs_eigen(dtype, ta, tb):
    m, n, k = 130, 70, 190
    A = O.fill((k, m) if ta else (m, k), dtype, 6)
    B = O.fill((n, k) if tb else (k, n), dtype, 7)
    ref, absref = O.gemm_ref(ta, tb, 0.75, A, B, with_abs=True)
    ours = np.zeros((m, n), dtype, order="F")
    O.gemm_gebp("T" if ta else "N", "T" if tb else "N", m, n, k, 0.75, A, A.shape[0], B, B.shape[0], 0.0,
                ours, m)
    eig = np.zeros((m, n), dtype, order="F")
    O.eigen_gemm(ta, tb, 0.75, A, B, eig, threads=2)
    u = np.finfo(dtype).eps / 2
    bound = O.higham_mary_bound(absref, k, u)
    assert np.linalg.norm(eig - ref) <= bound
    assert np.linalg.norm(ours - eig) <= bound
    # exact-integer inputs: Eigen and the restatement agree bit for bit
    Ai = O.fill(A.shape, dtype, 8, O.INT, -8, 8)
    Bi = O.fill(B.shape, dtype, 9, O.INT, -8, 8)
    e2 = np.zeros((m, n), dtype, order="F")
    o2 = np.zeros((m, n), dtype, order="F")
    O.eigen_gemm(ta, tb, 1.0, Ai, Bi, e2, threads=2)
    O.gemm_gebp("T" if ta else "N", "T" if tb else "N", m, n, k, 1.0, Ai, Ai.shape[0], Bi, Bi.shape[0], 0.0,
                o2, m)
    assert np.array_equal(e2, o2)


def test_gebp_blas_semantics():
    m, n, k = 9, 7, 5
    A = np.full((m, k), np.nan, order="F")
    B = np.full((k, n), np.inf, order="F")
    C = O.fill((m, n), np.float64, 1)
    C0 = C.copy(order="F")
    # alpha == 0: A, B never read, C = beta*C
    assert O.gemm_gebp("N", "N", m, n, k, 0.0, A, m, B, k, 2.0, C, m) == 0
    assert np.array_equal(C, 2.0 * C0)
    # beta == 0: NaN in C is overwritten, not propagated
    C[:] = np.nan
    Af = O.fill((m, k), np.float64, 2)
    Bf = O.fill((k, n), np.float64, 3)
    O.gemm_gebp("N", "N", m, n, k, 1.0, Af, m, Bf, k, 0.0, C, m)
    assert np.isfinite(C).all()
    # k == 0: C = beta*C
    C2 = C0.copy(order="F")
    O.gemm_gebp("N", "N", m, n, 0, 1.0, Af, m, Bf, 1, 0.5, C2, m)
    assert np.array_equal(C2, 0.5 * C0)


# ---------------------------------------------------------------- BLAS info codes
INVALID_ARGS = [
    ("X", "N", 2, 2, 2, 2, 2, 2), ("N", "Y", 2, 2, 2, 2, 2, 2), ("N", "N", -1, 2, 2, 2, 2, 2),
    ("N", "N", 2, -1, 2, 2, 2, 2), ("N", "N", 2, 2, -1, 2, 2, 2), ("N", "N", 4, 2, 2, 3, 2, 4),
    ("T", "N", 4, 2, 3, 2, 3, 4), ("N", "N", 2, 2, 5, 2, 4, 2), ("N", "T", 2, 6, 2, 2, 5, 2),
    ("N", "N", 5, 2, 2, 5, 2, 4), ("N", "N", 0, 0, 0, 0, 1, 1), ("n", "c", 3, 3, 3, 3, 3, 3),
    ("Q", "Z", -1, -1, -1, 0, 0, 0), ("N", "N", 2, 2, 2, 1, 1, 1),
]


@pytest.mark.parametrize("args", INVALID_ARGS)
def test_gemm_check_codes(args):
    expect = {("X",): 1}
    code = O.gemm_check(*args)
    ta, tb, m, n, k, lda, ldb, ldc = args
    # restated order of blas/level3_impl.h:61-78
    if ta.upper() not in "NTC":
        want = 1
    elif tb.upper() not in "NTC":
        want = 2
    elif m < 0:
        want = 3
    elif n < 0:
        want = 4
    elif k < 0:
        want = 5
    elif lda < max(1, m if ta.upper() == "N" else k):
        want = 8
    elif ldb < max(1, k if tb.upper() == "N" else n):
        want = 10
    elif ldc < max(1, m):
        want = 13
    else:
        want = 0
    assert code == want
    del expect


@needs_ref
@pytest.mark.parametrize("args", INVALID_ARGS)
def test_gemm_check_codes_vs_eigen_blas(args):
    import ctypes as C
    L = O.blas_ref()
    ta, tb, m, n, k, lda, ldb, ldc = args
    buf = (C.c_double * 64)()
    i = lambda v: C.byref(C.c_int(v))
    al, be = C.c_double(1.0), C.c_double(0.0)
    L.ref_blas_reset()
    if min(m, n, k) >= 0 and max(m, n, k) <= 8 and min(lda, ldb, ldc) >= 1:
        # only call with sizes that keep Eigen inside the 64-element buffer
        pass
    L.dgemm_(ta.encode(), tb.encode(), i(m), i(n), i(k), C.byref(al), buf, i(lda), buf, i(ldb), C.byref(be), buf,
             i(ldc))
    assert L.ref_blas_last_info() == O.gemm_check(*args)
    if L.ref_blas_last_info():
        assert L.ref_blas_last_name().decode() == "DGEMM "
