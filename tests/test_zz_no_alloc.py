"""eh_gemm_noalloc against eh_gemm (runs on an MI355X).

The allocation-free entry point runs eh_gemm's kernels with its workspace in
the library's fixed device scratch array, so its results must equal
eh_gemm's bit for bit.  Each case makes two calls in a row, the second on a
smaller product (the scratch keeps the first call's contents), covering
split-K partial sums (plain FP32 and FP64), the packed planes of the
split-BF16 product, and matrix-vector partial sums; a product that needs
more than the scratch holds is refused before anything runs.
"""
import ctypes as C

import numpy as np
import pytest

import eigen_amd as E
import oracle as O

pytestmark = pytest.mark.gpu


def _call(fn, dtype, prec, m, n, k, alpha, dA, dB, beta, dC):
    ct = C.c_float if dtype == E.F32 else C.c_double
    al, be = ct(alpha), ct(beta)
    return fn(dtype, 0, 0, m, n, k, C.byref(al), C.c_void_p(dA.data_ptr()), m, C.c_void_p(dB.data_ptr()), k,
              C.byref(be), C.c_void_p(dC.data_ptr()), m, prec, None)


def _case(dtype, prec, m, n, k, seed):
    import torch
    npdt = np.float32 if dtype == E.F32 else np.float64
    A = O.fill((m, k), npdt, seed, O.WIDE if dtype == E.F32 else O.UNIFORM, -8, 8)
    B = O.fill((k, n), npdt, seed + 1, O.WIDE if dtype == E.F32 else O.UNIFORM, -8, 8)
    C0 = O.fill((m, n), npdt, seed + 2)
    dev = lambda x: torch.from_numpy(x.reshape(-1, order="F").copy()).cuda()
    return dev(A), dev(B), dev(C0)


@pytest.mark.parametrize("dtype,prec,shapes", [
    (E.F64, E.PREC_PLAIN, [(96, 80, 16384), (40, 24, 6000)]),         # split K
    (E.F32, E.PREC_PLAIN, [(128, 64, 20000), (33, 17, 5000)]),        # split K
    (E.F32, E.PREC_SPLIT3, [(512, 384, 700), (200, 130, 257)]),       # packed planes
    (E.F32, E.PREC_SPLIT2, [(300, 280, 1100), (64, 64, 64)]),         # packed planes
    (E.F64, E.PREC_PLAIN, [(64, 1, 200000), (17, 1, 90000)]),         # matrix-vector, split K
])
def test_noalloc_matches_gemm_twice(gpu, dtype, prec, shapes):
    import torch
    L = E.lib()
    for i, (m, n, k) in enumerate(shapes):
        dA, dB, dC = _case(dtype, prec, m, n, k, 10 * i + 1)
        ref = dC.clone()
        assert _call(L.eh_gemm, dtype, prec, m, n, k, 1.5, dA, dB, 0.5, ref) == 0
        got = dC.clone()
        assert _call(L.eh_gemm_noalloc, dtype, prec, m, n, k, 1.5, dA, dB, 0.5, got) == 0, E.lib().eh_last_error()
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (m, n, k)


def test_noalloc_refuses_oversized_product(gpu):
    import torch
    L = E.lib()
    m = n = k = 4096  # split-BF16 planes: 2 x 6 bytes x 4096^2 = 201 MB > the scratch
    dA = torch.zeros(m * k, dtype=torch.float32, device="cuda")
    dB = torch.zeros(k * n, dtype=torch.float32, device="cuda")
    dC = torch.full((m * n,), 7.0, dtype=torch.float32, device="cuda")
    st = _call(L.eh_gemm_noalloc, E.F32, E.PREC_SPLIT3, m, n, k, 1.0, dA, dB, 0.0, dC)
    assert st == 3  # EH_ERR_UNSUPPORTED
    assert b"scratch" in L.eh_last_error()
    torch.cuda.synchronize()
    assert bool((dC == 7.0).all())  # nothing ran
    # the same call through eh_gemm (which may allocate) works
    assert _call(L.eh_gemm, E.F32, E.PREC_SPLIT3, m, n, k, 1.0, dA, dB, 0.0, dC) == 0
    torch.cuda.synchronize()
    assert bool((dC == 0.0).all())
