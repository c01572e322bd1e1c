"""Concurrent callers and the per-stream workspace (runs on an MI355X).

Split-K products keep their partial sums in the library's per-stream
workspace, which grows when a larger product needs more.  Several host
threads call dgemm_ (the BLAS export, null stream) and eh_gemm on
hipStreamPerThread with growing split-K shapes; every result is checked
exactly (small-integer operands, all partial sums exact).  A buffer freed
while another thread's launch still used it, or one workspace shared by two
per-thread streams, would show as a wrong result.
"""
import ctypes as C
import threading

import numpy as np
import pytest

import eigen_amd as E
import oracle as O

pytestmark = pytest.mark.gpu

HIP_STREAM_PER_THREAD = 2  # the hipStreamPerThread handle value

# small results with deep k: the planner splits K, so every call uses workspace
SHAPES = [(32, 24, 4096), (48, 40, 8192), (64, 64, 12288), (96, 80, 16384), (128, 96, 20480)]


def _inputs(seed):
    import torch
    out = []
    for i, (m, n, k) in enumerate(SHAPES):
        A = O.fill((m, k), np.float64, seed + 10 * i, O.INT, -4, 4)
        B = O.fill((k, n), np.float64, seed + 10 * i + 1, O.INT, -4, 4)
        want = O.gemm_ref(False, False, 1.0, A, B)
        dA = torch.from_numpy(A.reshape(-1, order="F").copy()).cuda()
        dB = torch.from_numpy(B.reshape(-1, order="F").copy()).cuda()
        dC = torch.empty(m * n, dtype=torch.float64, device="cuda")
        out.append((m, n, k, dA, dB, dC, want))
    return out


def _run_threads(work, nthreads=6, rounds=4):
    errors = []

    def body(t):
        try:
            work(t, rounds)
        except Exception as e:  # reported to the main thread
            errors.append(repr(e))
    ts = [threading.Thread(target=body, args=(t,)) for t in range(nthreads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors


def test_threads_dgemm_growing_shapes_null_stream(gpu):
    import torch
    L = E.blas()
    data = [_inputs(100 + 7 * t) for t in range(6)]
    torch.cuda.synchronize()
    i = lambda v: C.byref(C.c_int(v))

    def work(t, rounds):
        for r in range(rounds):
            for m, n, k, dA, dB, dC, want in data[t]:
                al, be = C.c_double(1.0), C.c_double(0.0)
                L.dgemm_(b"N", b"N", i(m), i(n), i(k), C.byref(al), C.c_void_p(dA.data_ptr()), i(m),
                         C.c_void_p(dB.data_ptr()), i(k), C.byref(be), C.c_void_p(dC.data_ptr()), i(m))
                got = dC.cpu().numpy().reshape((n, m)).T  # dgemm_ returns synchronised
                if not np.array_equal(got, want):
                    raise AssertionError(f"thread {t} round {r} shape {(m, n, k)} mismatch")
    _run_threads(work)


def test_threads_eh_gemm_per_thread_streams(gpu):
    import torch
    L = E.lib()
    data = [_inputs(300 + 7 * t) for t in range(6)]
    torch.cuda.synchronize()
    one, zero = C.c_double(1.0), C.c_double(0.0)
    sps = C.c_void_p(HIP_STREAM_PER_THREAD)

    def work(t, rounds):
        for r in range(rounds):
            for m, n, k, dA, dB, dC, want in data[t]:
                st = L.eh_gemm(E.F64, 0, 0, m, n, k, C.byref(one), C.c_void_p(dA.data_ptr()), m,
                               C.c_void_p(dB.data_ptr()), k, C.byref(zero), C.c_void_p(dC.data_ptr()), m,
                               E.PREC_PLAIN, sps)
                E.check(st, "eh_gemm")
                E.check(L.eh_synchronize(sps), "eh_synchronize")
                got = dC.cpu().numpy().reshape((n, m)).T
                if not np.array_equal(got, want):
                    raise AssertionError(f"thread {t} round {r} shape {(m, n, k)} mismatch")
    _run_threads(work)
