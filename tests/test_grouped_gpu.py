"""Grouped persistent kernel vs the oracle (runs on an MI355X).

Every product of a batch is checked on its own against the long-double
oracle: bit-exact for small-integer operands (all partial sums exact), within
the normwise bound of test/product_test_helpers.h for random operands.  The
same batch is run through the persistent kernel, the hipGraph replay (several
replays with eager work and synchronisation in between) and the
one-launch-per-product path.
"""
import ctypes as C

import numpy as np
import pytest

import eigen_amd as E
import oracle as O
from eigen_amd.batched import log_uniform_dims, operand, uneven_batch

pytestmark = pytest.mark.gpu


def _check_exact(batch, idx, alpha=1.0, beta=0.0, C0=None):
    for p in idx:
        A, B, Cg = (operand(batch, p, w) for w in range(3))
        want = O.gemm_ref(False, False, alpha, A, B, beta, None if C0 is None else C0[p])
        assert np.array_equal(Cg.astype(np.float64), want), (p, batch.dims[p], batch.dtypes[p])


def test_grouped_exact_mixed_batch(gpu):
    import torch
    dims = log_uniform_dims(300, seed=3, lo=1, hi=90)
    dtypes = np.where(np.arange(300) % 3 == 0, E.F64, E.F32)
    b = uneven_batch(dims=dims, dtypes=dtypes, dist=E.DIST_INT, p0=-8, p1=8)
    g = E.Grouped(b.problems)
    tiles, prods = g.info()
    assert prods == 300 and tiles == int(np.sum(((dims[:, 0] + 31) // 32) * ((dims[:, 1] + 31) // 32)))
    g.run()
    torch.cuda.synchronize()
    _check_exact(b, range(300))
    # hipGraph replays, with eager work in between: every replay must be right
    for rep in range(3):
        for p in range(0, 300, 7):  # clobber some outputs
            off = int(b.offsets[p, 2])
            m, n = int(dims[p, 0]), int(dims[p, 1])
            b.buffers[int(dtypes[p])][off: off + m * n].fill_(float("nan"))
        torch.cuda.synchronize()
        g.graph_launch()
        torch.cuda.synchronize()
        _check_exact(b, range(0, 300, 7))
    # the per-product reference path gives the same bits
    for p in range(0, 300, 5):
        off = int(b.offsets[p, 2])
        b.buffers[int(dtypes[p])][off: off + int(dims[p, 0] * dims[p, 1])].fill_(float("nan"))
    g.run_per_product()
    torch.cuda.synchronize()
    _check_exact(b, range(0, 300, 5))
    g.close()


def test_grouped_alpha_beta_and_transposes(gpu):
    import torch
    rng = np.random.default_rng(5)
    probs, host = [], []
    bufs = []
    for p in range(40):
        dt = E.F32 if p % 2 else E.F64
        npdt = np.float32 if dt == E.F32 else np.float64
        m, n, k = (int(x) for x in rng.integers(1, 70, 3))
        ta, tb = bool(p & 2), bool(p & 4)
        # every other problem stores A and B with padded leading dimensions
        pad = 3 if p % 3 == 0 else 0
        Af = O.fill((k + pad, m) if ta else (m + pad, k), npdt, 100 + p, O.INT, -8, 8)
        Bf = O.fill((n + pad, k) if tb else (k + pad, n), npdt, 200 + p, O.INT, -8, 8)
        A, B = np.asfortranarray(Af[:Af.shape[0] - pad]), np.asfortranarray(Bf[:Bf.shape[0] - pad])
        C0 = O.fill((m, n), npdt, 300 + p, O.INT, -8, 8)
        alpha, beta = [(1.0, 0.0), (2.0, -1.0), (0.0, 0.5), (-0.5, 1.0)][p % 4]
        dA, dB, dC = (torch.from_numpy(x.reshape(-1, order="F").copy()).cuda() for x in (Af, Bf, C0))
        bufs += [dA, dB, dC]
        probs.append(E.GemmProblem(dtype=dt, opa=int(ta), opb=int(tb), m=m, n=n, k=k, lda=Af.shape[0],
                                   ldb=Bf.shape[0], ldc=m, A=dA.data_ptr(), B=dB.data_ptr(), C=dC.data_ptr(),
                                   alpha=alpha, beta=beta))
        host.append((A, B, C0, ta, tb, alpha, beta, dC, (m, n)))
    g = E.Grouped(probs)
    g.graph_launch()
    torch.cuda.synchronize()
    for A, B, C0, ta, tb, alpha, beta, dC, (m, n) in host:
        want = O.gemm_ref(ta, tb, alpha, A, B, beta, C0)
        got = dC.cpu().numpy().reshape((n, m)).T
        assert np.array_equal(got.astype(np.float64), want)
    g.close()


def test_grouped_all_operand_types_many_runs_exact(gpu):
    """1,600 products over all eight operand types (dtype x op(A) x op(B)),
    interleaved in the input order and with K = 0 and alpha = 0 products among
    them, so that every type gets several workgroup runs; integer inputs make
    every result exact.  Checks every 5th product."""
    import torch
    rng = np.random.default_rng(21)
    probs, host, bufs = [], [], []
    for p in range(1600):
        dt = E.F32 if p % 2 else E.F64
        npdt = np.float32 if dt == E.F32 else np.float64
        m, n = (int(x) for x in rng.integers(1, 100, 2))
        k = 0 if p % 97 == 0 else int(rng.integers(1, 90))
        ta, tb = bool(p & 2), bool(p & 4)
        alpha, beta = (0.0, 2.0) if p % 89 == 0 else ((1.0, 0.0) if p % 3 else (-1.0, 0.5))
        A = O.fill((max(k, 1), m) if ta else (m, max(k, 1)), npdt, 1000 + p, O.INT, -8, 8)
        B = O.fill((n, max(k, 1)) if tb else (max(k, 1), n), npdt, 5000 + p, O.INT, -8, 8)
        C0 = O.fill((m, n), npdt, 9000 + p, O.INT, -8, 8)
        dA, dB, dC = (torch.from_numpy(x.reshape(-1, order="F").copy()).cuda() for x in (A, B, C0))
        bufs += [dA, dB, dC]
        probs.append(E.GemmProblem(dtype=dt, opa=int(ta), opb=int(tb), m=m, n=n, k=k, lda=A.shape[0],
                                   ldb=B.shape[0], ldc=m, A=dA.data_ptr(), B=dB.data_ptr(), C=dC.data_ptr(),
                                   alpha=alpha, beta=beta))
        if p % 5 == 0:
            Ak = A[:, :k] if not ta else A[:k, :]
            Bk = B[:k, :] if not tb else B[:, :k]
            host.append((np.asfortranarray(Ak), np.asfortranarray(Bk), C0, ta, tb, alpha, beta, dC, (m, n)))
    g = E.Grouped(probs)
    g.graph_launch()
    torch.cuda.synchronize()
    for A, B, C0, ta, tb, alpha, beta, dC, (m, n) in host:
        want = O.gemm_ref(ta, tb, alpha, A, B, beta, C0)
        got = dC.cpu().numpy().reshape((n, m)).T
        assert np.array_equal(got.astype(np.float64), want)
    g.close()


def test_grouped_schedules_bit_identical(gpu):
    """Every schedule (waves per workgroup x band width) computes each tile with
    the same code, so all give the same bits; checked on a mixed uneven batch
    (FP32 and FP64, random inputs), with a captured graph dropped and rebuilt
    on every change."""
    import torch
    b = uneven_batch(count=600, seed=13)
    g = E.Grouped(b.problems)
    ns = E.lib().eh_grouped_num_schedules()
    assert ns == 3
    results = []
    for sched in range(ns):
        g.set_schedule(sched)
        assert g.schedule() == sched
        g.graph_launch()
        torch.cuda.synchronize()
        results.append([t.clone() for t in b.buffers.values()])
    for r in results[1:]:
        for x, y in zip(results[0], r):
            assert torch.equal(x, y)
    with pytest.raises(E.EigenAmdError):
        g.set_schedule(ns)
    g.close()


def test_grouped_table_mode_tunes_schedule_without_touching_c(gpu):
    """Under EH_TUNE_TABLE, eh_grouped_create times every schedule on the batch
    (scratch results): C holds its old values until the first run, and the
    run under the chosen schedule matches schedule 0 bit for bit."""
    import torch
    b = uneven_batch(count=500, seed=17, beta=0.5)
    before = [t.clone() for t in b.buffers.values()]
    prev = E.lib().eh_get_tune_mode()
    E.set_tune_mode(E.TUNE_TABLE)
    try:
        g = E.Grouped(b.problems)
    finally:
        E.set_tune_mode(prev)
    torch.cuda.synchronize()
    for x, y in zip(before, b.buffers.values()):
        assert torch.equal(x, y)  # tuning never wrote C (nor A, B)
    chosen = g.schedule()
    assert 0 <= chosen < E.lib().eh_grouped_num_schedules()
    g.run()
    torch.cuda.synchronize()
    tuned = [t.clone() for t in b.buffers.values()]
    for t, x in zip(b.buffers.values(), before):  # restore C and run the default schedule
        t.copy_(x)
    g.set_schedule(0)
    g.run()
    torch.cuda.synchronize()
    for x, y in zip(tuned, b.buffers.values()):
        assert torch.equal(x, y)
    g.close()


def test_grouped_random_within_bound(gpu):
    import torch
    b = uneven_batch(count=400, seed=11)  # the batched workload's distribution, smaller count
    g = E.Grouped(b.problems)
    g.graph_launch()
    torch.cuda.synchronize()
    for p in range(0, 400, 9):
        A, B, Cg = (operand(b, p, w) for w in range(3))
        k = A.shape[1]
        ref, absref = O.gemm_ref(False, False, 1.0, A, B, with_abs=True)
        u = np.finfo(A.dtype).eps / 2
        assert np.linalg.norm(Cg.astype(np.float64) - ref) <= O.higham_mary_bound(absref, k, u) + 1e-300
    g.close()


def _dev(x):
    import torch
    return torch.from_numpy(np.asarray(x).reshape(-1, order="F").copy()).cuda()


def test_grouped_single_small_product(gpu):
    """One product of one tile: a grid of one workgroup, most waves idle."""
    import torch
    A = O.fill((5, 3), np.float64, 1, O.INT, -8, 8)
    B = O.fill((3, 7), np.float64, 2, O.INT, -8, 8)
    dA, dB, dC = _dev(A), _dev(B), torch.zeros(35, dtype=torch.float64, device="cuda")
    g = E.Grouped([E.GemmProblem(dtype=E.F64, opa=0, opb=0, m=5, n=7, k=3, lda=5, ldb=3, ldc=5, A=dA.data_ptr(),
                                 B=dB.data_ptr(), C=dC.data_ptr(), alpha=2.0, beta=0.0)])
    g.graph_launch()
    torch.cuda.synchronize()
    got = dC.cpu().numpy().reshape((7, 5)).T
    assert np.array_equal(got, O.gemm_ref(False, False, 2.0, A, B))
    g.close()


def test_grouped_operand_past_31_bit_offsets(gpu):
    """A product whose A spans more than 2^31 bytes (op(A) = A^T with its three
    stored columns 2^28 floats apart) runs through the single-product path
    after the kernel, in the same launch and the same captured graph, next to
    ordinary products."""
    import torch
    m, n, k, lda = 3, 5, 2, (1 << 28) + 16
    bigA = torch.zeros((m - 1) * lda + k, dtype=torch.float32, device="cuda")
    At = O.fill((k, m), np.float32, 3, O.INT, -8, 8)  # stored k x m, op(A) = At^T
    for i in range(m):
        bigA[i * lda:i * lda + k] = torch.from_numpy(At[:, i].copy()).cuda()
    B = O.fill((k, n), np.float32, 4, O.INT, -8, 8)
    dB, dC = _dev(B), torch.zeros(m * n, dtype=torch.float32, device="cuda")
    small = []
    host = []
    for p in range(6):
        As = O.fill((9, 20), np.float32, 10 + p, O.INT, -8, 8)
        Bs = O.fill((20, 11), np.float32, 20 + p, O.INT, -8, 8)
        dAs, dBs, dCs = _dev(As), _dev(Bs), torch.zeros(99, dtype=torch.float32, device="cuda")
        host.append((As, Bs, dAs, dBs, dCs))
        small.append(E.GemmProblem(dtype=E.F32, opa=0, opb=0, m=9, n=11, k=20, lda=9, ldb=20, ldc=9,
                                   A=dAs.data_ptr(), B=dBs.data_ptr(), C=dCs.data_ptr(), alpha=1.0, beta=0.0))
    big = E.GemmProblem(dtype=E.F32, opa=1, opb=0, m=m, n=n, k=k, lda=lda, ldb=k, ldc=m, A=bigA.data_ptr(),
                        B=dB.data_ptr(), C=dC.data_ptr(), alpha=1.0, beta=0.0)
    g = E.Grouped(small[:3] + [big] + small[3:])
    assert g.info() == (6, 7)  # six one-tile products in the kernel, seven products in all
    for _ in range(2):
        dC.zero_()
        g.graph_launch()
        torch.cuda.synchronize()
        got = dC.cpu().numpy().reshape((n, m)).T
        assert np.array_equal(got.astype(np.float64), O.gemm_ref(True, False, 1.0, At, B))
        for As, Bs, _, _, dCs in host:
            assert np.array_equal(dCs.cpu().numpy().reshape((11, 9)).T.astype(np.float64),
                                  O.gemm_ref(False, False, 1.0, As, Bs))
    g.close()
    del bigA
    torch.cuda.empty_cache()
