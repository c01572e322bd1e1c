"""The benchmark's own workloads at their stated sizes (runs on an MI355X).

* sgemm 8192^3, column-major, entries sign*U(1,2)*2^U(-8,8) from the
  benchmark's seeds (A: 1, B: 2): plain FP32 MFMA, 2-term and 3-term
  split-BF16, each checked on sampled rows against the long-double oracle.
  The 3-term split must be no less accurate than the plain product, the
  2-term split less accurate than the 3-term one.
* The 20,000 uneven products (M, N, K log-uniform in [4, 256], half FP32,
  half FP64) of the benchmark's batch, through the persistent grouped kernel
  captured in a hipGraph: random entries within the normwise bound of
  test/product_test_helpers.h:56-74 on every 50th product, and small-integer
  entries of the same shapes bit-equal to the oracle and to one launch per
  product, over every product of the batch and over repeated replays.
"""
import numpy as np
import pytest

import eigen_amd as E
import oracle as O
from eigen_amd.batched import operand, uneven_batch
from helpers import unit_roundoff

pytestmark = pytest.mark.gpu


def test_sgemm_8192_wide_range_plain_split2_split3(gpu):
    import torch
    m = n = k = 8192
    dA = torch.empty(m * k, dtype=torch.float32, device="cuda")
    dB = torch.empty(k * n, dtype=torch.float32, device="cuda")
    dC = torch.empty(m * n, dtype=torch.float32, device="cuda")
    E.fill(E.F32, dA, m * k, 1, E.DIST_WIDE, -8, 8)
    E.fill(E.F32, dB, k * n, 2, E.DIST_WIDE, -8, 8)
    rows = np.arange(0, m, m // 32, dtype=np.int64)
    idx = torch.from_numpy(rows).cuda()
    A_rows = np.asfortranarray(dA.view(k, m)[:, idx].T.cpu().numpy())  # (32, k)
    B = dB.view(n, k).T.cpu().numpy()                                    # (k, n), column-major
    assert B.flags.f_contiguous
    nA = float(torch.linalg.norm(dA.double()))
    nB = float(torch.linalg.norm(dB.double()))
    ref, absref = O.gemm_ref(False, False, 1.0, A_rows, B, with_abs=True)
    u = unit_roundoff(np.float32)
    errs = {}
    for name, prec in (("plain", E.PREC_PLAIN), ("split2", E.PREC_SPLIT2), ("split3", E.PREC_SPLIT3)):
        dC.fill_(float("nan"))
        E.gemm(E.F32, E.OP_N, E.OP_N, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, prec=prec)
        got = dC.view(n, m)[:, idx].T.cpu().numpy().astype(np.float64)
        diff = np.linalg.norm(got - ref)
        assert diff <= O.higham_mary_bound(absref, k, u), name
        # the normwise metric ||C - C_ref||_F / (K u ||A||_F ||B||_F), the sampled
        # rows' difference scaled to the whole matrix
        errs[name] = diff * np.sqrt(m / len(rows)) / (k * u * nA * nB)
        assert errs[name] <= 1.0, (name, errs[name])
    print("normwise errors at 8192^3:", errs, flush=True)
    assert errs["split3"] <= errs["plain"], errs
    assert errs["split2"] > errs["split3"], errs
    del dA, dB, dC
    torch.cuda.empty_cache()


def _clobber_outputs(batch, every=1):
    for p in range(0, len(batch.problems), every):
        off = int(batch.offsets[p, 2])
        size = int(batch.dims[p, 0] * batch.dims[p, 1])
        batch.buffers[int(batch.dtypes[p])][off: off + size].fill_(float("nan"))


def test_batched_20000_random_within_bound(gpu):
    import torch
    b = uneven_batch(20000, seed=0)  # the benchmark's batch
    assert len(b.problems) == 20000
    g = E.Grouped(b.problems)
    for _ in range(3):  # every replay recomputes every product
        _clobber_outputs(b, every=3)
        torch.cuda.synchronize()
        g.graph_launch()
        torch.cuda.synchronize()
    for p in range(0, 20000, 50):
        A, B, Cg = (operand(b, p, w) for w in range(3))
        ref, absref = O.gemm_ref(False, False, 1.0, A, B, with_abs=True)
        u = unit_roundoff(A.dtype)
        assert np.linalg.norm(Cg.astype(np.float64) - ref) <= O.higham_mary_bound(absref, A.shape[1], u) + 1e-300, \
            (p, b.dims[p], b.dtypes[p])
    g.close()


def test_batched_20000_exact_every_product(gpu):
    import torch
    b = uneven_batch(20000, seed=0, dist=E.DIST_INT, p0=-8, p1=8)  # same shapes, exact arithmetic
    g = E.Grouped(b.problems)
    tiles, prods = g.info()
    assert prods == 20000
    assert tiles == int(np.sum(((b.dims[:, 0] + 31) // 32) * ((b.dims[:, 1] + 31) // 32)))
    snaps = []
    for rep in range(2):
        _clobber_outputs(b)
        torch.cuda.synchronize()
        g.graph_launch()
        torch.cuda.synchronize()
        snaps.append({d: t.clone() for d, t in b.buffers.items()})
    for d in b.buffers:
        assert torch.equal(snaps[0][d], snaps[1][d]), d  # NaN left anywhere fails too
    for p in range(0, 20000, 20):
        A, B, Cg = (operand(b, p, w) for w in range(3))
        assert np.array_equal(Cg.astype(np.float64), O.gemm_ref(False, False, 1.0, A, B)), (p, b.dims[p])
    # one launch per product gives the same bits for every product
    _clobber_outputs(b)
    g.run_per_product()
    torch.cuda.synchronize()
    for d, t in b.buffers.items():
        assert torch.equal(t, snaps[0][d]), d
    g.close()
