"""Batches of independent uneven products on the device.

`uneven_batch(count, seed)` builds the batched workload: `count` products
C = A*B with M, N, K each round(exp(U(ln lo, ln hi))), the first half FP32
and the second half FP64, entries U[-1, 1] from the device generator,
column-major with lda = M, ldb = K, ldc = M.  Operands of one dtype live in
one device buffer; the returned problems point into it.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import DIST_UNIFORM, F32, F64, GemmProblem, fill


@dataclass
class Batch:
    problems: list
    dims: np.ndarray            # (count, 3) int64: m, n, k
    dtypes: np.ndarray          # (count,) 0 = FP32, 1 = FP64
    offsets: np.ndarray         # (count, 3) element offsets of A, B, C in their dtype buffer
    buffers: dict = field(default_factory=dict)  # dtype -> torch tensor

    @property
    def flops(self) -> float:
        return float(np.sum(2.0 * self.dims[:, 0] * self.dims[:, 1] * self.dims[:, 2]))

    @property
    def bytes(self) -> float:
        """Algorithmic HBM bytes: read A and B, write C (beta = 0)."""
        m, n, k = self.dims[:, 0], self.dims[:, 1], self.dims[:, 2]
        size = np.where(self.dtypes == F32, 4, 8)
        return float(np.sum((m * k + k * n + m * n) * size))


def log_uniform_dims(count: int, seed: int, lo: int = 4, hi: int = 256) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=(count, 3)))).astype(np.int64)


def uneven_batch(count: int = 20000, seed: int = 0, lo: int = 4, hi: int = 256, alpha: float = 1.0,
                 beta: float = 0.0, dims: np.ndarray | None = None, dtypes: np.ndarray | None = None,
                 device: str = "cuda:0", data_seed: int = 7, dist: int = DIST_UNIFORM,
                 p0: float = -1.0, p1: float = 1.0) -> Batch:
    import torch

    if dims is None:
        dims = log_uniform_dims(count, seed, lo, hi)
    if dtypes is None:
        dtypes = np.where(np.arange(count) < count // 2, F32, F64)
    count = len(dims)
    offsets = np.zeros((count, 3), dtype=np.int64)
    totals = {F32: 0, F64: 0}
    for p in range(count):
        m, n, k = (int(x) for x in dims[p])
        d = int(dtypes[p])
        base = totals[d]
        offsets[p] = (base, base + m * k, base + m * k + k * n)
        totals[d] = base + m * k + k * n + m * n
    buffers = {}
    for d, tdt in ((F32, torch.float32), (F64, torch.float64)):
        n_el = max(totals[d], 1)
        t = torch.empty(n_el, dtype=tdt, device=device)
        fill(d, t, n_el, data_seed + d, dist, p0, p1)
        buffers[d] = t
    problems = []
    for p in range(count):
        m, n, k = (int(x) for x in dims[p])
        d = int(dtypes[p])
        t = buffers[d]
        esz = t.element_size()
        base = t.data_ptr()
        problems.append(GemmProblem(dtype=d, opa=0, opb=0, reserved=0, m=m, n=n, k=k, lda=max(m, 1),
                                    ldb=max(k, 1), ldc=max(m, 1), A=C.c_void_p(base + esz * int(offsets[p, 0])),
                                    B=C.c_void_p(base + esz * int(offsets[p, 1])),
                                    C=C.c_void_p(base + esz * int(offsets[p, 2])), alpha=alpha, beta=beta))
    return Batch(problems=problems, dims=dims, dtypes=np.asarray(dtypes), offsets=offsets, buffers=buffers)


def operand(batch: Batch, p: int, which: int) -> np.ndarray:
    """Host copy of operand `which` (0 = A, 1 = B, 2 = C) of product p, column-major."""
    m, n, k = (int(x) for x in batch.dims[p])
    shape = [(m, k), (k, n), (m, n)][which]
    t = batch.buffers[int(batch.dtypes[p])]
    off = int(batch.offsets[p, which])
    flat = t[off: off + shape[0] * shape[1]].cpu().numpy()
    return flat.reshape(shape, order="F")
