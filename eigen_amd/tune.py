"""Exhaustive autotuning sweep: time every candidate configuration on many shapes.

    python -m eigen_amd.tune --shapes 4500 --out tune.csv [--resume] [--max-seconds S]

Shapes: M, N, K = round(exp(U(ln 64, ln 16384))), dtype FP32 or FP64 and the
two transpose flags uniformly at random (seeded), followed by a fixed list of
canonical shapes (canonical_shapes).  For every shape each of the
24 configurations is timed on the device with HIP events (median of `reps`
after a warm-up; reps shrink for large products).  Output is a CSV with one
row per shape:  index,dtype,ta,tb,m,n,k,t0,...,t23  (milliseconds).
Rows are appended and flushed as they complete; --resume skips rows present.
"""
from __future__ import annotations

import argparse
import csv
import ctypes as C
import os
import sys
import time

import numpy as np


def sample_shapes(count: int, seed: int = 2024, lo: int = 64, hi: int = 16384) -> np.ndarray:
    """(count, 6) int64 rows: dtype, ta, tb, m, n, k."""
    rng = np.random.default_rng(seed)
    dims = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=(count, 3)))).astype(np.int64)
    flags = rng.integers(0, 2, size=(count, 3)).astype(np.int64)
    return np.concatenate([flags, dims], axis=1)


def canonical_shapes() -> np.ndarray:
    """Common shapes the random sample rarely hits: squares 256..16384 in both
    dtypes and all four transpose combinations, and tall / skinny / deep
    non-transposed products."""
    rows = []
    for d in (0, 1):
        for ta in (0, 1):
            for tb in (0, 1):
                for e in range(8, 15):
                    rows.append((d, ta, tb, 1 << e, 1 << e, 1 << e))
        for m, n, k in ((16384, 256, 16384), (256, 16384, 16384), (16384, 16384, 256), (8192, 8192, 1024),
                        (1024, 1024, 16384)):
            rows.append((d, 0, 0, m, n, k))
    return np.array(rows, dtype=np.int64)


def all_shapes(count: int, seed: int = 2024) -> np.ndarray:
    """The random sample (indices 0 .. count-1) followed by the canonical shapes."""
    return np.concatenate([sample_shapes(count, seed), canonical_shapes()], axis=0)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--shapes", type=int, default=4500)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--out", required=True)
    ap.add_argument("--resume", action="store_true")
    ap.add_argument("--max-seconds", type=float, default=1e9)
    args = ap.parse_args(argv)

    import torch
    from . import F32, check, fill, lib, num_configs

    shapes = all_shapes(args.shapes, args.seed)
    done = set()
    if args.resume and os.path.exists(args.out):
        with open(args.out) as f:
            for row in csv.reader(f):
                if row and row[0].isdigit():
                    done.add(int(row[0]))
    L = lib()
    ncfg = num_configs()
    # one buffer set per dtype, sized for the largest operand of the sweep
    bufs = {}
    for d, tdt in ((0, torch.float32), (1, torch.float64)):
        sel = shapes[shapes[:, 0] == d]
        if len(sel) == 0:
            continue
        m, n, k = sel[:, 3], sel[:, 4], sel[:, 5]
        sizes = [int((m * k).max()), int((k * n).max()), int((m * n).max())]
        bs = [torch.empty(s, dtype=tdt, device="cuda:0") for s in sizes]
        for i, b in enumerate(bs):
            fill(d, b, b.numel(), 100 + i)
        bufs[d] = bs
    torch.cuda.synchronize()
    new = not os.path.exists(args.out) or not args.resume
    f = open(args.out, "w" if new else "a", newline="")
    w = csv.writer(f)
    if new:
        w.writerow(["index", "dtype", "ta", "tb", "m", "n", "k"] + [f"t{c}" for c in range(ncfg)])
    t_start = time.time()
    ms = C.c_float()
    for i, (d, ta, tb, m, n, k) in enumerate(shapes.tolist()):
        if i in done:
            continue
        if time.time() - t_start > args.max_seconds:
            print(f"stopping at shape {i}: time budget reached", flush=True)
            break
        A, B, Cb = bufs[d]
        flop = 2.0 * m * n * k
        reps = 1 if flop > 2e11 else (2 if flop > 2e10 else 3)
        row = []
        for c in range(ncfg):
            check(L.eh_time_config(d, ta, tb, m, n, k, A.data_ptr(), B.data_ptr(), Cb.data_ptr(), c, reps,
                                   C.byref(ms)), "eh_time_config")
            row.append(f"{ms.value:.5f}")
        w.writerow([i, d, ta, tb, m, n, k] + row)
        if i % 25 == 0:
            f.flush()
            print(f"shape {i}/{len(shapes)} ({d},{ta},{tb},{m},{n},{k}) best {min(map(float, row)):.3f} ms "
                  f"elapsed {time.time() - t_start:.0f}s", flush=True)
    f.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
