#!/usr/bin/env python3
"""Short driver for profiler runs: a few launches of one workload.

    python profiles/run_kernels.py {split3,plain,dgemm,batched,copy} [reps]

Used as the program under rocprofv3 for the summaries in this directory
(the benchmark itself runs every workload and the CPU baseline)."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import torch  # noqa: E402

import eigen_amd as E  # noqa: E402


def main() -> int:
    what = sys.argv[1]
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 5
    E.lib()
    if what in ("split3", "plain"):
        n = 8192
        A = torch.empty(n * n, dtype=torch.float32, device="cuda")
        B = torch.empty(n * n, dtype=torch.float32, device="cuda")
        C = torch.empty(n * n, dtype=torch.float32, device="cuda")
        E.fill(E.F32, A, n * n, 1, E.DIST_WIDE, -8, 8)
        E.fill(E.F32, B, n * n, 2, E.DIST_WIDE, -8, 8)
        prec = E.PREC_SPLIT3 if what == "split3" else E.PREC_PLAIN
        for _ in range(reps):
            E.gemm(E.F32, 0, 0, n, n, n, 1.0, A, n, B, n, 0.0, C, n, prec=prec)
    elif what == "dgemm":
        n = 4096
        A, B, C = (torch.empty(n * n, dtype=torch.float64, device="cuda") for _ in range(3))
        for i, t in enumerate((A, B, C)):
            E.fill(E.F64, t, n * n, 11 + i, E.DIST_NORMAL, 0, 1)
        for _ in range(reps):
            E.gemm(E.F64, 0, 0, n, n, n, 1.5, A, n, B, n, 0.5, C, n)
    elif what == "batched":
        from eigen_amd.batched import uneven_batch
        b = uneven_batch(20000, seed=0)
        g = E.Grouped(b.problems)
        for _ in range(reps):
            g.graph_launch()
    elif what == "copy":
        # calibration for the HBM byte counters: a plain 1 GiB device copy
        x = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(reps):
            y.copy_(x)
    else:
        raise SystemExit(f"unknown workload {what}")
    torch.cuda.synchronize()
    print("done", what, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
