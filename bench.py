#!/usr/bin/env python3
"""Benchmark of the MI355X dense-product path (one JSON line on stdout).

Headline workload: FP32 sgemm C = A*B, M = N = K = 8192, column-major,
entries sign*U(1,2)*2^U(-8,8) from a seed, computed by the split-BF16 path
(3 bf16 terms per operand, 6 cross products on BF16 MFMA, FP32
accumulation).  One step = one such product with inputs resident in HBM.
`value` is its TFLOP/s counting 2*M*N*K once.

Beside it the same line reports the plain FP32-MFMA product on the same
inputs (TFLOP/s and normwise error against an FP64 host product on sampled
rows), the 2-term split, the FP64 dgemm drop-in (4096^3 and 16384x256x16384,
C = 1.5*A*B + 0.5*C), the batched uneven products, and Eigen's CPU GEBP
timed on this host's cores (`cpu_baseline`).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"bf16": 2500.0, "fp16": 2500.0, "fp32": 157.3, "fp64": 78.6}  # dense MFMA TFLOP/s, MI355X
HBM_PEAK_GBS = 8000.0
WARM_MS = 100.0  # untimed warm-up of each leg after the headline (at least, beyond --warmup steps)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed(fn, steps: int, warmup: int, kernels=(), warm_ms: float = 0.0):
    """Wall-clock ms per step and HIP-event ms per step on the launch stream.

    `kernels`: kernel classes (eigen_amd.KC_*) whose launches inside the timed
    steps are bracketed by HIP events on their own stream; their mean launch
    time in ms is returned as a dict (warm-up launches are not recorded).
    `warm_ms`: after the `warmup` steps, keep running untimed steps until this
    much time has passed (the legs after the headline start from a GPU left
    idle by the previous leg's host-side error check, and a short leg would
    otherwise time the clock ramp)."""
    import torch
    import eigen_amd as E
    t_warm = time.perf_counter()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    while (time.perf_counter() - t_warm) * 1e3 < warm_ms:
        fn()
        torch.cuda.synchronize()
    if kernels:
        E.kernel_timing(True)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(steps):
        fn()
    ev1.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    kt = {}
    for c in kernels:
        ms, n = E.kernel_timing_read(c)
        kt[c] = (ms / n if n else 0.0, n)
    if kernels:
        E.kernel_timing(False)
    if not kernels:
        return wall, ev0.elapsed_time(ev1) / steps
    return wall, ev0.elapsed_time(ev1) / steps, kt


def exact_product(A, B):
    """A @ B for float64 (or narrower) operands, accurate to ~2^-60 relative.

    Each operand is cut into four slices of 16 significant bits relative to
    the largest magnitude of its row (A) or column (B), so every slice product
    (at most 34 bits) and every partial sum of a k-term dot product of slices
    (k <= 2^17) is an exact float64; each of the 16 slice products is
    therefore computed exactly by float64 BLAS in any summation order, and the
    exact results are summed in extended precision (Ozaki, Ogita, Oishi &
    Rump, "Error-free transformations of matrix multiplication", Numer.
    Algorithms 59, 2012).  What the four slices leave of an operand is below
    2^-64 of its row's (column's) largest entry."""
    A = np.asarray(A, np.float64)
    B = np.asarray(B, np.float64)
    k = A.shape[1]
    assert k <= 1 << 17

    def slices(X, axis):
        mx = np.max(np.abs(X), axis=axis, keepdims=True)
        e = np.where(mx > 0, np.ceil(np.log2(np.where(mx > 0, mx, 1.0))), 0.0)
        out, R = [], X.copy()
        for s_ in range(4):
            # adding 1.5 * 2^q rounds R to a multiple of 2^(q - 52) = 2^(e - 16(s+1))
            sigma = np.ldexp(1.5, (e + 52 - 16 * (s_ + 1)).astype(np.int64))
            hi = (R + sigma) - sigma
            out.append(hi)
            R = R - hi
        return out
    sa, sb = slices(A, 1), slices(B, 0)
    acc = np.zeros((A.shape[0], B.shape[1]), np.longdouble)
    for a in sa:
        for b in sb:
            acc += (a @ b).astype(np.longdouble)
    return acc


def normwise(got, ref, k, u, normA, normB, scale=1.0):
    """||C - C_ref||_F / (K u ||A||_F ||B||_F) -- the sampled part of C scaled
    by `scale` (sqrt of the sampling ratio) to the whole matrix."""
    den = k * u * normA * normB
    d = float(np.linalg.norm(np.asarray(got, np.longdouble) - np.asarray(ref, np.longdouble))) * scale
    return d / den if den > 0 else d


def sampled_rows_error(dA, dB, dC, m, n, k, rows, dtype, alpha=1.0, beta=0.0, C0_rows=None):
    """Normwise error of C (m x n, column-major device buffers) on sampled
    rows against the exact product of the same inputs (FP64 host arithmetic,
    error-free for FP64 inputs), rows scaled to the whole matrix."""
    import torch
    idx = torch.from_numpy(np.asarray(rows, np.int64)).to(dA.device)
    A_rows = dA.view(k, m)[:, idx].T.double().cpu().numpy()
    B = dB.view(n, k).T.double().cpu().numpy()
    got = dC.view(n, m)[:, idx].T.double().cpu().numpy()
    ref = alpha * exact_product(A_rows, B)
    if C0_rows is not None and beta != 0.0:
        ref = ref + beta * np.asarray(C0_rows, np.longdouble)
    u = float(np.finfo(dtype).eps) / 2
    nA = float(torch.linalg.norm(dA.double()))
    nB = float(torch.linalg.norm(dB.double()))
    return normwise(got, ref, k, u, abs(alpha) * nA, nB, np.sqrt(m / len(rows)))


def measured_traffic(workload: str):
    """HBM bytes per launch of the workload's dominant kernel, as measured by the
    FETCH_SIZE / WRITE_SIZE counter runs summarised in profiles/traffic.json
    (each counter scaled by its reading for a copy kernel of known size; written
    by `python -m eigen_amd.profile_kernels summarize`)."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            entry = json.load(f).get(workload)
        return int(entry["bytes"]) if entry else None
    except (OSError, ValueError, KeyError, TypeError):
        return None


def host_cpu() -> str:
    """CPU model, logical CPUs visible and OMP_NUM_THREADS of this host."""
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return (f"{model}, {os.cpu_count()} logical CPUs visible, "
            f"OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS', 'unset')}")


def cpu_baseline(reps: int = 7):
    """Eigen's CPU GEBP (built from the reference tree into oracle/_ref) on this host."""
    try:
        import oracle as O
        if not O.ref_available():
            return None
        L = O.ref()
        threads = L.ref_get_threads()
        m = n = k = 8192
        A = O.fill((m, k), np.float32, 1, O.WIDE, -8, 8)
        B = O.fill((k, n), np.float32, 2, O.WIDE, -8, 8)
        Cm = np.zeros((m, n), np.float32, order="F")
        t0 = time.perf_counter()
        O.eigen_gemm(False, False, 1.0, A, B, Cm)  # warm-up
        first = time.perf_counter() - t0
        reps = max(5, min(reps, int(20.0 // max(first, 1e-3))))  # median of >= 5, ~20 s at most
        ts = []
        for _ in range(reps):
            Cm[:] = 0
            t0 = time.perf_counter()
            O.eigen_gemm(False, False, 1.0, A, B, Cm)
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        return {"value": round(2.0 * m * n * k / t / 1e12, 4), "unit": "TFLOP/s", "cores": int(threads),
                "kind": "reference",
                "sample": f"Eigen C.noalias() += A*B, FP32 8192^3 on host Maps, same input distribution, "
                          f"OpenMP {threads} threads, {L.ref_simd().decode()}, median of {reps} after 1 warm-up; "
                          f"{host_cpu()}"}
    except Exception as e:  # the baseline is reported, never required
        log("cpu baseline failed:", e)
        return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-extra", action="store_true", help="headline only")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the product path here has no multi-GPU mode "
              "(one process per GPU would each run this path on its block of the product); run with --gpus 1",
              file=sys.stderr)
        return 2

    import torch
    import eigen_amd as E

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    E.lib()
    m = n = k = 8192
    flop = 2.0 * m * n * k
    dA = torch.empty(m * k, dtype=torch.float32, device=dev)
    dB = torch.empty(k * n, dtype=torch.float32, device=dev)
    dC = torch.zeros(m * n, dtype=torch.float32, device=dev)
    E.fill(E.F32, dA, m * k, 1, E.DIST_WIDE, -8, 8)
    E.fill(E.F32, dB, k * n, 2, E.DIST_WIDE, -8, 8)
    torch.cuda.synchronize()

    def sgemm(prec):
        return lambda: E.gemm(E.F32, E.OP_N, E.OP_N, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, prec=prec)

    rows = np.arange(0, m, m // 64)
    res = {}
    for name, prec in (("split3", E.PREC_SPLIT3), ("plain", E.PREC_PLAIN), ("split2", E.PREC_SPLIT2)):
        if args.no_extra and name != "split3":
            continue
        steps = args.steps if name == "split3" else max(3, args.steps // 2)
        prod_cls = E.KC_PRODUCT if name == "plain" else E.KC_PACKED_PRODUCT
        # HIP events around each kernel launch of the timed steps, on its stream
        wall, kern, kt = timed(sgemm(prec), steps, args.warmup, (prod_cls, E.KC_PACK))
        err = sampled_rows_error(dA, dB, dC, m, n, k, rows, np.float32)
        res[name] = {"ms": wall, "event_ms": kern, "tflops": flop / (wall * 1e-3) / 1e12, "normwise_err": err,
                     "product_kernel_ms": kt[prod_cls][0], "product_launches": kt[prod_cls][1],
                     "pack_kernel_ms": kt[E.KC_PACK][0] * kt[E.KC_PACK][1] / steps}
        log(f"{name}: {res[name]}")

    head = res["split3"]
    out = {
        "metric": "TFLOP/s+%peak: 8192^3 sgemm split-BF16 vs FP32 MFMA @equal err; batched GFLOP/s",
        "value": round(head["tflops"], 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "arithmetic": "bf16 x bf16 products, f32 accumulation (3-term split of each f32 operand)",
        "data": "synthetic (device splitmix64, sign*U(1,2)*2^U(-8,8))",
        "config": {"workload": "sgemm C=A*B M=N=K=8192 column-major, split-BF16 3 terms / 6 products on BF16 "
                               "MFMA, FP32 accumulate", "M": m, "N": n, "K": k,
                   "flop_per_step": flop, "flop_counting": "2MNK once"},
        # roofline of the dominant kernel (the packed BF16 product): the
        # algorithmic 2MNK counted once over its own event-timed launch time,
        # against the BF16 MFMA peak
        "roofline": {"bound": "mfma",
                     "achieved": round(flop / (head["product_kernel_ms"] * 1e-3) / 1e12, 2),
                     "peak": PEAK["bf16"], "unit": "TFLOP/s",
                     "frac": round(flop / (head["product_kernel_ms"] * 1e-3) / 1e12 / PEAK["bf16"], 4),
                     "traffic": measured_traffic("split3")},
        # BF16 MFMA work actually issued (6 cross products of 2MNK each) over the
        # same kernel time: the matrix cores' utilisation, not the roofline
        "mfma_utilisation": round(6 * flop / (head["product_kernel_ms"] * 1e-3) / 1e12 / PEAK["bf16"], 4),
        "traffic_source": "profiles/traffic.json: FETCH_SIZE + WRITE_SIZE per launch of this kernel and workload "
                          "from rocprofv3 --pmc runs, each scaled by its reading for a copy of known size",
        "cpu_baseline": None,
        "detail": {"split3": res["split3"]},
    }
    if not args.no_extra:
        out["detail"].update({k_: v for k_, v in res.items() if k_ != "split3"})
        out["detail"]["split3_speedup_vs_plain"] = round(res["split3"]["tflops"] / res["plain"]["tflops"], 3)
        out["detail"]["plain_fp32_frac_of_peak"] = round(res["plain"]["tflops"] / PEAK["fp32"], 4)
        del dA, dB, dC
        torch.cuda.empty_cache()
        out["detail"].update(bench_dgemm(E, torch, dev, args))
        out["detail"]["gemv_16384"] = bench_gemv(E, torch, dev, args)
        out["detail"]["level3_fp64"] = bench_level3(E, torch, dev, args)
        torch.cuda.empty_cache()
        out["detail"]["types"] = bench_types(E, torch, dev, args)
        out["detail"]["batched_strided"] = bench_batched_strided(E, torch, dev, args)
        torch.cuda.empty_cache()
        out["detail"]["batched"] = bench_batched(E, torch, args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
        if not args.no_extra:
            out["detail"]["batched"]["cpu_eigen"] = cpu_baseline_batched(out["detail"]["batched"].pop("_batch"))
            out["detail"]["dgemm_4096"]["cpu_eigen"] = cpu_baseline_dgemm(4096, 4096, 4096)
            out["detail"]["dgemm_tall"]["cpu_eigen"] = cpu_baseline_dgemm(16384, 256, 16384)
    elif not args.no_extra:
        out["detail"]["batched"].pop("_batch", None)
    print(json.dumps(out), flush=True)
    return 0


def bench_dgemm(E, torch, dev, args):
    """FP64 dgemm drop-in: C = 1.5*A*B + 0.5*C, entries N(0,1); normwise error on
    16 sampled rows against the exact product of the same inputs."""
    out = {}
    for name, (m, n, k) in (("dgemm_4096", (4096, 4096, 4096)), ("dgemm_tall", (16384, 256, 16384))):
        dA = torch.empty(m * k, dtype=torch.float64, device=dev)
        dB = torch.empty(k * n, dtype=torch.float64, device=dev)
        dC = torch.empty(m * n, dtype=torch.float64, device=dev)
        E.fill(E.F64, dA, m * k, 11, E.DIST_NORMAL, 0, 1)
        E.fill(E.F64, dB, k * n, 12, E.DIST_NORMAL, 0, 1)
        E.fill(E.F64, dC, m * n, 13, E.DIST_NORMAL, 0, 1)
        fn = lambda: E.gemm(E.F64, E.OP_N, E.OP_N, m, n, k, 1.5, dA, m, dB, k, 0.5, dC, m)
        wall, kern, kt = timed(fn, max(3, args.steps // 2), args.warmup, (E.KC_PRODUCT,), warm_ms=WARM_MS)
        tf = 2.0 * m * n * k / (wall * 1e-3) / 1e12
        cfg = E.plan(E.F64, 0, 0, m, n, k)
        out[name] = {"ms": wall, "event_ms": kern, "product_kernel_ms": kt[E.KC_PRODUCT][0], "tflops": tf,
                     "frac_of_fp64_peak": tf / PEAK["fp64"], "config": cfg, "config_tile_split": E.config_info(cfg),
                     "planner": ["heuristic", "mlp", "table"][E.lib().eh_get_tune_mode()]}
        # one more product from a fresh C for the error
        rows = np.arange(0, m, m // 16)
        E.fill(E.F64, dC, m * n, 13, E.DIST_NORMAL, 0, 1)
        idx = torch.from_numpy(rows).to(dev)
        C0_rows = dC.view(n, m)[:, idx].T.cpu().numpy()
        fn()
        torch.cuda.synchronize()
        out[name]["normwise_err"] = sampled_rows_error(dA, dB, dC, m, n, k, rows, np.float64, 1.5, 0.5, C0_rows)
        log(f"{name}: {out[name]}")
        del dA, dB, dC
    return out


def _exact_cols(T, X, cols):
    """T @ X[:, cols] exactly (host), T and X float64 arrays."""
    return exact_product(T, X[:, cols])


def bench_level3(E, torch, dev, args):
    """The other FP64 level-3 routines at order 8192 (B, C 8192 x 8192): TRSM and
    TRMM (left, lower, no transpose; m^2 n flops), SYMM (left, lower; 2 m^2 n),
    SYRK and SYR2K (lower, no transpose, k = 8192; n^2 k and 2 n^2 k).  The
    triangle has a diagonal of 2 and off-diagonal entries U[-1/n, 1/n] (well
    conditioned, so repeated solves stay in the normal range).  After the
    timing, each routine runs once more on fresh inputs and its normwise error
    on 8 sampled columns is taken against the exact product (TRSM: the
    backward error ||T X - B|| / (n u ||T|| ||X||))."""
    n = 8192
    A = torch.empty(n * n, dtype=torch.float64, device=dev)
    B = torch.empty(n * n, dtype=torch.float64, device=dev)
    Cm = torch.empty(n * n, dtype=torch.float64, device=dev)

    def fresh():
        E.fill(E.F64, A, n * n, 31, E.DIST_UNIFORM, -1.0 / n, 1.0 / n)
        A.view(n, n).diagonal().fill_(2.0)
        E.fill(E.F64, B, n * n, 32, E.DIST_UNIFORM, -1, 1)
        E.fill(E.F64, Cm, n * n, 33, E.DIST_UNIFORM, -1, 1)
    fresh()
    nn = float(n) ** 3
    cases = {
        "dtrsm_8192": (nn, lambda: E.trsm(E.F64, E.LEFT, E.LOWER, E.OP_N, E.NON_UNIT, n, n, 1.0, A, n, B, n)),
        "dtrmm_8192": (nn, lambda: E.trmm(E.F64, E.LEFT, E.LOWER, E.OP_N, E.NON_UNIT, n, n, 1.0, A, n, B, n)),
        "dsymm_8192": (2 * nn, lambda: E.symm(E.F64, E.LEFT, E.LOWER, n, n, 1.0, A, n, B, n, 0.5, Cm, n)),
        "dsyrk_8192": (nn, lambda: E.syrk(E.F64, E.LOWER, E.OP_N, n, n, 1.0, B, n, 0.5, Cm, n)),
        "dsyr2k_8192": (2 * nn, lambda: E.syr2k(E.F64, E.LOWER, E.OP_N, n, n, 1.0, A, n, B, n, 0.5, Cm, n)),
    }
    u = float(np.finfo(np.float64).eps) / 2
    cols = np.arange(5, n, n // 8)
    host = lambda t: t.view(n, n).T.cpu().numpy()  # column-major n x n
    out = {}
    for name, (flop, fn) in cases.items():
        wall, kern = timed(fn, 3, 1, warm_ms=WARM_MS)
        tf = flop / (wall * 1e-3) / 1e12
        out[name] = {"ms": wall, "tflops": tf, "frac_of_fp64_peak": tf / PEAK["fp64"]}
        fresh()
        Ah, Bh, Ch = host(A), host(B), host(Cm)
        fn()
        torch.cuda.synchronize()
        lo = np.tril(Ah)
        sym = lo + np.tril(lo, -1).T
        if name == "dtrsm_8192":
            X = host(B)
            res = _exact_cols(lo, X, cols) - Bh[:, cols]
            err = float(np.linalg.norm(np.asarray(res, np.float64))) * np.sqrt(n / len(cols)) / (
                n * u * np.linalg.norm(lo) * np.linalg.norm(X))
        elif name == "dtrmm_8192":
            err = normwise(host(B)[:, cols], _exact_cols(lo, Bh, cols), n, u, np.linalg.norm(lo), np.linalg.norm(Bh),
                           np.sqrt(n / len(cols)))
        elif name == "dsymm_8192":
            ref = _exact_cols(sym, Bh, cols) + 0.5 * np.asarray(Ch[:, cols], np.longdouble)
            err = normwise(host(Cm)[:, cols], ref, n, u, np.linalg.norm(sym), np.linalg.norm(Bh), np.sqrt(n / len(cols)))
        else:  # lower triangle of the rank-k / rank-2k update, on the sampled columns
            if name == "dsyrk_8192":
                ref = exact_product(Bh, Bh[cols, :].T) + 0.5 * np.asarray(Ch[:, cols], np.longdouble)
                nab = np.linalg.norm(Bh) ** 2
            else:
                ref = (exact_product(Ah, Bh[cols, :].T) + exact_product(Bh, Ah[cols, :].T)
                       + 0.5 * np.asarray(Ch[:, cols], np.longdouble))
                nab = 2 * np.linalg.norm(Ah) * np.linalg.norm(Bh)
            got = host(Cm)[:, cols]
            mask = np.arange(n)[:, None] >= cols[None, :]
            d = np.where(mask, np.asarray(got, np.longdouble) - ref, 0)
            err = float(np.linalg.norm(np.asarray(d, np.float64))) * np.sqrt(n / len(cols)) / (n * u * nab)
        out[name]["normwise_err"] = err
        log(f"{name}: {out[name]}")
    del A, B, Cm
    return out


def bench_batched_strided(E, torch, dev, args):
    """Arrays of fixed-size matrices through eh_gemm_batched_strided: 100,000
    products of 4x4 FP32 and 20,000 of 16x16 FP64 (C = A*B, U[-1,1]);
    algorithmic bytes = A, B read and C written once; `ms` is the wall time
    per call (for such small batches mostly the host's call and launch),
    `kernel_ms` the product kernel's own HIP-event time.  Error: the largest
    normwise error over every 100th product against the exact product."""
    out = {}
    for name, dtype, tdt, es, size, count in (("f32_4x4x4", E.F32, torch.float32, 4, 4, 100000),
                                              ("f64_16x16x16", E.F64, torch.float64, 8, 16, 20000)):
        e = size * size
        A = torch.rand(e * count, dtype=tdt, device=dev) * 2 - 1
        B = torch.rand(e * count, dtype=tdt, device=dev) * 2 - 1
        Cm = torch.zeros(e * count, dtype=tdt, device=dev)
        fn = lambda: E.gemm_batched_strided(dtype, 0, 0, size, size, size, 1.0, A, size, e, B, size, e, 0.0, Cm, size,
                                            e, count)
        wall, kern, kt = timed(fn, max(10, args.steps), args.warmup, (E.KC_GROUPED,), warm_ms=WARM_MS)
        kms = kt[E.KC_GROUPED][0]
        flop, nbytes = 2.0 * size ** 3 * count, 3.0 * e * es * count
        u = float(np.finfo(np.float32 if dtype == E.F32 else np.float64).eps) / 2
        Ah = A.view(count, size, size).transpose(1, 2).cpu().numpy()
        Bh = B.view(count, size, size).transpose(1, 2).cpu().numpy()
        Ch = Cm.view(count, size, size).transpose(1, 2).cpu().numpy()
        errs = [normwise(Ch[p], exact_product(Ah[p], Bh[p]), size, u, np.linalg.norm(Ah[p]), np.linalg.norm(Bh[p]))
                for p in range(0, count, 100)]
        out[name] = {"products": count, "ms": wall, "gflops": flop / (wall * 1e-3) / 1e9,
                     "gbs": nbytes / (wall * 1e-3) / 1e9, "frac_of_hbm": nbytes / (wall * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "kernel_ms": kms, "kernel_gbs": nbytes / (kms * 1e-3) / 1e9 if kms > 0 else None,
                     "kernel_frac_of_hbm": nbytes / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS if kms > 0 else None,
                     "normwise_err_max": float(np.max(errs))}
        del A, B, Cm
    log(f"batched_strided: {out}")
    return out


def bench_types(E, torch, dev, args):
    """Other scalar types: bf16 and IEEE half inputs with an FP32 result (8192^3,
    U[-1,1] rounded to bf16 / half; 2mnk flops) and complex products cgemm / zgemm (4096^3,
    real and imaginary parts U[-1,1]; 8mnk real flops).  Errors on sampled
    rows against the exact product of the same (widened) inputs."""
    out = {}
    n = 8192
    A = (torch.rand(n * n, device=dev) * 2 - 1).to(torch.bfloat16)
    B = (torch.rand(n * n, device=dev) * 2 - 1).to(torch.bfloat16)
    Cm = torch.zeros(n * n, device=dev)
    fn = lambda: E.gemm_bf16(0, 0, n, n, n, 1.0, A.data_ptr(), n, B.data_ptr(), n, 0.0, Cm, n)
    wall, kern, kt = timed(fn, max(3, args.steps // 2), args.warmup, (E.KC_PACKED_PRODUCT,), warm_ms=WARM_MS)
    tf = 2.0 * n ** 3 / (wall * 1e-3) / 1e12
    out["bf16_8192"] = {"ms": wall, "tflops": tf, "frac_of_bf16_peak": tf / PEAK["bf16"],
                        "product_kernel_ms": kt[E.KC_PACKED_PRODUCT][0],
                        "normwise_err": sampled_rows_error(A, B, Cm, n, n, n, np.arange(0, n, n // 32), np.float32)}
    del A, B, Cm
    # IEEE half inputs, FP32 result, on the FP16 matrix cores (same tiles)
    A = (torch.rand(n * n, device=dev) * 2 - 1).to(torch.float16)
    B = (torch.rand(n * n, device=dev) * 2 - 1).to(torch.float16)
    Cm = torch.zeros(n * n, device=dev)
    fn = lambda: E.gemm_f16(0, 0, n, n, n, 1.0, A.data_ptr(), n, B.data_ptr(), n, 0.0, Cm, n)
    wall, kern, kt = timed(fn, max(3, args.steps // 2), args.warmup, (E.KC_PACKED_PRODUCT,), warm_ms=WARM_MS)
    tf = 2.0 * n ** 3 / (wall * 1e-3) / 1e12
    out["f16_8192"] = {"ms": wall, "tflops": tf, "frac_of_fp16_peak": tf / PEAK["fp16"],
                       "product_kernel_ms": kt[E.KC_PACKED_PRODUCT][0],
                       "normwise_err": sampled_rows_error(A, B, Cm, n, n, n, np.arange(0, n, n // 32), np.float32)}
    del A, B, Cm
    n = 4096
    rows = np.arange(0, n, n // 16)
    for name, dtype, tdt, peak in (("cgemm_4096", E.F32, torch.float32, "fp32"),
                                   ("zgemm_4096", E.F64, torch.float64, "fp64")):
        A = torch.rand(2 * n * n, dtype=tdt, device=dev) * 2 - 1
        B = torch.rand(2 * n * n, dtype=tdt, device=dev) * 2 - 1
        Cm = torch.rand(2 * n * n, dtype=tdt, device=dev) * 2 - 1
        fn = lambda: E.gemm_complex(dtype, 0, 0, n, n, n, 1.0 + 0.5j, A, n, B, n, 0.5, Cm, n, prec=E.PREC_PLAIN)
        wall, kern = timed(fn, max(3, args.steps // 2), args.warmup, warm_ms=WARM_MS)
        tf = 8.0 * n ** 3 / (wall * 1e-3) / 1e12
        out[name] = {"ms": wall, "tflops": tf, f"frac_of_{peak}_peak": tf / PEAK[peak]}
        # error of one more product from a fresh C
        Cm.uniform_(-1, 1)
        cz = lambda t: t.view(-1, 2)[:, 0].double() + 1j * t.view(-1, 2)[:, 1].double()
        Ah = cz(A).view(n, n).T[rows].cpu().numpy()
        Bh = cz(B).view(n, n).T.cpu().numpy()
        C0 = cz(Cm).view(n, n).T[rows].cpu().numpy()
        fn()
        torch.cuda.synchronize()
        got = cz(Cm).view(n, n).T[rows].cpu().numpy()
        pr = (exact_product(Ah.real, Bh.real) - exact_product(Ah.imag, Bh.imag)
              + 1j * (exact_product(Ah.real, Bh.imag) + exact_product(Ah.imag, Bh.real)))
        ref = (1.0 + 0.5j) * pr + 0.5 * C0
        u = float(np.finfo(np.float32 if dtype == E.F32 else np.float64).eps) / 2
        nA = float(torch.linalg.norm(A.double())) * abs(1.0 + 0.5j)
        out[name]["normwise_err"] = float(np.linalg.norm(np.asarray(got - ref, np.complex128))) * np.sqrt(
            n / len(rows)) / (n * u * nA * float(torch.linalg.norm(B.double())))
        del A, B, Cm
    log(f"types: {out}")
    return out


def bench_gemv(E, torch, dev, args):
    """Matrix-vector products (the n == 1 path): y = A*x and y = A^T*x with A
    16384 x 16384, FP32 and FP64, U[-1,1]; algorithmic bytes = the matrix read
    once (the vectors are 0.01% of it).  Error: normwise against the exact
    product of the same inputs."""
    out = {}
    n = 16384
    for dtype, name, tdt, es in ((E.F32, "f32", torch.float32, 4), (E.F64, "f64", torch.float64, 8)):
        A = torch.empty(n * n, dtype=tdt, device=dev)
        x = torch.empty(n, dtype=tdt, device=dev)
        y = torch.zeros(n, dtype=tdt, device=dev)
        E.fill(dtype, A, n * n, 21, E.DIST_UNIFORM, -1, 1)
        E.fill(dtype, x, n, 22, E.DIST_UNIFORM, -1, 1)
        Ah = A.view(n, n).T.double().cpu().numpy()
        xh = x.double().cpu().numpy()[:, None]
        u = float(np.finfo(np.float32 if dtype == E.F32 else np.float64).eps) / 2
        for ta in (0, 1):
            fn = lambda: E.gemm(dtype, ta, 0, n, 1, n, 1.0, A, n, x, n, 0.0, y, n)
            wall, kern, kt = timed(fn, max(10, args.steps), args.warmup, (E.KC_GEMV,), warm_ms=WARM_MS)
            kms = kt[E.KC_GEMV][0]
            gbs = n * n * es / (kms * 1e-3) / 1e9
            ref = exact_product(Ah.T if ta else Ah, xh)[:, 0]
            err = normwise(y.double().cpu().numpy(), ref, n, u, np.linalg.norm(Ah), np.linalg.norm(xh))
            out[f"{name}_{'T' if ta else 'N'}"] = {"ms": wall, "event_ms": kern, "kernel_ms": kms,
                                                   "gbs": gbs, "frac_of_hbm": gbs / HBM_PEAK_GBS,
                                                   "normwise_err": err}
        del A, x, y
    log(f"gemv: {out}")
    return out


def bench_batched(E, torch, args):
    """20,000 uneven products (M, N, K log-uniform in [4, 256], half FP32, half FP64, U[-1,1]):
    the persistent grouped kernel replayed as a hipGraph vs one launch per product.
    Error: the largest normwise error over every 100th product against the exact product."""
    from eigen_amd.batched import operand, uneven_batch
    b = uneven_batch(20000, seed=0)
    g = E.Grouped(b.problems)
    wall_g, kern_g = timed(g.graph_launch, max(5, args.steps), args.warmup, warm_ms=WARM_MS)
    wall_r, _, kt = timed(g.run, max(5, args.steps // 2), 1, (E.KC_GROUPED,))
    wall_p, _ = timed(g.run_per_product, 2, 1)
    g.graph_launch()
    torch.cuda.synchronize()
    errs = []
    for p in range(0, len(b.problems), 100):
        A_, B_, C_ = (operand(b, p, w) for w in range(3))
        u = float(np.finfo(A_.dtype).eps) / 2
        errs.append(normwise(C_, exact_product(A_, B_), A_.shape[1], u, np.linalg.norm(A_), np.linalg.norm(B_)))
    kernel_ms = kt[E.KC_GROUPED][0]
    res = {"products": len(b.problems), "flop": b.flops, "bytes": b.bytes,
           "graph_ms": wall_g, "graph_event_ms": kern_g, "run_ms": wall_r, "kernel_ms": kernel_ms,
           "per_product_ms": wall_p,
           "gflops": b.flops / (wall_g * 1e-3) / 1e9, "gbs": b.bytes / (wall_g * 1e-3) / 1e9,
           "frac_of_hbm": b.bytes / (wall_g * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "kernel_frac_of_hbm": b.bytes / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "speedup_vs_per_product": wall_p / wall_g, "normwise_err_max": float(np.max(errs))}
    log(f"batched: {res}")
    g.close()
    res["_batch"] = b
    return res


def cpu_baseline_dgemm(m: int, n: int, k: int):
    """Eigen's CPU dgemm semantics C = 1.5*A*B + 0.5*C (C *= beta, then
    C.noalias() += alpha*A*B, as blas/level3_impl.h does) on host Maps of the
    same shape and distribution, all host cores."""
    try:
        import oracle as O
        if not O.ref_available():
            return None
        L = O.ref()
        A = O.fill((m, k), np.float64, 11, O.NORMAL, 0, 1)
        B = O.fill((k, n), np.float64, 12, O.NORMAL, 0, 1)
        C0 = O.fill((m, n), np.float64, 13, O.NORMAL, 0, 1)
        ts = []
        for r in range(6):  # one warm-up, median of five
            Cm = np.asfortranarray(C0.copy())
            t0 = time.perf_counter()
            L.ref_gemm_beta_f64(0, 0, m, n, k, 1.5, O._ptr(A), m, O._ptr(B), k, 0.5, O._ptr(Cm), m)
            if r:
                ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        return {"tflops": round(2.0 * m * n * k / t / 1e12, 4), "ms": round(t * 1e3, 2),
                "threads": int(L.ref_get_threads()), "kind": "reference",
                "sample": "Eigen C = 1.5*A*B + 0.5*C (dgemm semantics), FP64 N(0,1), median of 5 after 1 warm-up; "
                          + host_cpu()}
    except Exception as e:
        log("cpu dgemm baseline failed:", e)
        return None


def cpu_baseline_batched(b):
    """Eigen's CPU loop over the same 20,000 products (OpenMP dynamic over products,
    Eigen threading off inside each), on host copies of the device operands."""
    try:
        import ctypes as C
        import oracle as O
        if not O.ref_available():
            return None
        L = O.ref()
        host = {d: t.cpu().numpy() for d, t in b.buffers.items()}
        cnt = len(b.problems)
        ptrs = [(C.c_void_p * cnt)() for _ in range(3)]
        for p in range(cnt):
            h = host[int(b.dtypes[p])]
            for w in range(3):
                ptrs[w][p] = h.ctypes.data + h.itemsize * int(b.offsets[p, w])
        dt = np.ascontiguousarray(b.dtypes, dtype=np.int32)
        ms, ns, ks = (np.ascontiguousarray(b.dims[:, i], dtype=np.int32) for i in range(3))
        args_ = [cnt, O._ptr(dt), O._ptr(ms), O._ptr(ns), O._ptr(ks), ptrs[0], ptrs[1], ptrs[2]]
        L.ref_grouped(*args_, 1)  # warm-up
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            L.ref_grouped(*args_, 1)
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        return {"ms": t * 1e3, "gflops": b.flops / t / 1e9, "threads": int(L.ref_get_threads()),
                "kind": "reference", "sample": "Eigen c.noalias() = a*b per product, OpenMP schedule(dynamic) "
                                                "over the 20,000 products, median of 5; " + host_cpu()}
    except Exception as e:
        log("cpu batched baseline failed:", e)
        return None


if __name__ == "__main__":
    sys.exit(main())
