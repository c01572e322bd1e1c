"""Short driver for profiler runs, and the summariser of their output.

    python3 -m eigen_amd.profile_kernels run {split3,plain,dgemm,batched,dtrsm,zgemm,bf16,copy} [reps] [warmup]
    python3 -m eigen_amd.profile_kernels summarize <rocprofv3 output root> [traffic.json] [--skip N]

`run` launches one workload `warmup` times (default 2), then `reps` times
with the library's kernel event timing on (HIP events around each launch of
the dominant kernel class, on its stream), and prints the mean event-timed
launch of the dominant class ("event_ms ...") and the mean event-timed step
("step_ms ...", events around each whole step): the in-process figures the
profiler's kernel trace of the same command is compared with.  It is the
program under `rocprofv3 --kernel-trace --stats` or `rocprofv3 --pmc`:

  split3   sgemm 8192^3, split-BF16 three terms (pack A, pack B, packed product)
  plain    sgemm 8192^3 on FP32 MFMA
  dgemm    dgemm 4096^3, C = 1.5*A*B + 0.5*C
  batched  20,000 uneven products in one grouped launch (captured graph)
  dtrsm    dtrsm 8192 x 8192 (left, lower, no transpose), diagonal 2, off-diagonal U[-1/n, 1/n]
  zgemm    zgemm 4096^3 (complex128, one real product over 2x2 real blocks)
  bf16     bf16 inputs, FP32 result, 8192^3 (U[-1,1] rounded to bf16)
  copy     a 1 GiB device-to-device copy: calibrates the HBM byte counters

`summarize` reads directories named trace_<workload> (kernel-trace csv; with
--skip N the first N dispatches of every kernel -- the warm-up launches --
are left out of the per-dispatch averages; a stdout.txt there holding the
run's own output adds its event-timed kernel and step figures),
pmc_<workload>_<FETCH_SIZE|WRITE_SIZE> and util_<workload> (counter csv of
one pass over UTIL_COUNTERS) under the given root and
prints per-kernel duration tables and counter readings as markdown, plus the
HBM bytes per launch of each workload's dominant kernel: each counter
reading scaled by bytes-moved / reading of the copy kernel (the copy reads
and writes exactly 2^30 bytes).  With a traffic.json path those entries are
written into that file, replacing the same workloads' entries and keeping
the others.
"""
from __future__ import annotations

import csv
import glob
import json
import os
import sys

COPY_BYTES = 1 << 30
# dominant kernel of each workload (substring of its name)
DOMINANT = {"split3": "gemm_split_packed_kernel", "plain": "gemm_mfma_kernel<float", "dgemm": "gemm_mfma_kernel<double",
            "batched": "grouped_gemm_kernel", "dtrsm": "gemm_mfma_kernel<double", "zgemm": "gemm_mfma_kernel<double",
            "bf16": "gemm_split_packed_kernel", "copy": "copyBuffer"}


KCLASS = {"split3": "KC_PACKED_PRODUCT", "bf16": "KC_PACKED_PRODUCT", "plain": "KC_PRODUCT", "dgemm": "KC_PRODUCT",
          "zgemm": "KC_PRODUCT", "dtrsm": "KC_PRODUCT", "batched": "KC_GROUPED"}


def run(what: str, reps: int, warmup: int = 2) -> None:
    import torch

    import eigen_amd as E

    E.lib()
    step = _workload(E, torch, what)
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    cls = getattr(E, KCLASS[what]) if what in KCLASS else None
    if cls is not None:
        E.kernel_timing(True)
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for e0, e1 in ev:
        e0.record(stream)
        step()
        e1.record(stream)
    torch.cuda.synchronize()
    step_ms = sum(e0.elapsed_time(e1) for e0, e1 in ev) / max(reps, 1)
    if cls is not None:
        ms, n = E.kernel_timing_read(cls)
        E.kernel_timing(False)
        print(f"event_ms {what} {KCLASS[what]} {ms / max(n, 1):.4f} over {n} launches", flush=True)
    print(f"step_ms {what} {step_ms:.4f} over {reps} steps (HIP events around each step on its stream)", flush=True)
    print("done", what, flush=True)


def _workload(E, torch, what: str):
    """The workload's inputs on the device and a function running one step."""
    if what in ("split3", "plain"):
        n = 8192
        A = torch.empty(n * n, dtype=torch.float32, device="cuda")
        B = torch.empty(n * n, dtype=torch.float32, device="cuda")
        C = torch.empty(n * n, dtype=torch.float32, device="cuda")
        E.fill(E.F32, A, n * n, 1, E.DIST_WIDE, -8, 8)
        E.fill(E.F32, B, n * n, 2, E.DIST_WIDE, -8, 8)
        prec = E.PREC_SPLIT3 if what == "split3" else E.PREC_PLAIN
        return lambda: E.gemm(E.F32, 0, 0, n, n, n, 1.0, A, n, B, n, 0.0, C, n, prec=prec)
    elif what == "dgemm":
        n = 4096
        A, B, C = (torch.empty(n * n, dtype=torch.float64, device="cuda") for _ in range(3))
        for i, t in enumerate((A, B, C)):
            E.fill(E.F64, t, n * n, 11 + i, E.DIST_NORMAL, 0, 1)
        return lambda: E.gemm(E.F64, 0, 0, n, n, n, 1.5, A, n, B, n, 0.5, C, n)
    elif what == "batched":
        from eigen_amd.batched import uneven_batch
        b = uneven_batch(20000, seed=0)
        g = E.Grouped(b.problems)
        return lambda: (g.run(), b)  # b owns the operand buffers
    elif what == "dtrsm":
        n = 8192
        A = torch.empty(n * n, dtype=torch.float64, device="cuda")
        B = torch.empty(n * n, dtype=torch.float64, device="cuda")
        E.fill(E.F64, A, n * n, 31, E.DIST_UNIFORM, -1.0 / n, 1.0 / n)
        A.view(n, n).diagonal().fill_(2.0)
        E.fill(E.F64, B, n * n, 32, E.DIST_UNIFORM, -1, 1)
        return lambda: E.trsm(E.F64, E.LEFT, E.LOWER, E.OP_N, E.NON_UNIT, n, n, 1.0, A, n, B, n)
    elif what == "zgemm":
        n = 4096
        A, B, Cm = (torch.rand(2 * n * n, dtype=torch.float64, device="cuda") * 2 - 1 for _ in range(3))
        return lambda: E.gemm_complex(E.F64, 0, 0, n, n, n, 1.0 + 0.5j, A, n, B, n, 0.5, Cm, n)
    elif what == "bf16":
        n = 8192
        A = (torch.rand(n * n, device="cuda") * 2 - 1).to(torch.bfloat16)
        B = (torch.rand(n * n, device="cuda") * 2 - 1).to(torch.bfloat16)
        Cm = torch.zeros(n * n, device="cuda")
        return lambda: E.gemm_bf16(0, 0, n, n, n, 1.0, A.data_ptr(), n, B.data_ptr(), n, 0.0, Cm, n)
    elif what == "copy":
        x = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        return lambda: y.copy_(x)
    raise SystemExit(f"unknown workload {what}")


def _one(d: str, pattern: str):
    f = sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))
    return list(csv.DictReader(open(f[0]))) if f else None


def kernel_table(d: str, skip: int = 0):
    """Per-kernel table of a --kernel-trace --stats run.  The stats columns
    cover every dispatch; 'timed avg' leaves out the first `skip` dispatches
    of each kernel (the warm-up launches) using the per-dispatch trace."""
    rows = _one(d, "*kernel_stats.csv")
    if rows is None:
        return None, {}
    per = {}
    for r in _one(d, "*kernel_trace.csv") or []:
        per.setdefault(r["Kernel_Name"], []).append(
            (int(r["Dispatch_Id"]), (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3))
    out = ["| kernel | calls | avg (us) | min (us) | max (us) | % of GPU time | timed avg (us) |",
           "|---|---|---|---|---|---|---|"]
    avg = {}
    for r in rows:
        name = r["Name"]
        ds = [t for _, t in sorted(per.get(name, []))][skip:]
        avg[name] = sum(ds) / len(ds) if ds else float(r["AverageNs"]) / 1e3
        short = name if len(name) <= 100 else name[:100] + "..."
        out.append(f"| `{short}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | {float(r['MinNs']) / 1e3:.1f} | "
                   f"{float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} | "
                   f"{avg[name]:.1f} ({len(ds)}) |")
    return "\n".join(out), avg


def counter_means(d: str, counter: str) -> dict:
    """Mean over dispatches of the counter (summed over its instances) per kernel name, in KiB."""
    rows = _one(d, "*counter_collection.csv") or []
    per_dispatch = {}
    for r in rows:
        if r["Counter_Name"] != counter:
            continue
        key = (r["Dispatch_Id"], r["Kernel_Name"])
        per_dispatch[key] = per_dispatch.get(key, 0.0) + float(r["Counter_Value"])
    per_kernel = {}
    for (_, name), v in per_dispatch.items():
        per_kernel.setdefault(name, []).append(v)
    return {n: sum(v) / len(v) for n, v in per_kernel.items()}


UTIL_COUNTERS = ("SQ_VALU_MFMA_BUSY_CYCLES GRBM_GUI_ACTIVE SQ_WAVE_CYCLES SQ_WAIT_ANY SQ_WAIT_INST_ANY "
                 "SQ_ACTIVE_INST_ANY SQ_BUSY_CYCLES")
XCDS, SIMDS = 8, 1024  # MI355X: GRBM_GUI_ACTIVE is summed over the 8 XCDs; 256 CUs x 4 SIMDs


def utilisation(d: str, workload: str):
    """Clock and MFMA utilisation of the dominant kernel from one UTIL_COUNTERS pass:
    clock = GRBM_GUI_ACTIVE / XCDS / duration, MFMA busy = SQ_VALU_MFMA_BUSY_CYCLES /
    (GRBM_GUI_ACTIVE / XCDS * SIMDS), and the wave-cycle fractions spent waiting.
    Counters and durations are summed over the dispatches before dividing, so
    a workload of many launches is weighted by time (the mean of per-dispatch
    ratios would let the shortest launches, whose counters barely resolve,
    count as much as the longest)."""
    rows = _one(d, "*counter_collection.csv") or []
    per, dur = {}, {}
    for r in rows:
        if DOMINANT[workload] not in r["Kernel_Name"]:
            continue
        c = per.setdefault(r["Dispatch_Id"], {})
        c[r["Counter_Name"]] = c.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        dur[r["Dispatch_Id"]] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    if not per:
        return None
    tot = {}
    for v in per.values():
        for k, x in v.items():
            tot[k] = tot.get(k, 0.0) + x
    ns = sum(dur.values())
    gui = tot["GRBM_GUI_ACTIVE"] / XCDS
    return {"duration_us": ns / 1e3 / len(per), "dispatches": len(per), "clock_mhz": gui / ns * 1e3,
            "mfma_busy": tot["SQ_VALU_MFMA_BUSY_CYCLES"] / (gui * SIMDS),
            "wait_any": tot["SQ_WAIT_ANY"] / tot["SQ_WAVE_CYCLES"],
            "wait_inst_any": tot["SQ_WAIT_INST_ANY"] / tot["SQ_WAVE_CYCLES"],
            "active_inst_any": tot["SQ_ACTIVE_INST_ANY"] / tot["SQ_WAVE_CYCLES"]}


def _dominant(values: dict, workload: str):
    for name, v in values.items():
        if DOMINANT[workload] in name:
            return name, v
    return None, None


def summarize(root: str, skip: int = 0) -> dict:
    traffic = {}
    scale = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        _, kib = _dominant(counter_means(os.path.join(root, f"pmc_copy_{c}"), c), "copy")
        if kib:
            scale[c] = COPY_BYTES / kib  # bytes per counter unit
    print("## Counter calibration (1 GiB device copy)\n")
    for c, s in scale.items():
        print(f"- {c}: {COPY_BYTES / s:.1f} KiB read by the counter for {COPY_BYTES} bytes -> x{s / 1024:.4f} KiB")
    print()
    for d in sorted(glob.glob(os.path.join(root, "trace_*"))):
        table, _ = kernel_table(d, skip)
        if table:
            print(f"## {os.path.basename(d)}\n\n{table}\n")
            # the same process's own event timings (printed by `run` into stdout.txt)
            log = os.path.join(d, "stdout.txt")
            if os.path.exists(log):
                for line in open(log):
                    if line.startswith(("event_ms", "step_ms")):
                        print(f"- {line.strip()}")
                print()
    workloads = sorted({os.path.basename(p).split("_")[1] for p in glob.glob(os.path.join(root, "pmc_*"))})
    for w in workloads:
        if w == "copy":
            continue
        got = {}
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            name, kib = _dominant(counter_means(os.path.join(root, f"pmc_{w}_{c}"), c), w)
            if kib is not None and c in scale:
                got[c] = (name, kib, kib * scale[c])
        if len(got) == 2:
            name = got["FETCH_SIZE"][0]
            traffic[w] = {"kernel": name, "fetch_bytes": round(got["FETCH_SIZE"][2]),
                          "write_bytes": round(got["WRITE_SIZE"][2]),
                          "bytes": round(got["FETCH_SIZE"][2] + got["WRITE_SIZE"][2]),
                          "fetch_kib_raw": round(got["FETCH_SIZE"][1], 1), "write_kib_raw": round(got["WRITE_SIZE"][1], 1)}
            print(f"## HBM traffic per launch: {w}\n\n`{name[:100]}`: FETCH_SIZE {got['FETCH_SIZE'][1]:.1f} KiB raw -> "
                  f"{got['FETCH_SIZE'][2]:.4g} B; WRITE_SIZE {got['WRITE_SIZE'][1]:.1f} KiB raw -> "
                  f"{got['WRITE_SIZE'][2]:.4g} B; total {traffic[w]['bytes']:.4g} B\n")
    for d in sorted(glob.glob(os.path.join(root, "util_*"))):
        w = os.path.basename(d)[5:]
        u = utilisation(d, w)
        if u:
            print(f"## Clock and MFMA utilisation: {w}\n")
            print("| dispatches | mean duration (us) | clock (MHz) | MFMA busy | waves waiting (waitcnt/barrier) | "
                  "waves waiting to issue | waves issuing |")
            print("|---|---|---|---|---|---|---|")
            print(f"| {u['dispatches']} | {u['duration_us']:.0f} | {u['clock_mhz']:.0f} | {u['mfma_busy']:.1%} | "
                  f"{u['wait_any']:.1%} | {u['wait_inst_any']:.1%} | {u['active_inst_any']:.1%} |\n")
    return traffic


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) >= 2 and argv[0] == "run":
        run(argv[1], int(argv[2]) if len(argv) > 2 else 5, int(argv[3]) if len(argv) > 3 else 2)
        return 0
    if len(argv) >= 2 and argv[0] == "summarize":
        skip = 0
        if "--skip" in argv:
            i = argv.index("--skip")
            skip = int(argv[i + 1])
            argv = argv[:i] + argv[i + 2:]
        traffic = summarize(argv[1], skip)
        if len(argv) > 2:  # merged into the file: workloads profiled in other sessions keep their entries
            merged = {}
            if os.path.exists(argv[2]):
                with open(argv[2]) as f:
                    merged = json.load(f)
            merged.update(traffic)
            with open(argv[2], "w") as f:
                json.dump(dict(sorted(merged.items())), f, indent=1)
                f.write("\n")
        return 0
    print(__doc__)
    return 2


if __name__ == "__main__":
    sys.exit(main())
