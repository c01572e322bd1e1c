"""The learned tile / schedule selector: training and evaluation.

The network (7 -> 64 ReLU -> 24 softmax, FP32) lives in libeigen_amd and is
trained on the device by the library's own HIP kernels (eh_mlp_train:
forward, back-propagation, Adam).  This module prepares its data from the
autotuner's timing table and measures it:

* targets: soft labels q_c proportional to exp(-(t_c / t_best - 1) / tau)
  over the 24 measured configurations of each shape;
* regret of a choice c on a shape: t_c / t_best - 1 against the exhaustive
  sweep (the yardstick);
* the method without the network is the library's closed-form heuristic
  (eh_plan in TUNE_HEURISTIC mode).

    python -m eigen_amd.selector --db tune.csv --out selector.json
"""
from __future__ import annotations

import argparse
import csv
import ctypes as C
import json
import os
import sys

import numpy as np

from . import TUNE_HEURISTIC, check, lib, plan, set_tune_mode

N_TRAIN = 4000   # random shapes used for training
N_SAMPLE = 4500  # random shapes in the table (the rest are held out)


def load_db(path: str, with_index: bool = False):
    """Returns (shapes (S,6) int64: dtype, ta, tb, m, n, k;  times (S,24) float64 ms)
    sorted by row index, and the indices themselves when with_index."""
    rows = []
    with open(path) as f:
        r = csv.reader(f)
        next(r)
        for row in r:
            if row:
                rows.append(row)
    rows.sort(key=lambda x: int(x[0]))
    index = np.array([int(row[0]) for row in rows], dtype=np.int64)
    shapes = np.array([[int(v) for v in row[1:7]] for row in rows], dtype=np.int64)
    times = np.array([[float(v) for v in row[7:]] for row in rows], dtype=np.float64)
    return (shapes, times, index) if with_index else (shapes, times)


def features(shapes: np.ndarray) -> np.ndarray:
    L = lib()
    X = np.empty((len(shapes), 7), dtype=np.float32)
    for i, (d, ta, tb, m, n, k) in enumerate(shapes.tolist()):
        check(L.eh_mlp_features(d, ta, tb, m, n, k, X[i].ctypes.data_as(C.c_void_p)), "eh_mlp_features")
    return X


def soft_targets(times: np.ndarray, tau: float = 0.05) -> np.ndarray:
    rel = times / times.min(axis=1, keepdims=True) - 1.0
    w = np.exp(-rel / tau)
    return (w / w.sum(axis=1, keepdims=True)).astype(np.float32)


def regret(choice: np.ndarray, times: np.ndarray) -> np.ndarray:
    return times[np.arange(len(times)), choice] / times.min(axis=1) - 1.0


def summary(r: np.ndarray) -> dict:
    return {"mean": float(r.mean()), "median": float(np.median(r)), "p90": float(np.quantile(r, 0.9)),
            "max": float(r.max()), "geomean_slowdown": float(np.exp(np.mean(np.log1p(r)))) - 1.0}


def heuristic_choices(shapes: np.ndarray) -> np.ndarray:
    prev = lib().eh_get_tune_mode()
    set_tune_mode(TUNE_HEURISTIC)
    try:
        return np.array([plan(d, ta, tb, m, n, k) for d, ta, tb, m, n, k in shapes.tolist()], dtype=np.int64)
    finally:
        set_tune_mode(prev)


class Selector:
    """A device-resident network (handle of eh_mlp_create)."""

    def __init__(self, capacity: int = 8192):
        h = C.c_void_p()
        check(lib().eh_mlp_create(7, 64, 24, capacity, C.byref(h)), "eh_mlp_create")
        self.h = h
        self.n_params = int(lib().eh_mlp_num_params(h))

    def params(self) -> np.ndarray:
        P = np.empty(self.n_params, dtype=np.float32)
        check(lib().eh_mlp_get_params(self.h, P.ctypes.data_as(C.c_void_p)), "eh_mlp_get_params")
        return P

    def set_params(self, P) -> None:
        P = np.ascontiguousarray(P, dtype=np.float32)
        check(lib().eh_mlp_set_params(self.h, P.ctypes.data_as(C.c_void_p)), "eh_mlp_set_params")

    def train(self, X: np.ndarray, Q: np.ndarray, epochs: int = 3000, lr: float = 1e-2, wd: float = 1e-4):
        import torch
        dX = torch.from_numpy(np.ascontiguousarray(X, np.float32)).cuda()
        dQ = torch.from_numpy(np.ascontiguousarray(Q, np.float32)).cuda()
        curve = torch.empty(epochs, dtype=torch.float32, device="cuda")
        s = torch.cuda.current_stream().cuda_stream
        check(lib().eh_mlp_train(self.h, dX.data_ptr(), dQ.data_ptr(), len(X), epochs, lr, wd, curve.data_ptr(), s),
              "eh_mlp_train")
        torch.cuda.synchronize()
        return curve.cpu().numpy()

    def predict(self, X: np.ndarray) -> np.ndarray:
        import torch
        dX = torch.from_numpy(np.ascontiguousarray(X, np.float32)).cuda()
        out = torch.empty(len(X), dtype=torch.int32, device="cuda")
        s = torch.cuda.current_stream().cuda_stream
        check(lib().eh_mlp_predict(self.h, dX.data_ptr(), len(X), out.data_ptr(), s), "eh_mlp_predict")
        torch.cuda.synchronize()
        return out.cpu().numpy().astype(np.int64)

    def install(self) -> None:
        """Make this network the planner's selector (TUNE_MLP mode)."""
        import torch
        check(lib().eh_planner_set_mlp(self.h, torch.cuda.current_stream().cuda_stream), "eh_planner_set_mlp")

    def close(self) -> None:
        if self.h:
            lib().eh_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def train_and_evaluate(db: str, epochs: int = 3000, lr: float = 1e-2, wd: float = 1e-4, tau: float = 0.05):
    """Trains on the random shapes 0 .. N_TRAIN-1 and the canonical shapes
    (indices past the random sample), holds out random shapes N_TRAIN .. N_SAMPLE-1."""
    shapes, times, index = load_db(db, with_index=True)
    te = (index >= N_TRAIN) & (index < N_SAMPLE)
    tr = ~te
    X = features(shapes)
    Q = soft_targets(times, tau)
    sel = Selector(capacity=len(shapes))
    curve = sel.train(X[tr], Q[tr], epochs=epochs, lr=lr, wd=wd)
    pred = sel.predict(X[te])
    heur = heuristic_choices(shapes[te])
    res = {"train_shapes": int(shapes[tr].shape[0]), "heldout_shapes": int(shapes[te].shape[0]),
           "loss_first": float(curve[0]), "loss_last": float(curve[-1]),
           "mlp_regret": summary(regret(pred, times[te])),
           "heuristic_regret": summary(regret(heur, times[te])),
           "mlp_train_regret": summary(regret(sel.predict(X[tr]), times[tr]))}
    return sel, res


def tuned_table(db: str):
    """(dtype, ta, tb, m, n, k, best config) of the canonical shapes of a timing table."""
    shapes, times, index = load_db(db, with_index=True)
    sel = index >= N_SAMPLE
    return [tuple(int(v) for v in s) + (int(t.argmin()),) for s, t in zip(shapes[sel], times[sel])]


def emit_inc(params, path: str, source: str, tuned=()) -> None:
    """Write the parameters as the C array the library compiles in as its default
    selector, and the measured best configuration of each tuned canonical shape."""
    vals = [float(v) for v in params]
    assert len(vals) == 2072
    with open(path, "w") as f:
        f.write("// Default tile / schedule selector parameters (7 -> 64 -> 24 network) and the measured\n")
        f.write("// best configuration of the tuned canonical shapes, generated by\n")
        f.write(f"// `python -m eigen_amd.selector --emit-inc` from {source}.\n")
        f.write("static const float kDefaultSelectorParams[2072] = {\n")
        for i in range(0, len(vals), 6):
            f.write("    " + ", ".join(f"{v:.9e}f" for v in vals[i:i + 6]) + ",\n")
        f.write("};\n")
        f.write("// dtype, transA, transB, m, n, k, best configuration\n")
        f.write(f"static const long long kTunedShapes[{max(len(tuned), 1)}][7] = {{\n")
        for row in tuned:
            f.write("    {" + ", ".join(str(v) for v in row) + "},\n")
        if not tuned:
            f.write("    {-1, 0, 0, 0, 0, 0, 0},\n")
        f.write("};\n")
        f.write(f"static const int kNumTunedShapes = {len(tuned)};\n")


def _device_arch() -> str:
    import torch
    return torch.cuda.get_device_properties(0).gcnArchName.split(":")[0] if torch.cuda.is_available() else "unknown"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--db")
    ap.add_argument("--out")
    ap.add_argument("--emit-inc", nargs=2, metavar=("JSON", "INC"),
                    help="write the parameters of a trained JSON as the library's compiled-in default "
                         "(with --db: also the best configuration of its canonical shapes)")
    ap.add_argument("--epochs", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=1e-2)
    ap.add_argument("--wd", type=float, default=1e-4)
    args = ap.parse_args(argv)
    if args.emit_inc:
        src, inc = args.emit_inc
        tuned = tuned_table(args.db) if args.db else ()
        source = os.path.basename(src) + (f" and {os.path.basename(args.db)}" if args.db else "")
        emit_inc(json.load(open(src))["params"], inc, source, tuned)
        return 0
    if not args.db or not args.out:
        ap.error("--db and --out are required for training")
    sel, res = train_and_evaluate(args.db, args.epochs, args.lr, args.wd)
    # what the network's outputs mean: the device it was timed on and the
    # (tile_m, tile_n, split_k) of each configuration it chooses among
    from . import config_info, num_configs
    res = {"arch": _device_arch(), "configs": [list(config_info(c)) for c in range(num_configs())], **res}
    res["params"] = [float(v) for v in sel.params()]
    with open(args.out, "w") as f:
        json.dump(res, f)
    print(json.dumps({k: v for k, v in res.items() if k != "params"}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
