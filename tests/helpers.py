"""Host <-> device plumbing for the tests (column-major numpy arrays <-> flat device buffers)."""
import numpy as np


def to_dev(a: np.ndarray):
    import torch
    flat = np.asfortranarray(a).reshape(-1, order="F")
    return torch.from_numpy(flat.copy()).to("cuda:0")


def from_dev(t, shape, dtype=None) -> np.ndarray:
    import torch
    torch.cuda.synchronize()
    h = t.cpu().numpy()
    if dtype is not None:
        h = h.astype(dtype)
    return h[: shape[0] * shape[1]].reshape(shape, order="F")


def empty_dev(n: int, dtype):
    import torch
    tdt = torch.float32 if np.dtype(dtype) == np.float32 else torch.float64
    return torch.empty(max(n, 1), dtype=tdt, device="cuda:0")


def eh_dtype(dtype) -> int:
    return 0 if np.dtype(dtype) == np.float32 else 1


def unit_roundoff(dtype) -> float:
    return float(np.finfo(dtype).eps) / 2.0


def op_shape(trans: bool, rows: int, cols: int):
    """Stored shape of an operand whose op() is rows x cols."""
    return (cols, rows) if trans else (rows, cols)
