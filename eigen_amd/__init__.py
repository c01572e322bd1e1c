"""eigen_amd -- MI355X (gfx950) dense matrix products behind Eigen's interfaces.

The product lives in two in-tree shared libraries built by ``eigen_amd.build``:

* ``_lib/libeigen_amd.so``      -- the C-ABI of ``include/eigen_amd.h``
* ``_lib/libeigen_amd_blas.so`` -- ``sgemm_``/``dgemm_`` (``include/eigen_amd_blas.h``)

This module loads them with ctypes and offers thin helpers over device
buffers (torch tensors on ``cuda`` are used only as device memory and
streams).  Matrices are column-major: an m x n matrix with leading
dimension ld occupies ``ld * n`` elements of a flat buffer.

There is no CPU fallback: if the libraries are missing every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.environ.get("EIGEN_AMD_LIBDIR") or os.path.join(HERE, "_lib")

F32, F64 = 0, 1
OP_N, OP_T, OP_C = 0, 1, 2
PREC_AUTO, PREC_PLAIN, PREC_SPLIT2, PREC_SPLIT3 = 0, 1, 2, 3
TUNE_HEURISTIC, TUNE_MLP, TUNE_TABLE = 0, 1, 2
DIST_UNIFORM, DIST_NORMAL, DIST_WIDE, DIST_INT = 0, 1, 2, 3
SUCCESS, ERR_INVALID, ERR_HIP, ERR_UNSUPPORTED, ERR_ALLOC = 0, 1, 2, 3, 4
UPPER, LOWER = 0, 1
LEFT, RIGHT = 0, 1
NON_UNIT, UNIT = 0, 1

_lib = None
_blas = None


class GemmProblem(C.Structure):
    """eh_gemm_problem (include/eigen_amd.h)."""
    _fields_ = [("dtype", C.c_int32), ("opa", C.c_int32), ("opb", C.c_int32), ("reserved", C.c_int32),
                ("m", C.c_int64), ("n", C.c_int64), ("k", C.c_int64),
                ("lda", C.c_int64), ("ldb", C.c_int64), ("ldc", C.c_int64),
                ("A", C.c_void_p), ("B", C.c_void_p), ("C", C.c_void_p),
                ("alpha", C.c_double), ("beta", C.c_double)]


class Grouped:
    """A batch of independent products run by one persistent kernel."""

    def __init__(self, problems):
        arr = (GemmProblem * len(problems))(*problems)
        h = C.c_void_p()
        check(lib().eh_grouped_create(arr, len(problems), C.byref(h)), "eh_grouped_create")
        self.handle = h

    def run(self, stream=None):
        check(lib().eh_grouped_run(self.handle, current_stream() if stream is None else stream), "eh_grouped_run")

    def graph_launch(self, stream=None):
        check(lib().eh_grouped_graph_launch(self.handle, current_stream() if stream is None else stream),
              "eh_grouped_graph_launch")

    def run_per_product(self, stream=None):
        check(lib().eh_grouped_run_per_product(self.handle, current_stream() if stream is None else stream),
              "eh_grouped_run_per_product")

    def set_schedule(self, schedule: int) -> None:
        check(lib().eh_grouped_set_schedule(self.handle, schedule), "eh_grouped_set_schedule")

    def schedule(self) -> int:
        v = C.c_int()
        check(lib().eh_grouped_get_schedule(self.handle, C.byref(v)), "eh_grouped_get_schedule")
        return v.value

    def info(self):
        t, p = C.c_int64(), C.c_int64()
        check(lib().eh_grouped_info(self.handle, C.byref(t), C.byref(p)), "eh_grouped_info")
        return t.value, p.value

    def close(self):
        if self.handle:
            lib().eh_grouped_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class EigenAmdError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        msg = f"{what}: status {status}"
        if _lib is not None:
            msg += f" ({_lib.eh_status_string(status).decode()}: {_lib.eh_last_error().decode()})"
        super().__init__(msg)
        self.status = status


def lib_path() -> str:
    return os.path.join(LIBDIR, "libeigen_amd.so")


def blas_path() -> str:
    return os.path.join(LIBDIR, "libeigen_amd_blas.so")


def lib():
    """ctypes handle of libeigen_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built: run `python -m eigen_amd.build` (no CPU fallback exists)")
        # One HIP runtime per process: torch bundles its own libamdhip64 with
        # the same soname as the system one this library links.  Loading torch
        # first makes the library bind to that already-loaded runtime; loading
        # the library first would leave torch with a second runtime that finds
        # no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(path)
        vp, i64, i32, u64, f64 = C.c_void_p, C.c_int64, C.c_int, C.c_uint64, C.c_double
        L.eh_gemm.restype = i32
        L.eh_gemm.argtypes = [i32, i32, i32, i64, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, i32, vp]
        L.eh_gemm_noalloc.restype = i32
        L.eh_gemm_noalloc.argtypes = [i32, i32, i32, i64, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, i32, vp]
        L.eh_gemm_config.restype = i32
        L.eh_gemm_config.argtypes = [i32, i32, i32, i64, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, i32, i32, vp]
        L.eh_syrk.restype = i32
        L.eh_syrk.argtypes = [i32, i32, i32, i64, i64, vp, vp, i64, vp, vp, i64, vp]
        L.eh_gemmtr.restype = i32
        L.eh_gemmtr.argtypes = [i32, i32, i32, i32, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, vp]
        L.eh_syr2k.restype = i32
        L.eh_syr2k.argtypes = [i32, i32, i32, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, vp]
        L.eh_symm.restype = i32
        L.eh_symm.argtypes = [i32, i32, i32, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, vp]
        for f in (L.eh_trmm, L.eh_trsm):
            f.restype = i32
            f.argtypes = [i32, i32, i32, i32, i32, i64, i64, vp, vp, i64, vp, i64, vp]
        L.eh_gemm_batched_strided.restype = i32
        L.eh_gemm_batched_strided.argtypes = [i32, i32, i32, i64, i64, i64, vp, vp, i64, i64, vp, i64, i64, vp, vp,
                                              i64, i64, i64, vp]
        L.eh_gemv.restype = i32
        L.eh_gemv.argtypes = [i32, i32, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, vp]
        L.eh_gemv_complex.restype = i32
        L.eh_gemv_complex.argtypes = [i32, i32, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, vp]
        L.eh_gemm_complex.restype = i32
        L.eh_gemm_complex.argtypes = [i32, i32, i32, i64, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, i32, vp]
        L.eh_gemm_bf16.restype = i32
        L.eh_gemm_bf16.argtypes = [i32, i32, i64, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, vp]
        L.eh_gemm_f16.restype = i32
        L.eh_gemm_f16.argtypes = [i32, i32, i64, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, vp]
        L.eh_gemm_strided.restype = i32
        L.eh_gemm_strided.argtypes = [i32, i32, i32, i64, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, i64, i32, vp]
        L.eh_num_configs.restype = i32
        L.eh_config_info.restype = i32
        L.eh_config_info.argtypes = [i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]
        L.eh_plan.restype = i32
        L.eh_plan.argtypes = [i32, i32, i32, i64, i64, i64, C.POINTER(i32)]
        L.eh_set_tune_mode.restype = i32
        L.eh_set_tune_mode.argtypes = [i32]
        L.eh_get_tune_mode.restype = i32
        L.eh_set_sgemm_precision.restype = i32
        L.eh_set_sgemm_precision.argtypes = [i32]
        L.eh_get_sgemm_precision.restype = i32
        L.eh_fill.restype = i32
        L.eh_fill.argtypes = [i32, vp, i64, u64, i32, f64, f64, vp]
        L.eh_split_bf16.restype = i32
        L.eh_split_bf16.argtypes = [vp, i64, i32, vp, vp, vp, vp]
        L.eh_synchronize.restype = i32
        L.eh_synchronize.argtypes = [vp]
        L.eh_status_string.restype = C.c_char_p
        L.eh_status_string.argtypes = [i32]
        L.eh_last_error.restype = C.c_char_p
        L.eh_version.restype = C.c_char_p
        _declare_extra(L)
        _lib = L
    return _lib


def _declare_extra(L) -> None:
    """Signatures of the grouped / tuner / MLP entry points (present in full builds)."""
    vp, i64, i32, f64, f32 = C.c_void_p, C.c_int64, C.c_int, C.c_double, C.c_float
    table = {
        "eh_grouped_create": (i32, [C.POINTER(GemmProblem), i64, C.POINTER(vp)]),
        "eh_grouped_run": (i32, [vp, vp]),
        "eh_grouped_graph_launch": (i32, [vp, vp]),
        "eh_grouped_run_per_product": (i32, [vp, vp]),
        "eh_grouped_destroy": (i32, [vp]),
        "eh_grouped_info": (i32, [vp, C.POINTER(i64), C.POINTER(i64)]),
        "eh_grouped_num_schedules": (i32, []),
        "eh_grouped_schedule_info": (i32, [i32, C.POINTER(i32), C.POINTER(i32)]),
        "eh_grouped_set_schedule": (i32, [vp, i32]),
        "eh_grouped_get_schedule": (i32, [vp, C.POINTER(i32)]),
        "eh_time_config": (i32, [i32, i32, i32, i64, i64, i64, vp, vp, vp, i32, i32, C.POINTER(f32)]),
        "eh_tuner_record": (i32, [i32, i32, i32, i64, i64, i64, i32]),
        "eh_tuner_load": (i32, [C.c_char_p]),
        "eh_tuner_save": (i32, [C.c_char_p]),
        "eh_tuner_count": (i64, []),
        "eh_tuner_clear": (i32, []),
        "eh_num_packed_configs": (i32, [i32]),
        "eh_packed_config_info": (i32, [i32, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
        "eh_plan_packed": (i32, [i32, i32, i32, i64, i64, i64, C.POINTER(i32)]),
        "eh_time_config_packed": (i32, [i32, i32, i32, i64, i64, i64, vp, vp, vp, i32, i32, C.POINTER(f32)]),
        "eh_tuner_record_packed": (i32, [i32, i32, i32, i64, i64, i64, i32]),
        "eh_mlp_create": (i32, [i32, i32, i32, i64, C.POINTER(vp)]),
        "eh_mlp_destroy": (i32, [vp]),
        "eh_mlp_set_params": (i32, [vp, vp]),
        "eh_mlp_get_params": (i32, [vp, vp]),
        "eh_mlp_num_params": (i64, [vp]),
        "eh_mlp_features": (i32, [i32, i32, i32, i64, i64, i64, vp]),
        "eh_mlp_forward": (i32, [vp, vp, i64, vp, vp]),
        "eh_mlp_backward": (i32, [vp, vp, vp, i64, vp, vp, vp]),
        "eh_mlp_adam_step": (i32, [vp, vp, f32, f32, f32, f32, f32, vp]),
        "eh_mlp_train": (i32, [vp, vp, vp, i64, i32, f32, f32, vp, vp]),
        "eh_mlp_predict": (i32, [vp, vp, i64, vp, vp]),
        "eh_planner_set_mlp": (i32, [vp, vp]),
        "eh_planner_set_mlp_params": (i32, [vp]),
        "eh_trsm_complex": (i32, [i32] * 5 + [i64, i64, vp, vp, i64, vp, i64, vp]),
        "eh_trmm_complex": (i32, [i32] * 5 + [i64, i64, vp, vp, i64, vp, i64, vp]),
        "eh_symm_complex": (i32, [i32, i32, i32, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, vp]),
        "eh_hemm": (i32, [i32, i32, i32, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, vp]),
        "eh_syrk_complex": (i32, [i32, i32, i32, i64, i64, vp, vp, i64, vp, vp, i64, vp]),
        "eh_herk": (i32, [i32, i32, i32, i64, i64, vp, vp, i64, vp, vp, i64, vp]),
        "eh_syr2k_complex": (i32, [i32, i32, i32, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, vp]),
        "eh_her2k": (i32, [i32, i32, i32, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, vp]),
        "eh_gemmtr_complex": (i32, [i32, i32, i32, i32, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, vp]),
        "eh_kernel_timing": (i32, [i32]),
        "eh_kernel_timing_read": (i32, [i32, C.POINTER(f64), C.POINTER(i64)]),
    }
    for name, (res, args) in table.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args


def blas():
    """ctypes handle of libeigen_amd_blas.so."""
    global _blas
    if _blas is None:
        lib()
        path = blas_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built")
        _blas = C.CDLL(path)
        _blas.eh_blas_set_stream.argtypes = [C.c_void_p]
        _blas.eh_blas_get_stream.restype = C.c_void_p
    return _blas


def check(status: int, what: str = "eigen_amd call") -> None:
    if status != 0:
        raise EigenAmdError(status, what)


def _scalar(dtype: int, v: float):
    return C.c_float(v) if dtype == F32 else C.c_double(v)


def _ptr(t) -> int | None:
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def current_stream() -> int | None:
    import torch
    return torch.cuda.current_stream().cuda_stream if torch.cuda.is_available() else None


def gemm(dtype: int, opa: int, opb: int, m: int, n: int, k: int, alpha: float, A, lda: int, B, ldb: int,
         beta: float, Cm, ldc: int, prec: int = PREC_AUTO, config: int | None = None, stream=None) -> None:
    """C = alpha*op(A)*op(B) + beta*C on device buffers (torch tensors or raw pointers)."""
    L = lib()
    a, b = _scalar(dtype, alpha), _scalar(dtype, beta)
    s = current_stream() if stream is None else stream
    if config is None:
        st = L.eh_gemm(dtype, opa, opb, m, n, k, C.byref(a), _ptr(A), lda, _ptr(B), ldb, C.byref(b), _ptr(Cm), ldc,
                       prec, s)
    else:
        st = L.eh_gemm_config(dtype, opa, opb, m, n, k, C.byref(a), _ptr(A), lda, _ptr(B), ldb, C.byref(b), _ptr(Cm),
                              ldc, prec, config, s)
    check(st, "eh_gemm")


def syrk(dtype: int, uplo: int, trans: int, n: int, k: int, alpha: float, A, lda: int, beta: float, Cm, ldc: int,
         stream=None) -> None:
    """uplo triangle of C = alpha*op(A)*op(A)^T + beta*C (eh_syrk)."""
    a, b = _scalar(dtype, alpha), _scalar(dtype, beta)
    s = current_stream() if stream is None else stream
    check(lib().eh_syrk(dtype, uplo, trans, n, k, C.byref(a), _ptr(A), lda, C.byref(b), _ptr(Cm), ldc, s), "eh_syrk")


def gemm_batched_strided(dtype: int, opa: int, opb: int, m: int, n: int, k: int, alpha: float, A, lda: int,
                         stride_a: int, B, ldb: int, stride_b: int, beta: float, Cm, ldc: int, stride_c: int,
                         count: int, stream=None) -> None:
    """`count` products of one shape at fixed element strides (eh_gemm_batched_strided)."""
    a, b = _scalar(dtype, alpha), _scalar(dtype, beta)
    s = current_stream() if stream is None else stream
    check(lib().eh_gemm_batched_strided(dtype, opa, opb, m, n, k, C.byref(a), _ptr(A), lda, stride_a, _ptr(B), ldb,
                                        stride_b, C.byref(b), _ptr(Cm), ldc, stride_c, count, s),
          "eh_gemm_batched_strided")


def gemv(dtype: int, trans: int, m: int, n: int, alpha: float, A, lda: int, x, incx: int, beta: float, y,
         incy: int, stream=None) -> None:
    """y = alpha*op(A)*x + beta*y, A m x n, strided vectors (eh_gemv). x and y
    may be device pointers (ints) already offset as BLAS expects."""
    a, b = _scalar(dtype, alpha), _scalar(dtype, beta)
    s = current_stream() if stream is None else stream
    check(lib().eh_gemv(dtype, trans, m, n, C.byref(a), _ptr(A), lda, _ptr(x), incx, C.byref(b), _ptr(y), incy, s),
          "eh_gemv")


def gemm_complex(dtype: int, opa: int, opb: int, m: int, n: int, k: int, alpha: complex, A, lda: int, B, ldb: int,
                 beta: complex, Cm, ldc: int, prec: int = PREC_AUTO, stream=None) -> None:
    """Complex C = alpha*op(A)*op(B) + beta*C on interleaved storage (eh_gemm_complex);
    dtype F32 = complex64, F64 = complex128; op OP_C conjugates."""
    T = C.c_float if dtype == F32 else C.c_double
    a = (T * 2)(complex(alpha).real, complex(alpha).imag)
    b = (T * 2)(complex(beta).real, complex(beta).imag)
    s = current_stream() if stream is None else stream
    check(lib().eh_gemm_complex(dtype, opa, opb, m, n, k, a, _ptr(A), lda, _ptr(B), ldb, b, _ptr(Cm), ldc, prec, s),
          "eh_gemm_complex")


def gemv_complex(dtype: int, trans: int, m: int, n: int, alpha: complex, A, lda: int, x, incx: int, beta: complex,
                 y, incy: int, stream=None) -> None:
    """Complex y = alpha*op(A)*x + beta*y with strided vectors (eh_gemv_complex)."""
    T = C.c_float if dtype == F32 else C.c_double
    a = (T * 2)(complex(alpha).real, complex(alpha).imag)
    b = (T * 2)(complex(beta).real, complex(beta).imag)
    s = current_stream() if stream is None else stream
    check(lib().eh_gemv_complex(dtype, trans, m, n, a, _ptr(A), lda, _ptr(x), incx, b, _ptr(y), incy, s),
          "eh_gemv_complex")


def gemm_bf16(opa: int, opb: int, m: int, n: int, k: int, alpha: float, A, lda: int, B, ldb: int, beta: float, Cm,
              ldc: int, stream=None) -> None:
    """C (float32) = alpha*op(A)*op(B) + beta*C with A, B bfloat16 (eh_gemm_bf16)."""
    a, b = C.c_float(alpha), C.c_float(beta)
    s = current_stream() if stream is None else stream
    check(lib().eh_gemm_bf16(opa, opb, m, n, k, C.byref(a), _ptr(A), lda, _ptr(B), ldb, C.byref(b), _ptr(Cm), ldc, s),
          "eh_gemm_bf16")


def gemm_f16(opa: int, opb: int, m: int, n: int, k: int, alpha: float, A, lda: int, B, ldb: int, beta: float, Cm,
             ldc: int, stream=None) -> None:
    """C (float32) = alpha*op(A)*op(B) + beta*C with A, B IEEE half (eh_gemm_f16)."""
    a, b = C.c_float(alpha), C.c_float(beta)
    s = current_stream() if stream is None else stream
    check(lib().eh_gemm_f16(opa, opb, m, n, k, C.byref(a), _ptr(A), lda, _ptr(B), ldb, C.byref(b), _ptr(Cm), ldc, s),
          "eh_gemm_f16")


def gemmtr(dtype: int, uplo: int, opa: int, opb: int, n: int, k: int, alpha: float, A, lda: int, B, ldb: int,
           beta: float, Cm, ldc: int, stream=None) -> None:
    """uplo triangle of C = alpha*op(A)*op(B) + beta*C (eh_gemmtr)."""
    a, b = _scalar(dtype, alpha), _scalar(dtype, beta)
    s = current_stream() if stream is None else stream
    check(lib().eh_gemmtr(dtype, uplo, opa, opb, n, k, C.byref(a), _ptr(A), lda, _ptr(B), ldb, C.byref(b), _ptr(Cm),
                          ldc, s), "eh_gemmtr")


def syr2k(dtype: int, uplo: int, trans: int, n: int, k: int, alpha: float, A, lda: int, B, ldb: int, beta: float,
          Cm, ldc: int, stream=None) -> None:
    """uplo triangle of C = alpha*(op(A)*op(B)^T + op(B)*op(A)^T) + beta*C (eh_syr2k)."""
    a, b = _scalar(dtype, alpha), _scalar(dtype, beta)
    s = current_stream() if stream is None else stream
    check(lib().eh_syr2k(dtype, uplo, trans, n, k, C.byref(a), _ptr(A), lda, _ptr(B), ldb, C.byref(b), _ptr(Cm), ldc,
                         s), "eh_syr2k")


def symm(dtype: int, side: int, uplo: int, m: int, n: int, alpha: float, A, lda: int, B, ldb: int, beta: float, Cm,
         ldc: int, stream=None) -> None:
    """C = alpha*S*B + beta*C (LEFT) or alpha*B*S + beta*C (RIGHT), S symmetric in A's uplo triangle (eh_symm)."""
    a, b = _scalar(dtype, alpha), _scalar(dtype, beta)
    s = current_stream() if stream is None else stream
    check(lib().eh_symm(dtype, side, uplo, m, n, C.byref(a), _ptr(A), lda, _ptr(B), ldb, C.byref(b), _ptr(Cm), ldc,
                        s), "eh_symm")


def trmm(dtype: int, side: int, uplo: int, transa: int, diag: int, m: int, n: int, alpha: float, A, lda: int, B,
         ldb: int, stream=None) -> None:
    """B = alpha*op(T)*B (LEFT) or alpha*B*op(T) (RIGHT), T the uplo triangle of A (eh_trmm)."""
    a = _scalar(dtype, alpha)
    s = current_stream() if stream is None else stream
    check(lib().eh_trmm(dtype, side, uplo, transa, diag, m, n, C.byref(a), _ptr(A), lda, _ptr(B), ldb, s), "eh_trmm")


def trsm(dtype: int, side: int, uplo: int, transa: int, diag: int, m: int, n: int, alpha: float, A, lda: int, B,
         ldb: int, stream=None) -> None:
    """Solve op(T)*X = alpha*B (LEFT) or X*op(T) = alpha*B (RIGHT) in place in B (eh_trsm)."""
    a = _scalar(dtype, alpha)
    s = current_stream() if stream is None else stream
    check(lib().eh_trsm(dtype, side, uplo, transa, diag, m, n, C.byref(a), _ptr(A), lda, _ptr(B), ldb, s), "eh_trsm")


def gemm_strided(dtype: int, opa: int, opb: int, m: int, n: int, k: int, alpha: float, A, lda: int, B, ldb: int,
                 beta: float, Cm, incc: int, ldc: int, prec: int = PREC_AUTO, stream=None) -> None:
    """Same as gemm with C(i, j) at C[i*incc + j*ldc] (eh_gemm_strided)."""
    a, b = _scalar(dtype, alpha), _scalar(dtype, beta)
    s = current_stream() if stream is None else stream
    check(lib().eh_gemm_strided(dtype, opa, opb, m, n, k, C.byref(a), _ptr(A), lda, _ptr(B), ldb, C.byref(b),
                                _ptr(Cm), incc, ldc, prec, s), "eh_gemm_strided")


def fill(dtype: int, X, n: int, seed: int, dist: int = DIST_UNIFORM, p0: float = -1.0, p1: float = 1.0,
         stream=None) -> None:
    s = current_stream() if stream is None else stream
    check(lib().eh_fill(dtype, _ptr(X), n, seed, dist, p0, p1, s), "eh_fill")


def split_bf16(x, n: int, terms: int, t0, t1=None, t2=None, stream=None) -> None:
    s = current_stream() if stream is None else stream
    check(lib().eh_split_bf16(_ptr(x), n, terms, _ptr(t0), _ptr(t1), _ptr(t2), s), "eh_split_bf16")


# kernel classes of eh_kernel_timing_read
KC_PRODUCT, KC_PACKED_PRODUCT, KC_PACK, KC_GROUPED, KC_GEMV = range(5)


def kernel_timing(enable: bool) -> None:
    """Start (clearing earlier records) or stop recording HIP events around kernel launches."""
    check(lib().eh_kernel_timing(1 if enable else 0), "eh_kernel_timing")


def kernel_timing_read(cls: int):
    """(total ms, launches) of the recorded launches of one kernel class; clears them."""
    ms, n = C.c_double(0.0), C.c_int64(0)
    check(lib().eh_kernel_timing_read(cls, C.byref(ms), C.byref(n)), "eh_kernel_timing_read")
    return ms.value, n.value


def num_configs() -> int:
    return lib().eh_num_configs()


def config_info(config: int):
    tm, tn, sk = C.c_int(), C.c_int(), C.c_int()
    check(lib().eh_config_info(config, C.byref(tm), C.byref(tn), C.byref(sk)), "eh_config_info")
    return tm.value, tn.value, sk.value


def plan(dtype: int, opa: int, opb: int, m: int, n: int, k: int) -> int:
    c = C.c_int()
    check(lib().eh_plan(dtype, opa, opb, m, n, k, C.byref(c)), "eh_plan")
    return c.value


PACKED_SPLIT2, PACKED_SPLIT3, PACKED_BF16, PACKED_F16 = 2, 3, 4, 5


def num_packed_configs(kind: int) -> int:
    """Configurations of a packed product kind (split-BF16 2/3 terms, bf16 / half inputs)."""
    return lib().eh_num_packed_configs(kind)


def plan_packed(kind: int, opa: int, opb: int, m: int, n: int, k: int) -> int:
    c = C.c_int()
    check(lib().eh_plan_packed(kind, opa, opb, m, n, k, C.byref(c)), "eh_plan_packed")
    return c.value


def time_config(dtype: int, opa: int, opb: int, m: int, n: int, k: int, A, B, Cm, config: int, reps: int = 3) -> float:
    """Median ms of `reps` runs of a plain configuration on the caller's buffers (eh_time_config)."""
    ms = C.c_float()
    check(lib().eh_time_config(dtype, opa, opb, m, n, k, _ptr(A), _ptr(B), _ptr(Cm), config, reps, C.byref(ms)),
          "eh_time_config")
    return ms.value


def time_config_packed(kind: int, opa: int, opb: int, m: int, n: int, k: int, A, B, Cm, config: int,
                       reps: int = 3) -> float:
    """Median ms of `reps` runs of a packed configuration (eh_time_config_packed)."""
    ms = C.c_float()
    check(lib().eh_time_config_packed(kind, opa, opb, m, n, k, _ptr(A), _ptr(B), _ptr(Cm), config, reps,
                                      C.byref(ms)), "eh_time_config_packed")
    return ms.value


def tuner_count() -> int:
    return lib().eh_tuner_count()


def tuner_clear() -> None:
    check(lib().eh_tuner_clear(), "eh_tuner_clear")


def tuner_save(path: str) -> None:
    check(lib().eh_tuner_save(path.encode()), "eh_tuner_save")


def tuner_load(path: str) -> None:
    check(lib().eh_tuner_load(path.encode()), "eh_tuner_load")


def set_tune_mode(mode: int) -> None:
    check(lib().eh_set_tune_mode(mode), "eh_set_tune_mode")


def set_sgemm_precision(prec: int) -> None:
    check(lib().eh_set_sgemm_precision(prec), "eh_set_sgemm_precision")
