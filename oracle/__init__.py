"""CPU oracle for the dense matrix-product path -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the benchmark's CPU-baseline leg import
this package, and only as the checker (or, for the baseline, as the timed
Eigen CPU GEBP).  The eigen_amd product never imports it.

Two layers:
  * ``lib()``  -- the C restatement in oracle.c (always built by ``make``).
  * ``ref()``  -- Eigen itself, compiled from the reference tree into
    oracle/_ref by ``make ref`` (present only where that build ran; the .so
    files travel with the tree).  ``blas_ref()`` is Eigen's BLAS library.

The helpers below wrap both with numpy arrays (column-major, Fortran order).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_TREE = "/root/reference"

UNIFORM, NORMAL, WIDE, INT = 0, 1, 2, 3

_lib = None
_ref = None
_blas = None


def build(force: bool = False, with_ref: bool | None = None) -> None:
    """Compile the C restatement (and the Eigen reference when the tree exists)."""
    args = ["make", "-C", HERE]
    if force:
        args.append("-B")
    subprocess.run(args + ["all"], check=True, stdout=subprocess.DEVNULL)
    if with_ref is None:
        with_ref = os.path.isdir(os.path.join(REF_TREE, "Eigen"))
    if with_ref:
        subprocess.run(["make", "-C", HERE, "-j8", "ref"], check=True, stdout=subprocess.DEVNULL)
        if os.path.exists(os.path.join(HERE, "..", "eigen_amd", "_lib", "libeigen_amd_blas.so")):
            subprocess.run(["make", "-C", HERE, "-B", "netlib"], check=True, stdout=subprocess.DEVNULL)


def _cpu_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    fl = line.split()
                    return all(x in fl for x in ("avx512f", "avx512bw", "avx512dq", "avx512vl", "avx512cd"))
    except OSError:
        pass
    return False


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "_build", "liboracle.so")
        if not os.path.exists(path):
            build(with_ref=False)
        L = C.CDLL(path)
        vp, i64, u64, f64, i32 = C.c_void_p, C.c_int64, C.c_uint64, C.c_double, C.c_int
        L.orc_splitmix64_at.restype = u64
        L.orc_splitmix64_at.argtypes = [u64, u64]
        L.orc_fill_f32.argtypes = [vp, i64, u64, i32, f64, f64]
        L.orc_fill_f64.argtypes = [vp, i64, u64, i32, f64, f64]
        L.orc_gemm_check.restype = i32
        L.orc_gemm_check.argtypes = [C.c_char, C.c_char] + [i32] * 6
        L.orc_gemm_gebp_f32.restype = i32
        L.orc_gemm_gebp_f32.argtypes = [C.c_char, C.c_char, i32, i32, i32, C.c_float, vp, i32, vp, i32,
                                        C.c_float, vp, i32, i32, i32, i32]
        L.orc_gemm_gebp_f64.restype = i32
        L.orc_gemm_gebp_f64.argtypes = [C.c_char, C.c_char, i32, i32, i32, f64, vp, i32, vp, i32,
                                        f64, vp, i32, i32, i32, i32]
        L.orc_gemm_ref.argtypes = [i32, i32, i32, i64, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64,
                                   vp, i64, vp, vp]
        L.orc_bf16_rtne.restype = C.c_uint16
        L.orc_bf16_rtne.argtypes = [C.c_float]
        L.orc_bf16_to_f32.restype = C.c_float
        L.orc_bf16_to_f32.argtypes = [C.c_uint16]
        L.orc_bf16_split.argtypes = [vp, i64, i32, vp, vp, vp]
        L.orc_bf16_rtne_n.argtypes = [vp, i64, vp]
        L.orc_gemm_split_ref.argtypes = [i32, i32, i32, i64, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64,
                                         vp, i64, vp]
        L.orc_gemm_split_pairs_ref.argtypes = [i32, i32, i32, i32, i64, i64, i64, f64, vp, i64, vp, i64, f64, vp,
                                               i64, vp, i64, vp]
        L.orc_half_rtne_n.argtypes = [vp, i64, vp]
        L.orc_half_to_f32_n.argtypes = [vp, i64, vp]
        L.orc_split_kept_pairs.restype = i32
        L.orc_split_kept_pairs.argtypes = [i32]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libeigen_ref_avx2.so"))


def ref():
    """Eigen's CPU product path compiled from the reference tree (AVX-512 build where the host has it)."""
    global _ref
    if _ref is None:
        name = "libeigen_ref_avx512.so" if _cpu_has_avx512() else "libeigen_ref_avx2.so"
        path = os.path.join(HERE, "_ref", name)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref` where the Eigen tree exists")
        L = C.CDLL(path)
        vp, lg, i32, i64 = C.c_void_p, C.c_long, C.c_int, C.c_int64
        for suf, T in (("f32", C.c_float), ("f64", C.c_double)):
            f = getattr(L, "ref_gemm_" + suf)
            f.argtypes = [i32, i32, lg, lg, lg, T, vp, lg, vp, lg, vp, lg]
            g = getattr(L, "ref_gemm_beta_" + suf)
            g.argtypes = [i32, i32, lg, lg, lg, T, vp, lg, vp, lg, T, vp, lg]
        L.ref_grouped.argtypes = [i64, vp, vp, vp, vp, vp, vp, vp, i32]
        L.ref_bf16_rtne.restype = C.c_uint16
        L.ref_bf16_rtne.argtypes = [C.c_float]
        L.ref_bf16_rtne_n.argtypes = [vp, i64, vp]
        L.ref_half_rtne_n.argtypes = [vp, i64, vp]
        L.ref_gemm_half.argtypes = [lg, lg, lg, vp, vp, vp]
        L.ref_max_threads.restype = i32
        L.ref_get_threads.restype = i32
        L.ref_set_threads.argtypes = [i32]
        L.ref_simd.restype = C.c_char_p
        L.ref_blocking_f32.argtypes = [lg, lg, lg, vp, vp, vp]
        _ref = L
    return _ref


def blas_ref():
    """Eigen's own BLAS library (sgemm_/dgemm_) with a recording xerbla_."""
    global _blas
    if _blas is None:
        path = os.path.join(HERE, "_ref", "libeigen_blas_ref.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        L.ref_blas_last_info.restype = C.c_int
        L.ref_blas_last_name.restype = C.c_char_p
        _blas = L
    return _blas


# --------------------------------------------------------------------------
# numpy helpers
# --------------------------------------------------------------------------
def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def fill(shape, dtype, seed: int, dist: int = UNIFORM, p0: float = -1.0, p1: float = 1.0) -> np.ndarray:
    """Column-major array from the seeded splitmix64 stream (element i = column-major index i)."""
    a = np.empty(shape, dtype=dtype, order="F")
    n = a.size
    flat = a.reshape(-1, order="F")
    f = lib().orc_fill_f32 if a.dtype == np.float32 else lib().orc_fill_f64
    f(_ptr(flat), n, seed, dist, p0, p1)
    return a


def gemm_check(ta: str, tb: str, m, n, k, lda, ldb, ldc) -> int:
    return lib().orc_gemm_check(ta.encode(), tb.encode(), m, n, k, lda, ldb, ldc)


def gemm_gebp(ta: str, tb: str, m, n, k, alpha, A, lda, B, ldb, beta, C_, ldc, kc=0, mc=0, nc=0) -> int:
    """C restatement of Eigen's BLAS gemm (in place on C_, column-major numpy array)."""
    f = lib().orc_gemm_gebp_f32 if C_.dtype == np.float32 else lib().orc_gemm_gebp_f64
    return f(ta.encode(), tb.encode(), m, n, k, alpha, _ptr(A), lda, _ptr(B), ldb, beta, _ptr(C_), ldc,
             kc, mc, nc)


def gemm_ref(ta: bool, tb: bool, alpha, A, B, beta=0.0, C0=None, rows=None, m=None, n=None, k=None,
             lda=None, ldb=None, ldc=None, with_abs=False):
    """Long-double reference of alpha*op(A)*op(B) + beta*C0 on the selected rows.

    A, B, C0 are column-major arrays; returns an (nrows, n) float64 array (and
    the |alpha| sum |a||b| array when with_abs)."""
    dtype = 0 if A.dtype == np.float32 else 1
    # the leading dimension defaults to shape[0]: the arrays must then be
    # column-major (a row-major array would be read transposed)
    for name, x, ld in (("A", A, lda), ("B", B, ldb), ("C0", C0, ldc)):
        if x is not None and ld is None and x.ndim == 2 and not x.flags.f_contiguous:
            raise ValueError(f"gemm_ref: {name} must be column-major (Fortran order)")
    if m is None:
        m = A.shape[1] if ta else A.shape[0]
    if k is None:
        k = A.shape[0] if ta else A.shape[1]
    if n is None:
        n = B.shape[0] if tb else B.shape[1]
    lda = lda or A.shape[0]
    ldb = ldb or B.shape[0]
    if C0 is None:
        C0 = np.zeros((1, 1), dtype=A.dtype, order="F")
        beta = 0.0
    ldc = ldc or C0.shape[0]
    if rows is None:
        rows_arr = None
        nrows = m
    else:
        rows_arr = np.ascontiguousarray(rows, dtype=np.int64)
        nrows = rows_arr.size
    out = np.empty((nrows, n), dtype=np.float64)
    absout = np.empty((nrows, n), dtype=np.float64) if with_abs else None
    lib().orc_gemm_ref(dtype, int(ta), int(tb), m, n, k, float(alpha), _ptr(A), lda, _ptr(B), ldb, float(beta),
                       _ptr(C0), ldc, None if rows_arr is None else _ptr(rows_arr), nrows, _ptr(out),
                       None if absout is None else _ptr(absout))
    return (out, absout) if with_abs else out


def bf16_rtne(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    t0 = np.empty(x.shape, np.uint16)
    lib().orc_bf16_rtne_n(_ptr(x), x.size, _ptr(t0))
    return t0


def half_rtne(x: np.ndarray) -> np.ndarray:
    """IEEE binary16 bits of float32 values, round to nearest even."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, np.uint16)
    lib().orc_half_rtne_n(_ptr(x), x.size, _ptr(out))
    return out


def half_to_f32(h: np.ndarray) -> np.ndarray:
    h = np.ascontiguousarray(h, dtype=np.uint16)
    out = np.empty(h.shape, np.float32)
    lib().orc_half_to_f32_n(_ptr(h), h.size, _ptr(out))
    return out


def bf16_split(x: np.ndarray, terms: int):
    x = np.ascontiguousarray(x, dtype=np.float32)
    ts = [np.zeros(x.shape, np.uint16) for _ in range(3)]
    lib().orc_bf16_split(_ptr(x), x.size, terms, _ptr(ts[0]), _ptr(ts[1]), _ptr(ts[2]))
    return ts[:terms]


def bf16_to_f32(h: np.ndarray) -> np.ndarray:
    return (np.asarray(h, dtype=np.uint32) << 16).view(np.float32)


def split_kept_pairs(terms: int) -> int:
    """Bit 3p+q set: the split product keeps t_p(A) * t_q(B) (p + q < terms)."""
    return int(lib().orc_split_kept_pairs(terms))


def gemm_split_ref(terms: int, ta: bool, tb: bool, alpha, A, B, beta=0.0, C0=None, rows=None, pairs=None):
    """Long-double value of the split product alpha*sum_{(p,q) kept} t_p(op(A)) t_q(op(B)) + beta*C0.

    `pairs` (9-bit mask, bit 3p+q) overrides the kept set of the split with
    `terms` terms (default: p + q < terms)."""
    m = A.shape[1] if ta else A.shape[0]
    k = A.shape[0] if ta else A.shape[1]
    n = B.shape[0] if tb else B.shape[1]
    if C0 is None:
        C0 = np.zeros((1, 1), np.float32, order="F")
        beta = 0.0
    rows_arr = None if rows is None else np.ascontiguousarray(rows, dtype=np.int64)
    nrows = m if rows is None else rows_arr.size
    out = np.empty((nrows, n), np.float64)
    mask = split_kept_pairs(terms) if pairs is None else int(pairs)
    lib().orc_gemm_split_pairs_ref(terms, mask, int(ta), int(tb), m, n, k, float(alpha), _ptr(A), A.shape[0],
                                   _ptr(B), B.shape[0], float(beta), _ptr(C0), C0.shape[0],
                                   None if rows_arr is None else _ptr(rows_arr), nrows, _ptr(out))
    return out


def eigen_gemm(ta: bool, tb: bool, alpha, A, B, C_, threads: int | None = None):
    """Eigen's own C.noalias() += alpha*op(A)*op(B) on host Maps (in place on C_)."""
    L = ref()
    if threads is not None:
        L.ref_set_threads(threads)
    m, n = C_.shape
    k = A.shape[0] if ta else A.shape[1]
    f = L.ref_gemm_f32 if C_.dtype == np.float32 else L.ref_gemm_f64
    f(int(ta), int(tb), m, n, k, alpha, _ptr(A), A.shape[0], _ptr(B), B.shape[0], _ptr(C_), C_.shape[0])


def normwise_error(Cgpu: np.ndarray, Cref: np.ndarray, k: int, u: float, normA: float, normB: float) -> float:
    """||C - C_ref||_F / (K * u * ||A||_F * ||B||_F)."""
    diff = np.linalg.norm(np.asarray(Cgpu, np.float64) - np.asarray(Cref, np.float64))
    den = k * u * normA * normB
    return float(diff / den) if den > 0 else float(diff)


def higham_mary_bound(absprod: np.ndarray, k: int, u: float, lam: float = 5.0, nprod: int = 2) -> float:
    """lambda * sqrt(k) * u * nprod * || |A||B| ||_F  (test/product_test_helpers.h:56-74 form)."""
    return float(lam * np.sqrt(k) * u * nprod * np.linalg.norm(absprod))


# --------------------------------------------------------------------------
# Other level-3 routines, real and complex (restated in numpy: float64 for
# real inputs, extended-precision complex for complex ones; tests compare the
# device with normwise bounds, not bit for bit).
# --------------------------------------------------------------------------
def _w(x) -> np.ndarray:
    """Working copy: float64 for real arrays, clongdouble for complex ones."""
    x = np.asarray(x)
    return np.array(x, np.clongdouble if np.iscomplexobj(x) else np.float64)


def _tri_dense(A: np.ndarray, n: int, lower: bool, unit: bool) -> np.ndarray:
    """The n x n triangle of A (column-major) as a dense matrix; the other
    triangle is never read, unit puts ones on the diagonal."""
    T = _w(A[:n, :n])
    T = np.tril(T) if lower else np.triu(T)
    if unit:
        np.fill_diagonal(T, 1.0)
    return T


def _sym_dense(A: np.ndarray, n: int, lower: bool) -> np.ndarray:
    """Symmetric matrix from the stored triangle (blas/level3_impl.h:550-557
    form: the mirror is the plain transpose, also for complex scalars)."""
    T = _tri_dense(A, n, lower, False)
    return T + np.tril(T, -1).T if lower else T + np.triu(T, 1).T


def _herm_dense(A: np.ndarray, n: int, lower: bool) -> np.ndarray:
    """Hermitian matrix from the stored triangle (?hemm_, blas/level3_impl.h:760-833
    through selfadjointView): the mirror is the conjugate transpose and only the
    real part of the diagonal is read."""
    T = _tri_dense(A, n, lower, False)
    np.fill_diagonal(T, np.real(np.diag(T)))
    return T + np.tril(T, -1).conj().T if lower else T + np.triu(T, 1).conj().T


def _op(x: np.ndarray, trans: str) -> np.ndarray:
    t = trans.upper()
    return x if t == "N" else (x.T if t == "T" else x.conj().T)


def _tri_mask(n: int, uplo: str) -> np.ndarray:
    return np.tril(np.ones((n, n), bool)) if uplo.upper() == "L" else np.triu(np.ones((n, n), bool))


def gemmtr_ref(uplo: str, ta: str, tb: str, alpha, A, B, beta, C0) -> np.ndarray:
    """?gemmtr_ (blas/level3_impl.h:198-227): beta-scale the uplo triangle
    (beta == 0 sets it to zero), then add alpha*op(A)*op(B) to it; the other
    triangle of C is returned unchanged."""
    C = _w(C0)
    n = C.shape[0]
    P = alpha * (_op(_w(A), ta) @ _op(_w(B), tb))
    mask = _tri_mask(n, uplo)
    scaled = np.zeros_like(C) if beta == 0 else beta * C
    C[mask] = (scaled + P)[mask]
    return C


def syrk_ref(uplo: str, trans: str, alpha, A, beta, C0) -> np.ndarray:
    """?syrk_ (blas/level3_impl.h:651-689): beta-scale the uplo triangle, then add
    alpha*A*A^T (trans N) or alpha*A^T*A (trans T; also C for real scalars)."""
    A_ = _w(A)
    P = A_ @ A_.T if trans.upper() == "N" else A_.T @ A_
    C = _w(C0)
    mask = _tri_mask(C.shape[0], uplo)
    scaled = np.zeros_like(C) if beta == 0 else beta * C
    C[mask] = (scaled + alpha * P)[mask]
    return C


def syr2k_ref(uplo: str, trans: str, alpha, A, B, beta, C0) -> np.ndarray:
    """?syr2k_ (blas/level3_impl.h:718-757): beta-scale the uplo triangle, then
    add alpha*(A*B^T + B*A^T) (trans N) or alpha*(A^T*B + B^T*A) (trans T;
    also C for real scalars)."""
    A_, B_ = _w(A), _w(B)
    if trans.upper() == "N":
        P = A_ @ B_.T + B_ @ A_.T
    else:
        P = A_.T @ B_ + B_.T @ A_
    C = _w(C0)
    mask = _tri_mask(C.shape[0], uplo)
    scaled = np.zeros_like(C) if beta == 0 else beta * C
    C[mask] = (scaled + alpha * P)[mask]
    return C


def herk_ref(uplo: str, trans: str, alpha: float, A, beta: float, C0) -> np.ndarray:
    """?herk_ (blas/level3_impl.h:886-911), alpha and beta real: when beta != 1
    the uplo triangle is scaled (beta == 0 sets it to zero) and the diagonal's
    imaginary part cleared; when k > 0 and alpha != 0, alpha*A*A^H (trans N) or
    alpha*A^H*A (trans C) is added and the diagonal's imaginary part cleared."""
    A_ = _w(A)
    C = _w(C0)
    n = C.shape[0]
    k = A_.shape[1] if trans.upper() == "N" else A_.shape[0]
    mask = _tri_mask(n, uplo)
    d = np.arange(n)
    if beta != 1:
        C[mask] = 0 if beta == 0 else (beta * C)[mask]
        C[d, d] = C[d, d].real
    if k > 0 and alpha != 0:
        P = A_ @ A_.conj().T if trans.upper() == "N" else A_.conj().T @ A_
        C[mask] = (C + alpha * P)[mask]
        C[d, d] = C[d, d].real
    return C


def her2k_ref(uplo: str, trans: str, alpha, A, B, beta: float, C0) -> np.ndarray:
    """?her2k_ (blas/level3_impl.h:944-985), beta real: when beta != 1 the uplo
    triangle is scaled (zero for beta == 0) and the diagonal's imaginary part
    cleared, else (k > 0, alpha != 0) only the latter; then, for k > 0, the
    triangle gets alpha*A*B^H + conj(alpha)*B*A^H (trans N) or
    alpha*A^H*B + conj(alpha)*B^H*A (trans C)."""
    A_, B_ = _w(A), _w(B)
    C = _w(C0)
    n = C.shape[0]
    k = A_.shape[1] if trans.upper() == "N" else A_.shape[0]
    mask = _tri_mask(n, uplo)
    d = np.arange(n)
    if beta != 1:
        C[mask] = 0 if beta == 0 else (beta * C)[mask]
        C[d, d] = C[d, d].real
    elif k > 0 and alpha != 0:
        C[d, d] = C[d, d].real
    if k == 0:
        return C
    ca = np.conj(alpha)
    if trans.upper() == "N":
        P = alpha * (A_ @ B_.conj().T) + ca * (B_ @ A_.conj().T)
    else:
        P = alpha * (A_.conj().T @ B_) + ca * (B_.conj().T @ A_)
    C[mask] = (C + P)[mask]
    return C


def symm_ref(side: str, uplo: str, alpha, A, B, beta, C0) -> np.ndarray:
    """?symm_ (blas/level3_impl.h:534-591): C = beta*C (beta == 0 sets zero),
    then C += alpha*S*B (side L) or alpha*B*S (side R), S the symmetric matrix
    of A's uplo triangle."""
    m, n = C0.shape
    ka = m if side.upper() == "L" else n
    S = _sym_dense(A, ka, uplo.upper() == "L")
    B_ = _w(B)[:m, :n]
    C = np.zeros_like(_w(C0)) if beta == 0 else beta * _w(C0)
    return C + alpha * (S @ B_ if side.upper() == "L" else B_ @ S)


def hemm_ref(side: str, uplo: str, alpha, A, B, beta, C0) -> np.ndarray:
    """?hemm_ (blas/level3_impl.h:760-833): as symm_ref with the Hermitian
    matrix of A's uplo triangle."""
    m, n = C0.shape
    ka = m if side.upper() == "L" else n
    H = _herm_dense(A, ka, uplo.upper() == "L")
    B_ = _w(B)[:m, :n]
    C = np.zeros_like(_w(C0)) if beta == 0 else beta * _w(C0)
    return C + alpha * (H @ B_ if side.upper() == "L" else B_ @ H)


def trmm_ref(side: str, uplo: str, trans: str, diag: str, alpha, A, B0) -> np.ndarray:
    """?trmm_ (blas/level3_impl.h:482-502): B = alpha*op(T)*B (side L) or
    alpha*B*op(T) (side R), T the uplo triangle of A (unit diagonal when diag
    is U); op C is the conjugate transpose."""
    m, n = B0.shape
    ka = m if side.upper() == "L" else n
    T = _op(_tri_dense(A, ka, uplo.upper() == "L", diag.upper() == "U"), trans)
    B_ = _w(B0)
    return alpha * (T @ B_ if side.upper() == "L" else B_ @ T)


def trsm_ref(side: str, uplo: str, trans: str, diag: str, alpha, A, B0) -> np.ndarray:
    """?trsm_ (blas/level3_impl.h:339-364): X with op(T)*X = B (side L) or
    X*op(T) = B (side R), then scaled by alpha (the reference's order); T the
    uplo triangle of A.  Substitution (float64, or extended complex), the
    textbook recurrence over the rows of X (side R: over the rows of X^T,
    solving op(T)^T X^T = B^T)."""
    m, n = B0.shape
    left = side.upper() == "L"
    ka = m if left else n
    T = _op(_tri_dense(A, ka, uplo.upper() == "L", diag.upper() == "U"), trans)
    lower = (uplo.upper() == "L") != (trans.upper() in "TC")  # triangle of op(T)
    Bt = _w(B0)
    if not left:
        T, Bt, lower = T.T, Bt.T.copy(), not lower
    X = np.zeros_like(Bt)
    for i in (range(ka) if lower else range(ka - 1, -1, -1)):
        acc = Bt[i].copy()
        acc -= T[i, :i] @ X[:i] if lower else T[i, i + 1:] @ X[i + 1:]
        X[i] = acc / T[i, i]
    return alpha * (X if left else X.T)


def cgemm_ref(ta: str, tb: str, alpha, A, B, beta, C0) -> np.ndarray:
    """Complex ?gemm_ (blas/level3_impl.h:12-96 for complex scalars): C = beta*C
    (beta == 0 sets zero, C not read) + alpha*op(A)*op(B), op in N / T / C
    (conjugate transpose).  Evaluated in extended precision (clongdouble)."""
    def op(x, t):
        x = np.asarray(x, np.clongdouble)
        return x if t.upper() == "N" else (x.T if t.upper() == "T" else x.conj().T)
    P = op(A, ta) @ op(B, tb)
    base = np.zeros(P.shape, np.clongdouble) if beta == 0 else complex(beta) * np.asarray(C0, np.clongdouble)
    return base + complex(alpha) * P
