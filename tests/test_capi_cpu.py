"""C-ABI checks that need no GPU: the libraries load, export every declared
symbol, and reject invalid arguments before any device work is issued."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import eigen_amd as E
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header: str):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:eh_|[sd](?:gemm|gemmtr|syrk|syr2k|symm|trmm|trsm)_|sbgemm_|[cz]gemm_|[sdcz]gemv_|xerbla_)\w*)\s*\(", text)))


def _exported(path: str):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if len(line.split()) >= 3}


def test_libraries_load():
    assert E.lib().eh_version().decode().startswith("eigen_amd")
    E.blas()


def test_every_declared_symbol_is_exported():
    main = _exported(E.lib_path())
    blas = _exported(E.blas_path())
    for name in _declared("eigen_amd.h"):
        assert name in main, name
    for name in _declared("eigen_amd_blas.h"):
        assert name in blas, name
    assert len(_declared("eigen_amd.h")) >= 15


def test_config_table():
    assert E.num_configs() == 24
    seen = set()
    for c in range(24):
        tm, tn, sk = E.config_info(c)
        assert tm in (64, 128, 256) and tn in (64, 128, 256) and sk in (1, 2, 4, 8)
        seen.add((tm, tn, sk))
    assert len(seen) == 24
    with pytest.raises(E.EigenAmdError):
        E.config_info(24)


def test_heuristic_plan_is_valid():
    prev = E.lib().eh_get_tune_mode()
    E.set_tune_mode(E.TUNE_HEURISTIC)
    try:
        for (m, n, k) in [(8192, 8192, 8192), (16384, 256, 16384), (64, 64, 64), (4096, 4096, 4096), (1, 1, 1)]:
            c = E.plan(E.F64, E.OP_N, E.OP_N, m, n, k)
            assert 0 <= c < 24
        # a product whose tiles cannot fill the chip gets split K
        assert E.config_info(E.plan(E.F64, E.OP_N, E.OP_N, 2048, 256, 32768))[2] > 1
        assert E.config_info(E.plan(E.F64, E.OP_N, E.OP_N, 8192, 8192, 8192))[2] == 1
    finally:
        E.set_tune_mode(prev)


def test_packed_configs_and_rule():
    """Packed products (split-BF16, bf16 / half inputs) have their own
    configuration spaces; outside the table mode the planner uses the fixed
    rule: the large tile once it fills the chip (256 workgroups), else
    128x128, then split K while the grid stays under one wave and each split
    keeps 512 of depth."""
    assert E.num_packed_configs(E.PACKED_SPLIT2) == E.num_packed_configs(E.PACKED_SPLIT3) == 16
    assert E.num_packed_configs(E.PACKED_BF16) == E.num_packed_configs(E.PACKED_F16) == 8
    assert E.num_packed_configs(7) == 0
    info = lambda kind, c: tuple(_packed_info(kind, c))
    assert info(E.PACKED_SPLIT3, 0) == (128, 128, 1) and info(E.PACKED_SPLIT3, 8) == (256, 128, 1)
    assert info(E.PACKED_SPLIT2, 13) == (256, 256, 2)
    assert info(E.PACKED_BF16, 4) == (256, 256, 1) and info(E.PACKED_F16, 3) == (128, 128, 8)
    prev = E.lib().eh_get_tune_mode()
    E.set_tune_mode(E.TUNE_HEURISTIC)
    try:
        # 8192^2: 1024 tiles of 256x256 fill the chip -> tile 3, no split;
        # 4096 x 2048: 128 of 256x256 but 256 of 256x128 -> tile 2
        assert E.plan_packed(E.PACKED_SPLIT3, 0, 0, 8192, 8192, 8192) == 12
        assert E.plan_packed(E.PACKED_SPLIT3, 0, 0, 4096, 2048, 4096) == 8
        # 1024^2 x 8192: 64 tiles of 128x128 -> split 4 (64*4 = 256)
        assert E.plan_packed(E.PACKED_SPLIT2, 0, 0, 1024, 1024, 8192) == 2
        # bf16 8192^3: 1024 tiles of 256x256 -> tile 1
        assert E.plan_packed(E.PACKED_BF16, 0, 0, 8192, 8192, 8192) == 4
        # 4096^3 bf16: 256 tiles of 256x256 -> tile 1; 2048^3: 64 -> 128x128 (256 tiles), no split
        assert E.plan_packed(E.PACKED_F16, 0, 0, 4096, 4096, 4096) == 4
        assert E.plan_packed(E.PACKED_F16, 0, 0, 2048, 2048, 2048) == 0
    finally:
        E.set_tune_mode(prev)
    with pytest.raises(E.EigenAmdError):
        E.plan_packed(9, 0, 0, 8, 8, 8)


def _packed_info(kind, c):
    import ctypes as C
    tm, tn, sk = C.c_int(), C.c_int(), C.c_int()
    E.check(E.lib().eh_packed_config_info(kind, c, C.byref(tm), C.byref(tn), C.byref(sk)), "eh_packed_config_info")
    return tm.value, tn.value, sk.value


def test_compiled_in_selector_is_the_trained_network():
    """The planner's default network (compiled into the library) is the one
    trained on the device's timing table: its host-side choices equal the
    argmax of the FP64 restatement run on the saved parameters, and the saved
    training run beat the heuristic on its held-out shapes."""
    import json
    import os
    from oracle import mlp as M
    from eigen_amd.selector import load_db
    path = os.path.join(os.path.dirname(E.__file__), "data", "selector_gfx950.json")
    res = json.load(open(path))
    assert res["mlp_regret"]["mean"] < res["heuristic_regret"]["mean"]
    # the network chooses among the library's configurations as they were when
    # it was trained (a changed configuration table needs a retrained network)
    assert res["arch"] == "gfx950"
    assert [tuple(c) for c in res["configs"]] == [E.config_info(c) for c in range(E.num_configs())]
    P = np.array(res["params"], dtype=np.float64)
    assert E.lib().eh_get_tune_mode() == E.TUNE_MLP  # the default when a network is compiled in
    shapes, _ = load_db(os.path.join(os.path.dirname(E.__file__), "data", "tune_gfx950.csv"))
    shapes = shapes[4000:4200]
    X = np.array([M.features(*s) for s in shapes.tolist()])
    _, Z, _ = M.forward(P, X)
    host = np.array([E.plan(*s) for s in shapes.tolist()])
    top2 = np.sort(Z, axis=1)[:, -2:]
    clear = (top2[:, 1] - top2[:, 0]) > 1e-4  # FP32 host inference; skip near-ties
    assert clear.sum() > 150
    assert np.array_equal(host[clear], Z.argmax(axis=1)[clear])


def _call_gemm(dtype=E.F32, opa=E.OP_N, opb=E.OP_N, m=4, n=4, k=4, lda=4, ldb=4, ldc=4, alpha=True, beta=True,
               A=1, B=1, Cp=1):
    L = E.lib()
    a = C.c_float(1) if dtype == E.F32 else C.c_double(1)
    return L.eh_gemm(dtype, opa, opb, m, n, k, C.byref(a) if alpha else None, A, lda, B, ldb,
                     C.byref(a) if beta else None, Cp, ldc, 0, None)


@pytest.mark.parametrize("kw", [
    dict(dtype=7), dict(opa=5), dict(opb=-1), dict(m=-1), dict(n=-2), dict(k=-3),
    dict(lda=3), dict(opa=E.OP_T, m=8, lda=4, k=5), dict(ldb=3), dict(opb=E.OP_T, n=9, ldb=4), dict(ldc=3),
    dict(alpha=False), dict(beta=False), dict(Cp=None), dict(A=None), dict(B=None),
])
def test_invalid_arguments_rejected(kw):
    assert _call_gemm(**kw) == 1  # EH_ERR_INVALID, before any device work
    assert E.lib().eh_last_error().decode() != ""


def test_quick_returns_do_not_touch_device():
    # m == 0 or n == 0: nothing happens, even with NULL buffers
    assert _call_gemm(m=0, A=None, B=None, Cp=None, ldc=1, lda=1) == 0
    assert _call_gemm(n=0, A=None, B=None, Cp=None) == 0


def test_fp64_rejects_split_precision():
    L = E.lib()
    one = C.c_double(1)
    st = L.eh_gemm(E.F64, 0, 0, 4, 4, 4, C.byref(one), 1, 4, 1, 4, C.byref(one), 1, 4, E.PREC_SPLIT3, None)
    assert st == 3  # EH_ERR_UNSUPPORTED


_XERBLA_SCRIPT = r"""
import ctypes as C, sys
sys.path.insert(0, {root!r})
import eigen_amd as E
L = E.blas()
i = lambda v: C.byref(C.c_int(v))
one = C.c_float(1.0)
buf = (C.c_float * 4)()
args = {args!r}
ta, tb, m, n, k, lda, ldb, ldc = args
L.sgemm_(ta.encode(), tb.encode(), i(m), i(n), i(k), C.byref(one), buf, i(lda), buf, i(ldb), C.byref(one), buf, i(ldc))
sys.stdout.flush()
"""

INVALID = [a for a in [
    ("X", "N", 2, 2, 2, 2, 2, 2), ("N", "Y", 2, 2, 2, 2, 2, 2), ("N", "N", -1, 2, 2, 2, 2, 2),
    ("N", "N", 2, -1, 2, 2, 2, 2), ("N", "N", 2, 2, -1, 2, 2, 2), ("N", "N", 4, 2, 2, 3, 2, 4),
    ("T", "N", 4, 2, 3, 2, 3, 4), ("N", "N", 2, 2, 5, 2, 4, 2), ("N", "T", 2, 6, 2, 2, 5, 2),
    ("N", "N", 5, 2, 2, 5, 2, 4)]]


@pytest.mark.parametrize("args", INVALID)
def test_sgemm_error_exits_match_reference(args):
    r = subprocess.run([sys.executable, "-c", _XERBLA_SCRIPT.format(root=ROOT, args=args)],
                       capture_output=True, text=True, timeout=120)
    want = O.gemm_check(*args)
    assert want != 0
    assert f"ERROR #{want}: SGEMM" in r.stdout, (r.stdout, r.stderr)


def test_sgemm_zero_size_quick_return():
    # m == 0: returns before touching any pointer (NULLs are fine)
    L = E.blas()
    i = lambda v: C.byref(C.c_int(v))
    one = C.c_float(1.0)
    L.sgemm_(b"N", b"N", i(0), i(3), i(3), C.byref(one), None, i(1), None, i(3), C.byref(one), None, i(1))


def test_grouped_rejects_bad_product():
    # validation happens before any device work: product 0 has lda < m
    probs = [E.GemmProblem(dtype=0, opa=0, opb=0, m=4, n=4, k=4, lda=2, ldb=4, ldc=4, A=1, B=1, C=1, alpha=1.0,
                           beta=0.0)]
    arr = (E.GemmProblem * 1)(*probs)
    h = C.c_void_p()
    assert E.lib().eh_grouped_create(arr, 1, C.byref(h)) == 1
    assert "product 0" in E.lib().eh_last_error().decode()
    assert E.lib().eh_grouped_create(arr, 0, C.byref(h)) == 1


def test_tuned_shapes_use_their_measured_best():
    """Shapes tuned with the compiled-in network plan to their measured best
    configuration; a network installed by the caller is used as is."""
    import os
    from oracle import mlp as M
    from eigen_amd.selector import N_SAMPLE, load_db
    S, T, I = load_db(os.path.join(os.path.dirname(E.__file__), "data", "tune_gfx950.csv"), with_index=True)
    can = I >= N_SAMPLE
    assert can.sum() >= 60
    for s, t in zip(S[can].tolist(), T[can]):
        assert E.plan(*s) == int(t.argmin()), s
    P = np.random.default_rng(3).normal(0, 0.5, M.N_PARAMS)
    L = E.lib()
    E.check(L.eh_planner_set_mlp_params(np.ascontiguousarray(P, np.float32).ctypes.data_as(C.c_void_p)))
    try:
        for s in S[can][:8].tolist():
            _, Z, _ = M.forward(P.astype(np.float32).astype(np.float64), np.array([M.features(*s)]))
            top = np.sort(Z[0])[-2:]
            if top[1] - top[0] > 1e-4:
                assert E.plan(*s) == int(Z[0].argmax()), s
    finally:
        E.check(L.eh_planner_set_mlp_params(None))


def test_strided_destination_argument_checks():
    L = E.lib()
    one = C.c_float(1.0)
    args = lambda incc, ldc: (E.F32, 0, 0, 4, 3, 2, C.byref(one), 1, 4, 1, 2, C.byref(one), 1, incc, ldc, 0, None)
    assert L.eh_gemm_strided(*args(0, 4)) == E.ERR_INVALID          # incc < 1
    assert L.eh_gemm_strided(*args(2, 6)) == E.ERR_INVALID          # columns would overlap: ldc < (m-1)*incc + 1
    assert L.eh_gemm_strided(*args(1, 3)) == E.ERR_INVALID          # incc == 1 is eh_gemm: ldc >= m
    assert "ldc" in E.lib().eh_last_error().decode()


def test_level3_argument_checks():
    """The level-3 entry points reject bad flags and leading dimensions, and
    return at once for empty problems, before any device work."""
    L = E.lib()
    one = C.c_double(1.0)
    a = C.byref(one)
    # eh_trsm / eh_trmm(dtype, side, uplo, transa, diag, m, n, alpha, A, lda, B, ldb, stream)
    for f in (L.eh_trsm, L.eh_trmm):
        assert f(E.F64, 2, 0, 0, 0, 4, 3, a, 1, 4, 1, 4, None) == E.ERR_INVALID       # side
        assert f(E.F64, 0, 2, 0, 0, 4, 3, a, 1, 4, 1, 4, None) == E.ERR_INVALID       # uplo
        assert f(E.F64, 0, 0, 3, 0, 4, 3, a, 1, 4, 1, 4, None) == E.ERR_INVALID       # transa
        assert f(E.F64, 0, 0, 0, 2, 4, 3, a, 1, 4, 1, 4, None) == E.ERR_INVALID       # diag
        assert f(E.F64, 0, 0, 0, 0, 4, 3, a, 1, 3, 1, 4, None) == E.ERR_INVALID       # lda < m (left)
        assert f(E.F64, 1, 0, 0, 0, 4, 3, a, 1, 2, 1, 4, None) == E.ERR_INVALID       # lda < n (right)
        assert f(E.F64, 1, 0, 0, 0, 4, 3, a, 1, 3, 1, 3, None) == E.ERR_INVALID       # ldb < m
        assert f(E.F64, 0, 0, 0, 0, 0, 3, a, 1, 1, 1, 1, None) == E.SUCCESS           # m == 0
        assert f(E.F64, 0, 0, 0, 0, 4, 0, a, 1, 4, 1, 4, None) == E.SUCCESS           # n == 0
    # eh_symm(dtype, side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc, stream)
    assert L.eh_symm(E.F64, 0, 0, 4, 3, a, 1, 3, 1, 4, a, 1, 4, None) == E.ERR_INVALID   # lda < m
    assert L.eh_symm(E.F64, 1, 0, 4, 3, a, 1, 3, 1, 3, a, 1, 4, None) == E.ERR_INVALID   # ldb < m
    assert L.eh_symm(E.F64, 1, 0, 4, 3, a, 1, 3, 1, 4, a, 1, 3, None) == E.ERR_INVALID   # ldc < m
    assert L.eh_symm(E.F64, 0, 0, 0, 3, a, 1, 1, 1, 1, a, 1, 1, None) == E.SUCCESS
    # eh_syr2k(dtype, uplo, trans, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream)
    assert L.eh_syr2k(E.F32, 0, 0, 4, 3, a, 1, 3, 1, 4, a, 1, 4, None) == E.ERR_INVALID  # lda < n
    assert L.eh_syr2k(E.F32, 0, 1, 4, 3, a, 1, 3, 1, 2, a, 1, 4, None) == E.ERR_INVALID  # ldb < k
    assert L.eh_syr2k(E.F32, 0, 0, 0, 3, a, 1, 1, 1, 1, a, 1, 1, None) == E.SUCCESS
    # eh_gemmtr(dtype, uplo, opa, opb, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream)
    assert L.eh_gemmtr(E.F32, 5, 0, 0, 4, 3, a, 1, 4, 1, 3, a, 1, 4, None) == E.ERR_INVALID  # uplo
    assert L.eh_gemmtr(E.F32, 0, 0, 0, 4, 3, a, 1, 4, 1, 2, a, 1, 4, None) == E.ERR_INVALID  # ldb < k
    assert L.eh_gemmtr(E.F32, 0, 0, 0, 0, 3, a, 1, 1, 1, 3, a, 1, 1, None) == E.SUCCESS


_GEMV_SCRIPT = r"""
import ctypes as C, sys
sys.path.insert(0, {root!r})
import eigen_amd as E
L = E.blas()
i = lambda v: C.byref(C.c_int(v))
one = C.c_double(1.0)
buf = (C.c_double * 64)()
t, m, n, lda, incx, incy = {args!r}
L.dgemv_(t.encode(), i(m), i(n), C.byref(one), buf, i(lda), buf, i(incx), C.byref(one), buf, i(incy))
sys.stdout.flush()
"""

GEMV_INVALID = [("X", 2, 2, 2, 1, 1), ("N", -1, 2, 2, 1, 1), ("T", 2, -3, 2, 1, 1), ("N", 4, 2, 3, 1, 1),
                ("T", 3, 2, 3, 0, 1), ("N", 3, 2, 3, 1, 0), ("C", 0, 2, 0, 1, 1)]


@pytest.mark.parametrize("args", GEMV_INVALID)
def test_dgemv_error_exits_match_eigen_blas(args):
    """Info codes of dgemv_ against Eigen's own dgemv_ (blas/level2_impl.h:45-59,
    compiled into oracle/_ref with a recording xerbla_)."""
    if not O.ref_available():
        pytest.skip("oracle/_ref absent")
    R = O.blas_ref()
    R.ref_blas_reset()
    t, m, n, lda, incx, incy = args
    i = lambda v: C.byref(C.c_int(v))
    one = C.c_double(1.0)
    buf = (C.c_double * 64)()
    R.dgemv_(t.encode(), i(m), i(n), C.byref(one), buf, i(lda), buf, i(incx), C.byref(one), buf, i(incy))
    want = R.ref_blas_last_info()
    assert want != 0
    r = subprocess.run([sys.executable, "-c", _GEMV_SCRIPT.format(root=ROOT, args=args)], capture_output=True,
                       text=True, timeout=120)
    assert f"ERROR #{want}: DGEMV" in r.stdout, (want, r.stdout, r.stderr)


def test_gemv_argument_checks():
    L = E.lib()
    one = C.c_double(1.0)
    a = C.byref(one)
    # eh_gemv(dtype, trans, m, n, alpha, A, lda, x, incx, beta, y, incy, stream)
    assert L.eh_gemv(E.F64, 3, 4, 3, a, 1, 4, 1, 1, a, 1, 1, None) == E.ERR_INVALID   # trans
    assert L.eh_gemv(E.F64, 0, 4, 3, a, 1, 3, 1, 1, a, 1, 1, None) == E.ERR_INVALID   # lda < m
    assert L.eh_gemv(E.F64, 0, 4, 3, a, 1, 4, 1, 0, a, 1, 1, None) == E.ERR_INVALID   # incx == 0
    assert L.eh_gemv(E.F64, 0, 4, 3, a, 1, 4, 1, 1, a, 1, 0, None) == E.ERR_INVALID   # incy == 0
    assert L.eh_gemv(E.F64, 0, 0, 3, a, None, 1, None, 1, a, None, 1, None) == E.SUCCESS  # m == 0


def test_triangular_tile_enumeration(tmp_path):
    """The compact grid of triangular updates (tri_tile_coords) covers every
    tile on or below the diagonal exactly once, for all grids up to 300 x 300
    tiles (host build of the device function)."""
    exe = tmp_path / "tri_tiles_check"
    src = os.path.join(ROOT, "tests", "tri_tiles_check.cpp")
    r = subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-o", str(exe), src], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and "ALL PASSED" in out.stdout, out.stdout


def test_batched_strided_argument_checks():
    L = E.lib()
    one = C.c_double(1.0)
    a = C.byref(one)
    # (dtype, opa, opb, m, n, k, alpha, A, lda, sa, B, ldb, sb, beta, C, ldc, sc, count, stream)
    assert L.eh_gemm_batched_strided(E.F64, 0, 0, 4, 3, 2, a, 1, 4, 8, 1, 2, 6, a, 1, 4, 11, 2, None) == E.ERR_INVALID
    assert L.eh_gemm_batched_strided(E.F64, 0, 0, 4, 3, 2, a, 1, 4, -1, 1, 2, 6, a, 1, 4, 12, 2, None) == E.ERR_INVALID
    assert L.eh_gemm_batched_strided(E.F64, 0, 0, 4, 3, 2, a, 1, 3, 8, 1, 2, 6, a, 1, 4, 12, 2, None) == E.ERR_INVALID
    assert L.eh_gemm_batched_strided(E.F64, 0, 0, 4, 3, 2, a, 1, 4, 8, 1, 2, 6, a, 1, 4, 12, 0, None) == E.SUCCESS
