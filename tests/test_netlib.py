"""Netlib level-3 known-answer testers (blas/testing/sblat3.f, dblat3.f,
cblat3.f, zblat3.f of the reference, compiled into oracle/_ref and linked
against the library's BLAS), every routine of every precision switched on.

The input deck mirrors the reference's sblat3.dat / dblat3.dat (N in
{0,1,2,3,5,9}, alpha in {0,1,0.7}, beta in {0,1,1.3}, threshold 16) with the
routines the library provides switched on: all six real level-3 routines
(GEMM, SYMM, TRMM, TRSM, SYRK, SYR2K).  The tester supplies its own XERBLA, which
overrides our weak xerbla_, to check the error exits; the computational tests
compare every call with its own column-by-column recomputation (DMMCH).
The error-exit run uses N = 0 only and needs no GPU."""
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")

# routines in the order of the reference decks (sblat3.dat / dblat3.dat) and
# whether the library provides them
ROUTINES = [("GEMM ", True), ("SYMM ", True), ("TRMM ", True), ("TRSM ", True), ("SYRK ", True),
            ("SYR2K", True)]
NAMES = ["GEMM", "SYMM", "TRMM", "TRSM", "SYRK", "SYR2K"]


# complex decks (cblat3.dat / zblat3.dat): all nine complex routines
ROUTINES_COMPLEX = [("GEMM ", True), ("HEMM ", True), ("SYMM ", True), ("TRMM ", True), ("TRSM ", True),
                    ("HERK ", True), ("SYRK ", True), ("HER2K", True), ("SYR2K", True)]
NAMES_COMPLEX = ["GEMM", "HEMM", "SYMM", "TRMM", "TRSM", "HERK", "SYRK", "HER2K", "SYR2K"]


def deck(prec: str, ns, error_exits=True) -> str:
    up = prec.upper()
    cplx = prec in "cz"
    lines = [f"'{prec}blat3.summ'     NAME OF SUMMARY OUTPUT FILE", "6                 UNIT NUMBER OF SUMMARY FILE",
             f"'{prec}blat3.snap'     NAME OF SNAPSHOT OUTPUT FILE", "-1                UNIT NUMBER OF SNAPSHOT FILE",
             "F        LOGICAL FLAG, T TO REWIND SNAPSHOT FILE AFTER EACH RECORD.",
             "F        LOGICAL FLAG, T TO STOP ON FAILURES.",
             f"{'T' if error_exits else 'F'}        LOGICAL FLAG, T TO TEST ERROR EXITS.",
             "16.0     THRESHOLD VALUE OF TEST RATIO",
             f"{len(ns)}                 NUMBER OF VALUES OF N", " ".join(str(n) for n in ns) + "       VALUES OF N",
             "3                 NUMBER OF VALUES OF ALPHA",
             "(0.0,0.0) (1.0,0.0) (0.7,-0.9)       VALUES OF ALPHA" if cplx else "0.0 1.0 0.7       VALUES OF ALPHA",
             "3                 NUMBER OF VALUES OF BETA",
             "(0.0,0.0) (1.0,0.0) (1.3,-1.1)       VALUES OF BETA" if cplx else "0.0 1.0 1.3       VALUES OF BETA"]
    lines += [f"{up}{name} {'T' if on else 'F'} PUT F FOR NO TEST. SAME COLUMNS."
              for name, on in (ROUTINES_COMPLEX if cplx else ROUTINES)]
    return "\n".join(lines) + "\n"


def run_tester(prec: str, ns, error_exits=True) -> str:
    exe = os.path.join(REF, f"{prec}blat3")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built (needs the reference tree and amdflang at build time)")
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([exe], input=deck(prec, ns, error_exits), capture_output=True, text=True, cwd=d,
                           timeout=600)
        summ = os.path.join(d, f"{prec}blat3.summ")  # the tester writes its report here
        out = (open(summ).read() if os.path.exists(summ) else "") + r.stdout + r.stderr
    assert r.returncode == 0, out
    return out


@pytest.mark.parametrize("prec", ["s", "d"])
def test_error_exits(prec):
    out = run_tester(prec, [0])
    for name in NAMES:
        assert re.search(rf"{prec.upper()}{name}\s+PASSED THE TESTS OF ERROR-EXITS", out), out
        # N = 0: quick returns only
        assert re.search(rf"{prec.upper()}{name}\s+PASSED THE COMPUTATIONAL TESTS \(\s*\d+ CALLS\)", out), out
    assert f"{prec.upper()}GEMM  PASSED THE COMPUTATIONAL TESTS (    81 CALLS)" in out, out
    assert "FAILED" not in out and "SUSPECT" not in out, out


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["s", "d"])
def test_computational(gpu, prec):
    out = run_tester(prec, [0, 1, 2, 3, 5, 9])
    for name in NAMES:
        assert re.search(rf"{prec.upper()}{name}\s+PASSED THE TESTS OF ERROR-EXITS", out), out
        m = re.search(rf"{prec.upper()}{name}\s+PASSED THE COMPUTATIONAL TESTS \(\s*(\d+) CALLS\)", out)
        assert m and int(m.group(1)) > 1000, (name, out)
    assert f"{prec.upper()}GEMM  PASSED THE COMPUTATIONAL TESTS ( 17496 CALLS)" in out, out
    assert "FAILED" not in out and "SUSPECT" not in out, out


@pytest.mark.parametrize("prec", ["c", "z"])
def test_complex_error_exits(prec):
    out = run_tester(prec, [0])
    for name in NAMES_COMPLEX:
        assert re.search(rf"{prec.upper()}{name}\s+PASSED THE TESTS OF ERROR-EXITS", out), (name, out)
        assert re.search(rf"{prec.upper()}{name}\s+PASSED THE COMPUTATIONAL TESTS \(\s*\d+ CALLS\)", out), (name, out)
    assert f"{prec.upper()}GEMM  PASSED THE COMPUTATIONAL TESTS (    81 CALLS)" in out, out
    assert "FAILED" not in out and "SUSPECT" not in out, out


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["c", "z"])
def test_complex_computational(gpu, prec):
    """All nine complex routines: every op / side / uplo / diag combination,
    complex alpha and beta (real for HERK, HER2K's beta)."""
    out = run_tester(prec, [0, 1, 2, 3, 5, 9])
    for name in NAMES_COMPLEX:
        assert re.search(rf"{prec.upper()}{name}\s+PASSED THE TESTS OF ERROR-EXITS", out), (name, out)
        m = re.search(rf"{prec.upper()}{name}\s+PASSED THE COMPUTATIONAL TESTS \(\s*(\d+) CALLS\)", out)
        assert m and int(m.group(1)) > 1000, (name, out)
    assert "FAILED" not in out and "SUSPECT" not in out, out
