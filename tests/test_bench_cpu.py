"""The benchmark's own arithmetic, on the CPU.

bench.py reports every leg's error against `exact_product`, an error-free
slicing of float64 operands (four 16-bit slices per row / column, every
slice product exact in float64, summed in extended precision).  Pinned here
against exact rational arithmetic (fractions.Fraction) on small matrices:
ordinary, wide-range (the headline's sign*U(1,2)*2^U(-8,8)) and heavily
cancelling inputs, to the stated ~2^-60 relative accuracy, and shown to be
sharp enough that keeping only three of the four slices fails the same
check.  `normwise` is checked on a closed form, and the one-GPU contract of
the command line (any --gpus but 1 exits with status 2 and a message)."""
import os
import subprocess
import sys
from fractions import Fraction

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _exact(A, B):
    """A @ B in rationals (float64 entries are exact binary fractions)."""
    Af = [[Fraction(float(x)) for x in row] for row in A]
    Bf = [[Fraction(float(x)) for x in row] for row in B]
    m, k, n = len(Af), len(Bf), len(Bf[0])
    return [[sum(Af[i][t] * Bf[t][j] for t in range(k)) for j in range(n)] for i in range(m)]


def _inputs(kind, rng, m, k, n):
    if kind == "uniform":
        return rng.uniform(-1, 1, (m, k)), rng.uniform(-1, 1, (k, n))
    if kind == "wide":
        w = lambda s: rng.choice([-1.0, 1.0], s) * rng.uniform(1, 2, s) * np.exp2(rng.uniform(-8, 8, s))
        return w((m, k)), w((k, n))
    # cancelling: B's columns nearly orthogonal to A's rows, so results are
    # many orders below the terms' magnitude
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n))
    B[-1, :] = -(A[:, :-1] @ B[:-1, :])[0] / A[0, -1]
    return A, B


def _rel_errors(got, A, B):
    """|got - exact| / (|A||B|) per entry, got taken exactly (long double)."""
    ex = _exact(A, B)
    absprod = np.abs(A) @ np.abs(B)
    return np.array([[float(abs(Fraction(*np.longdouble(got[i, j]).as_integer_ratio()) - ex[i][j])) / absprod[i, j]
                      for j in range(len(ex[0]))] for i in range(len(ex))])


@pytest.mark.parametrize("kind", ["uniform", "wide", "cancel"])
def test_exact_product_against_rationals(kind):
    """Within 2^-60 of |A||B| entry by entry (long double keeps 64 bits, and
    the slices leave less than 2^-64 of each row's largest entry)."""
    rng = np.random.default_rng({"uniform": 1, "wide": 2, "cancel": 3}[kind])
    A, B = _inputs(kind, rng, 6, 40, 5)
    assert _rel_errors(bench.exact_product(A, B), A, B).max() <= 2.0 ** -60, kind


def test_three_slices_are_not_enough():
    """The check above is sharp: cut into three 16-bit slices (48 of
    float64's 53 bits) the same product misses by far more than 2^-60."""
    rng = np.random.default_rng(4)
    A, B = _inputs("uniform", rng, 4, 30, 3)

    def three_slice_product(A, B):
        def cut(X, axis):
            mx = np.max(np.abs(X), axis=axis, keepdims=True)
            e = np.ceil(np.log2(mx))
            out, R = [], X.copy()
            for s in range(3):
                sigma = np.ldexp(1.5, (e + 52 - 16 * (s + 1)).astype(np.int64))
                hi = (R + sigma) - sigma
                out.append(hi)
                R = R - hi
            return out
        acc = np.zeros((A.shape[0], B.shape[1]), np.longdouble)
        for a in cut(A, 1):
            for b in cut(B, 0):
                acc += (a @ b).astype(np.longdouble)
        return acc

    assert _rel_errors(bench.exact_product(A, B), A, B).max() <= 2.0 ** -60
    assert _rel_errors(three_slice_product(A, B), A, B).max() > 2.0 ** -56


def test_normwise_closed_form():
    """||C - C_ref||_F / (K u ||A||_F ||B||_F), with the sampled-rows scale."""
    ref = np.zeros((2, 2))
    got = np.array([[3.0, 0.0], [0.0, 4.0]])
    k, u, na, nb = 8, 2.0 ** -24, 2.0, 0.5
    assert bench.normwise(got, ref, k, u, na, nb) == pytest.approx(5.0 / (k * u * na * nb))
    assert bench.normwise(got, ref, k, u, na, nb, scale=2.0) == pytest.approx(10.0 / (k * u * na * nb))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
