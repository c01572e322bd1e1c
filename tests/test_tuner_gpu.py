"""The autotuner on the device: eh_time_config / eh_time_config_packed, and
EH_TUNE_TABLE tuning a new shape on first use, for the plain kernels and the
packed products (split-BF16 with 3 terms, bf16 inputs).

Integer-valued inputs keep every configuration's result exact, so a product
run under the table's choice must equal the exact product (and, where a
configuration can be named through eh_gemm_config, that configuration's run
bit for bit)."""
import os

import numpy as np
import pytest

import eigen_amd as E
import oracle as O
from helpers import from_dev, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture
def table_mode():
    prev = E.lib().eh_get_tune_mode()
    E.tuner_clear()
    E.set_tune_mode(E.TUNE_TABLE)
    yield
    E.set_tune_mode(prev)
    E.tuner_clear()


def _ints(shape, seed, dtype=np.float32):
    return np.asfortranarray(np.random.default_rng(seed).integers(-8, 9, shape).astype(dtype))


def test_time_config_every_plain_config(gpu):
    m, n, k = 384, 256, 192
    A, B = _ints((m, k), 1), _ints((k, n), 2)
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(np.zeros((m, n), np.float32))
    want = A.astype(np.float64) @ B.astype(np.float64)
    for c in range(E.num_configs()):
        ms = E.time_config(E.F32, 0, 0, m, n, k, dA, dB, dC, c, reps=2)
        assert 0.0 < ms < 1000.0, (c, ms)
        assert np.array_equal(from_dev(dC, (m, n)).astype(np.float64), want), c  # C = A*B, beta = 0


@pytest.mark.parametrize("kind", [E.PACKED_SPLIT2, E.PACKED_SPLIT3, E.PACKED_BF16, E.PACKED_F16])
def test_time_config_every_packed_config(gpu, kind):
    m, n, k = 300, 260, 520
    A, B = _ints((m, k), 3), _ints((k, n), 4)
    if kind in (E.PACKED_BF16, E.PACKED_F16):  # integers in [-8, 8] are exact in bf16 and half
        conv = (lambda x: O.bf16_rtne(x)) if kind == E.PACKED_BF16 else (lambda x: O.half_rtne(x))
        dA = to_dev(conv(A).reshape(A.shape, order="F").view(np.int16))
        dB = to_dev(conv(B).reshape(B.shape, order="F").view(np.int16))
    else:
        dA, dB = to_dev(A), to_dev(B)
    dC = to_dev(np.full((m, n), np.nan, np.float32))
    want = A.astype(np.float64) @ B.astype(np.float64)
    nc = E.num_packed_configs(kind)
    assert nc == (8 if kind in (E.PACKED_BF16, E.PACKED_F16) else 16)
    for c in range(nc):
        ms = E.time_config_packed(kind, 0, 0, m, n, k, dA, dB, dC, c, reps=2)
        assert 0.0 < ms < 1000.0, (kind, c, ms)
        assert np.array_equal(from_dev(dC, (m, n)).astype(np.float64), want), (kind, c)


def test_table_mode_tunes_plain_shape_on_first_use(gpu, table_mode):
    m, n, k = 640, 384, 200
    A, B = _ints((m, k), 5), _ints((k, n), 6)
    dA, dB = to_dev(A), to_dev(B)
    dC = to_dev(np.zeros((m, n), np.float32))
    assert E.tuner_count() == 0
    E.gemm(E.F32, 0, 0, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, prec=E.PREC_PLAIN)
    assert E.tuner_count() == 1  # tuned on first use and remembered
    got = from_dev(dC, (m, n))
    c = E.plan(E.F32, 0, 0, m, n, k)
    assert E.tuner_count() == 1  # the plan reads the table, no second tuning
    dC2 = to_dev(np.zeros((m, n), np.float32))
    E.gemm(E.F32, 0, 0, m, n, k, 1.0, dA, m, dB, k, 0.0, dC2, m, prec=E.PREC_PLAIN, config=c)
    assert np.array_equal(got, from_dev(dC2, (m, n)))
    assert np.array_equal(got.astype(np.float64), A.astype(np.float64) @ B.astype(np.float64))


def test_table_mode_tunes_split_and_bf16_shapes(gpu, table_mode):
    m, n, k = 520, 300, 700
    A, B = _ints((m, k), 7), _ints((k, n), 8)
    want = A.astype(np.float64) @ B.astype(np.float64)
    dA, dB = to_dev(A), to_dev(B)
    dC = to_dev(np.zeros((m, n), np.float32))
    E.gemm(E.F32, 0, 0, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, prec=E.PREC_SPLIT3)
    assert E.tuner_count() == 1
    c = E.plan_packed(E.PACKED_SPLIT3, 0, 0, m, n, k)
    assert 0 <= c < E.num_packed_configs(E.PACKED_SPLIT3)
    got = from_dev(dC, (m, n))
    dC2 = to_dev(np.zeros((m, n), np.float32))
    E.gemm(E.F32, 0, 0, m, n, k, 1.0, dA, m, dB, k, 0.0, dC2, m, prec=E.PREC_SPLIT3, config=c)
    assert np.array_equal(got, from_dev(dC2, (m, n)))
    assert np.array_equal(got.astype(np.float64), want)
    # bf16 inputs: a separate kind of the same shape is tuned on its own
    hA = O.bf16_rtne(A).reshape(A.shape, order="F").view(np.int16)
    hB = O.bf16_rtne(B).reshape(B.shape, order="F").view(np.int16)
    dC3 = to_dev(np.zeros((m, n), np.float32))
    E.gemm_bf16(0, 0, m, n, k, 1.0, to_dev(hA), m, to_dev(hB), k, 0.0, dC3, m)
    assert E.tuner_count() == 2
    assert np.array_equal(from_dev(dC3, (m, n)).astype(np.float64), want)


def test_table_save_load_round_trip(gpu, table_mode, tmp_path):
    E.lib().eh_tuner_record(E.F64, 0, 1, 100, 200, 300, 5)
    E.lib().eh_tuner_record_packed(E.PACKED_SPLIT3, 1, 0, 400, 500, 600, 9)
    path = str(tmp_path / "table.txt")
    E.tuner_save(path)
    text = open(path).read().splitlines()
    assert text[0].startswith("# eigen_amd tuning table v2 arch=gfx950 kernels=")
    rows = [line.split() for line in text if not line.startswith("#")]
    assert sorted(rows) == sorted([["0", "1", "0", "1", "100", "200", "300", "5"],
                                   ["3", "0", "1", "0", "400", "500", "600", "9"]])
    E.tuner_clear()
    E.tuner_load(path)
    assert E.tuner_count() == 2
    assert E.plan(E.F64, 0, 1, 100, 200, 300) == 5
    assert E.plan_packed(E.PACKED_SPLIT3, 1, 0, 400, 500, 600) == 9
    # a v1 line (plain kernels, no kind column) still loads
    v1 = str(tmp_path / "v1.txt")
    with open(v1, "w") as f:
        f.write("# eigen_amd tuning table v1: dtype opa opb m n k config\n0 0 0 64 64 64 7\n")
    E.tuner_load(v1)
    assert E.plan(E.F32, 0, 0, 64, 64, 64) == 7
    # a table stamped for another architecture is refused
    other = str(tmp_path / "other.txt")
    with open(other, "w") as f:
        f.write("# eigen_amd tuning table v2 arch=gfx942\n0 0 0 0 64 64 64 3\n")
    with pytest.raises(E.EigenAmdError):
        E.tuner_load(other)
    # and one made for another kernel set
    stamp = text[0].split("kernels=")[1].split()[0]
    stale = str(tmp_path / "stale.txt")
    with open(stale, "w") as f:
        f.write(f"# eigen_amd tuning table v2 arch=gfx950 kernels={int(stamp) + 1}\n0 0 0 0 64 64 64 3\n")
    with pytest.raises(E.EigenAmdError):
        E.tuner_load(stale)
    assert E.plan(E.F32, 0, 0, 64, 64, 64) == 7
