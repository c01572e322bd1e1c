"""The packed product's LDS stage pipeline at its ends.

The packed kernel keeps up to NSTAGE - 1 direct-to-LDS stages in flight and,
before reading a stage, waits until only the stages issued after it may
still be loading.  Near the end of the k loop fewer stages are issued after
the one being read, so the wait must shrink with them; a wait that assumes a
full pipeline there lets a wave read a stage that has not landed.  (The
software-pipelined loop of the bf16 / half tiles and of two split tiles
reloads the last stage past the end, so its waits stay constant.)  These
cases run every configuration of the four split tiles (the pipelined
128x128 and 8-wave 256x128 ones, the plain 4-wave 256x128 one, and 256x256
with 16-deep stages) and of the pipelined bf16 / half tiles over k of one
to seven 32-deep slices, so every position of each loop's tail is hit, on
integer-valued inputs whose products are exact in every configuration: the
result must equal the exact product bit for bit, on every repetition."""
import numpy as np
import pytest

import eigen_amd as E
import oracle as O
from helpers import from_dev, to_dev

pytestmark = pytest.mark.gpu

KS = [32, 64, 96, 128, 160, 200, 224]


def _ints(shape, seed):
    return np.asfortranarray(np.random.default_rng(seed).integers(-8, 9, shape).astype(np.float32))


def _operands(kind, A, B):
    if kind == E.PACKED_BF16:
        return (to_dev(O.bf16_rtne(A).reshape(A.shape, order="F").view(np.int16)),
                to_dev(O.bf16_rtne(B).reshape(B.shape, order="F").view(np.int16)))
    if kind == E.PACKED_F16:
        return (to_dev(O.half_rtne(A).reshape(A.shape, order="F").view(np.int16)),
                to_dev(O.half_rtne(B).reshape(B.shape, order="F").view(np.int16)))
    return to_dev(A), to_dev(B)


@pytest.mark.parametrize("kind,configs", [
    (E.PACKED_BF16, range(0, 8)),    # 128x128 (four stages), 256x256 (three), both the pipelined loop
    (E.PACKED_F16, range(0, 8)),     # the same tiles on the FP16 instruction
    (E.PACKED_SPLIT2, range(0, 16)),  # 128x128, 256x128 (two layouts), 256x256 with 16-deep stages
    (E.PACKED_SPLIT3, range(0, 16)),  # the same four tiles with three terms
])
def test_stage_pipeline_tail_exact(gpu, kind, configs):
    m, n = 300, 270
    for k in KS:
        A, B = _ints((m, k), k), _ints((k, n), k + 1)
        want = A.astype(np.float64) @ B.astype(np.float64)
        dA, dB = _operands(kind, A, B)
        dC = to_dev(np.full((m, n), np.nan, np.float32))
        for c in configs:
            for rep in range(3):
                E.time_config_packed(kind, 0, 0, m, n, k, dA, dB, dC, c, reps=1)
                got = from_dev(dC, (m, n)).astype(np.float64)
                assert np.array_equal(got, want), (kind, c, k, rep, float(np.abs(got - want).max()))
