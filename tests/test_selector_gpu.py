"""The learned selector against the method without it (runs on an MI355X).

Trains the 7-64-24 network on the device (the library's own HIP forward /
backward / Adam kernels) from the committed autotuner table: 4,000 shapes,
M, N, K log-uniform in [64, 16384], all 24 configurations timed.  On the 500
held-out shapes the regret of its choice against exhaustive tuning
(t(choice) / t(best) - 1, measured against the table) must beat the closed-form
heuristic's."""
import os

import numpy as np
import pytest

import eigen_amd as E
from eigen_amd import selector as S

pytestmark = pytest.mark.gpu

DB = os.path.join(os.path.dirname(E.__file__), "data", "tune_gfx950.csv")


def test_selector_beats_heuristic_on_heldout(gpu):
    sel, res = S.train_and_evaluate(DB, epochs=3000)
    print(res)
    assert res["train_shapes"] >= 4000 and res["heldout_shapes"] == 500
    assert res["loss_last"] < res["loss_first"]
    assert res["mlp_regret"]["mean"] < res["heuristic_regret"]["mean"]
    assert res["mlp_regret"]["median"] <= res["heuristic_regret"]["median"]
    # the planner in MLP mode picks exactly what the device inference picks
    shapes, _ = S.load_db(DB)
    sel.install()
    prev = E.lib().eh_get_tune_mode()
    E.set_tune_mode(E.TUNE_MLP)
    try:
        held = shapes[4000:4100]
        dev = sel.predict(S.features(held))
        host = np.array([E.plan(d, ta, tb, m, n, k) for d, ta, tb, m, n, k in held.tolist()])
        assert np.mean(dev == host) >= 0.99  # FP32 ties aside, host and device inference agree
    finally:
        E.set_tune_mode(prev)
        E.check(E.lib().eh_planner_set_mlp(None, None))
        sel.close()
