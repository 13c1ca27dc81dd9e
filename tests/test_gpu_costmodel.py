"""The autotuner's cost model on the committed timing database
(impeller_amd/tuned/timings.json, measured by impeller_amd.tune_main on an
MI355X): for held-out problems, a network trained on the GPU by the
library's MLP kernels on every other problem's rows picks a configuration
without sweeping the problem; its measured time is compared with the
exhaustive best, with the configuration used without the model
(default_ntt_plan / default_msm), and with the same network untrained."""
import os

import numpy as np
import pytest

from impeller_amd import autotune as A
from impeller_amd.costmodel import CostModel

pytestmark = pytest.mark.gpu
TIMINGS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "impeller_amd", "tuned", "timings.json")
HOLDOUT = ["ntt_fwd:24", "ntt_coset:22", "ntt_inv:21", "msm:22:0", "msm:20:1", "msm:17:2"]


def _default(problem, rows):
    parts = problem.split(":")
    cfg = A.default_msm(int(parts[1])) if parts[0] == "msm" else A.default_ntt_plan(int(parts[1]))
    return next(r for r in rows if tuple(r["config"]) == tuple(cfg))


def test_cost_model_picks_vs_exhaustive_and_default(stream):
    db = A.TimingDB.load(TIMINGS)
    assert len(db.rows) > 1000, "committed timing database missing"
    model_r, default_r, untrained_r = [], [], []
    for p in HOLDOUT:
        rows = db.for_problem(p)
        best = min(r["ms"] for r in rows)
        model = A.fit_model(db, stream, [r for r in db.rows if r["problem"] != p], steps=1500)
        model_r.append(A.pick(model, rows, stream)["ms"] / best - 1.0)
        default_r.append(_default(p, rows)["ms"] / best - 1.0)
        untrained = CostModel(A.N_FEATURES, 64, seed=0)
        untrained.mu = model.mu
        untrained.sd = model.sd
        untrained_r.append(A.pick(untrained, rows, stream)["ms"] / best - 1.0)
    model_r, default_r, untrained_r = map(np.array, (model_r, default_r, untrained_r))
    # pick within 5% of the exhaustive best on average, and better than both
    # the no-model configuration and the untrained network
    assert model_r.mean() <= 0.05, model_r
    assert model_r.mean() < default_r.mean(), (model_r, default_r)
    assert model_r.mean() < untrained_r.mean(), (model_r, untrained_r)
