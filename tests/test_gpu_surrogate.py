"""The surrogate-screened ensemble end to end at a small size (GPU).

Dataset from the batched CA, training with the library's own kernels,
screening, CA reruns of the flagged members and of a seeded audit sample of
the unflagged ones. What the network gains over the
method without it: on held-out scenarios with CA truth, its hmax maps beat
the best storm-depth-only predictor (hmax = c x depth on LAND, c fitted on
the training rows) by at least 10% in RMSE, and its predicted flooded area
ranks the scenarios in the CA's order (Spearman > 0.6).
"""
import numpy as np
import pytest

from floodca_amd import synth
from floodca_amd import surrogate as SG

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pipeline():
    base = synth.make_grid(2048, 2048, seed=20261207)
    pc = SG.PipelineCfg(n_train=576, n_val=96, n_cand=4000, n_alone=6, epochs=60, batch=32,
                        lr=2e-3, side=(256, 640), dur=(1.0, 3.0), seed=5, ca_batch=288)
    return SG.run_pipeline(base, pc, log=lambda m: print(m, flush=True)), pc


def test_network_beats_the_method_without_it(pipeline):
    res, _ = pipeline
    assert np.isfinite(res["val_rmse_m"])
    assert res["val_rmse_m"] < 0.9 * res["val_rmse_depth_baseline_m"]
    assert res["val_rank_spearman"] > 0.6


def test_training_lowered_the_error(pipeline):
    res, _ = pipeline
    tr = res["train_rmse_by_epoch_m"]
    va = res["val_rmse_by_epoch_m"]
    assert tr[-1][1] < 0.8 * tr[0][1]
    assert va[-1][1] < va[0][1]


def test_selection_and_reruns(pipeline):
    res, pc = pipeline
    o = res["_objects"]
    sc, top_a, top_u = o["screen"], o["top_a"], o["top_u"]
    n = pc.n_cand
    assert len(sc.area) == n and np.isfinite(sc.area).all() and np.isfinite(sc.sigma).all()
    assert len(top_a) == len(top_u) == int(np.ceil(0.02 * n))
    assert not set(top_a) & set(top_u)
    assert sc.area[top_a].min() >= np.delete(sc.area, top_a).max()
    rest = np.setdiff1d(np.arange(n), np.concatenate([top_a, top_u]))
    assert sc.sigma[top_u].min() >= sc.sigma[rest].max()
    audit = o["audit"]
    assert len(audit) == int(np.ceil(0.01 * n)) and np.all(np.diff(audit) > 0)
    assert not set(audit) & (set(top_a) | set(top_u))
    assert res["judged"] == len(top_a) + len(top_u) + len(audit)
    for k in ("judged_rmse_m", "judged_rmse_area_m", "judged_rmse_uncertainty_m",
              "judged_rmse_audit_m"):
        assert np.isfinite(res[k]) and res[k] > 0, k
    assert res["scenarios_per_h"] > res["all_ca_alone_scenarios_per_h"]


def test_screening_is_deterministic(pipeline):
    res, pc = pipeline
    o = res["_objects"]
    again = SG.screen(o["net"], o["terrain"], o["cands"][:512], pc.R, pc.screen_batch)
    assert np.array_equal(again.area, o["screen"].area[:512])
    assert np.array_equal(again.sigma, o["screen"].sigma[:512])
