"""Host logic of the surrogate's selection (CPU): which candidates the CA
reruns. The top fraction by predicted flooded area, then the top fraction by
predicted spread among the rest (ties broken by index), and a seeded audit
sample drawn uniformly from the members neither ranking flagged."""
import numpy as np

from floodca_amd import surrogate as SG


def _screen(n, seed):
    rng = np.random.default_rng(seed)
    area = rng.integers(0, 50, n).astype(float)      # many ties
    sigma = rng.integers(0, 20, n).astype(float)
    return SG.Screen(area=area, sigma=sigma, volume=area * 0.0, seconds=0.0)


def test_select_takes_the_top_by_area_then_by_spread_among_the_rest():
    n = 1000
    sc = _screen(n, 3)
    top_a, top_u = SG.select(sc, 0.05, 0.03)
    assert len(top_a) == 50 and len(top_u) == 30
    # brute force: the same order written as a sort of (-score, index) pairs
    want_a = [i for _, i in sorted((-sc.area[i], i) for i in range(n))][:50]
    assert list(top_a) == want_a
    rest = [i for i in range(n) if i not in set(want_a)]
    want_u = [i for _, i in sorted((-sc.sigma[i], i) for i in rest)][:30]
    assert list(top_u) == want_u


def test_select_rounds_counts_up():
    sc = _screen(7, 1)
    top_a, top_u = SG.select(sc, 0.02, 0.02)
    assert len(top_a) == 1 and len(top_u) == 1


def test_audit_sample_is_seeded_uniform_and_disjoint():
    n = 20000
    flagged = np.arange(0, 2400, 3)
    a = SG.audit_sample(n, flagged, 0.01, 11)
    assert len(a) == 200 and np.all(np.diff(a) > 0)
    assert not set(a.tolist()) & set(flagged.tolist())
    assert np.array_equal(a, SG.audit_sample(n, flagged, 0.01, 11))
    assert not np.array_equal(a, SG.audit_sample(n, flagged, 0.01, 12))
    # uniform over the rest: over many seeds every unflagged member is drawn
    # about equally often (here: the mean index of a draw is the rest's mean)
    rest = np.setdiff1d(np.arange(n), flagged)
    means = [SG.audit_sample(n, flagged, 0.01, s).mean() for s in range(200)]
    assert abs(np.mean(means) - rest.mean()) < 0.02 * rest.mean()


def test_audit_sample_edge_cases():
    assert len(SG.audit_sample(10, [0, 1], 0.0, 1)) == 0
    assert sorted(SG.audit_sample(5, [0, 1, 2], 1.0, 1).tolist()) == [3, 4]
    assert len(SG.audit_sample(4, [0, 1, 2, 3], 0.5, 1)) == 0


def test_rmse_interval_brackets_the_rmse_and_narrows_with_more_rows():
    rng = np.random.default_rng(5)
    w = rng.uniform(100, 1000, 400)
    se = np.stack([w * rng.gamma(2.0, 0.0008, 400), w], 1)   # per-row (sum sq err, weight)
    r = SG.rmse(se)
    lo, hi = SG.rmse_interval(se, seed=1)
    assert lo < r < hi
    lo4, hi4 = SG.rmse_interval(np.tile(se, (4, 1)), seed=1)
    assert hi4 - lo4 < 0.7 * (hi - lo)       # ~1/2 for 4x the rows
    assert SG.rmse_interval(se, seed=1) == (lo, hi)
    assert all(np.isnan(SG.rmse_interval(np.zeros((0, 2)))))
