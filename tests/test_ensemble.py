"""Scenario ensembles: the seeded draw and per-scenario inputs (CPU), and the
batched session against one session per scenario on a small ensemble (GPU)."""
import numpy as np
import pytest

from floodca_amd import abi as A
from floodca_amd import ensemble as E
from floodca_amd import synth


@pytest.fixture(scope="module")
def base():
    return synth.make_grid(300, 420, seed=9)


def test_draw_is_seeded_and_in_range():
    a = E.draw(50, 3, (1000, 800), side=(64, 256), dur=(2.0, 12.0))
    b = E.draw(50, 3, (1000, 800), side=(64, 256), dur=(2.0, 12.0))
    assert [(s.r0, s.c0, s.side, s.factor, s.dur_h) for s in a] == \
           [(s.r0, s.c0, s.side, s.factor, s.dur_h) for s in b]
    for s in a:
        assert 64 <= s.side <= 256 and 0 <= s.r0 <= 1000 - s.side and 0 <= s.c0 <= 800 - s.side
        assert 2.0 <= s.dur_h <= 12.0 and s.factor > 0
        assert np.all((s.nscale >= 0.8) & (s.nscale <= 1.2))
    f = np.array([s.factor for s in E.draw(4000, 1, (10, 10), side=(1, 1))])
    assert abs(np.median(f) - 1.0) < 0.05 and abs(np.std(np.log(f)) - 0.5) < 0.03


def test_day_draw_follows_the_reference_generator_distributions():
    """draw_days uses the distributions of the reference's scenario generator
    (prep/floodca_prep/scenarios.py:44-46): lambda U[0.8, 1.5], D_h one of
    6, 12, 24 h, phi U[0.25, 0.75]; seeded."""
    a = E.draw_days(4000, 7)
    assert [(s.lam, s.dur_h, s.phi) for s in a] == \
           [(s.lam, s.dur_h, s.phi) for s in E.draw_days(4000, 7)]
    lam = np.array([s.lam for s in a])
    phi = np.array([s.phi for s in a])
    dur = np.array([s.dur_h for s in a])
    assert lam.min() >= 0.8 and lam.max() < 1.5 and abs(lam.mean() - 1.15) < 0.01
    assert phi.min() >= 0.25 and phi.max() < 0.75 and abs(phi.mean() - 0.5) < 0.01
    assert set(dur.tolist()) == {6, 12, 24}
    assert all(abs(np.mean(dur == d) - 1 / 3) < 0.03 for d in (6, 12, 24))


def test_day_crop_and_config():
    sc = E.DayScenario(0, 1.25, 12, 0.4)
    cr = E.day_crop(E.DAY_GRID, sc)
    assert (cr["r0"], cr["c0"], cr["nrows"], cr["ncols"]) == (0, 0) + E.DAY_GRID
    assert cr["rday"] == float(np.float32(E.DAY_RDAY * 1.25))
    assert cr["n_class"] == [float(v) for v in synth.MANNING_CLASSES.astype(np.float32)]
    cfg = E.day_config(sc, 90.0)
    assert cfg.hyeto == A.FCA_HYETO_TRIANGLE and cfg.D_h == 12.0
    assert cfg.phi == pytest.approx(0.4) and cfg.T_sim == 86400.0 and cfg.steps == 0
    assert cfg.phys.dx == 90.0 and cfg.phys.A == 8100.0


def test_build_crops_scales_roughness_and_sets_the_storm(base):
    sc = E.Scenario(idx=0, r0=10, c0=20, side=100, factor=1.5, dur_h=4.0,
                    nscale=np.linspace(0.8, 1.2, 8))
    g, cfg = E.build(base, sc, "f32")
    assert g.shape == (100, 100)
    assert np.array_equal(g.z, base.z[10:110, 20:120])
    assert np.array_equal(g.mask, base.mask[10:110, 20:120])
    cls = np.searchsorted(synth.MANNING_CLASSES.astype(np.float32), base.n[10:110, 20:120])
    assert np.array_equal(g.n, (synth.MANNING_CLASSES[cls] * sc.nscale[cls]).astype(np.float32))
    assert np.all(g.Rday == np.float32(0.05 * 1.5 * 4.0)) and np.all(g.C == 1) and not g.h.any()
    assert cfg.hyeto == A.FCA_HYETO_TRIANGLE and cfg.D_h == 4.0 and cfg.phi == 0.5
    assert cfg.T_sim == 4.0 * 3600.0 and cfg.steps == 0


def test_acc_start_matches_the_reference_init(base):
    from oracle import Oracle
    g = base.copy()
    g.h[...] = np.where(g.mask == A.FCA_LAND, 0.01, 0.0).astype(np.float32)
    a = A.abi("f32")
    cfg = A.default_cfg(a)
    A.set_dx(a, cfg, g.dx)
    acc = E.acc_start("f32", g, cfg)
    ref = Oracle("f32").acc_init(g, cfg)
    assert acc.V0 == pytest.approx(ref.V0, rel=1e-12) and acc.force_day == ref.force_day
    assert acc.t == ref.t and acc.step == ref.step and acc.dt_base == ref.dt_base


@pytest.mark.gpu
@pytest.mark.timeout(300)
def test_batched_ensemble_equals_one_session_per_scenario(base):
    scen = E.draw(8, 11, base.shape, side=(40, 200), dur=(0.3, 1.0))
    cases = [E.build(base, s, "f32") for s in scen]
    tb, rb = E.run_batched(cases, "f32", strip_rows=16)
    ta, ra = E.run_alone(cases, "f32", strip_rows=16)
    assert rb == ra
    assert all(r[2] > 0 for r in rb)
