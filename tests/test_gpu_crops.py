"""Crop scenarios (GPU).

* Config 3's own scenarios against the oracle: the ensemble generator's draw
  (side, storm duration, lognormal storm factor, per-class Manning scale)
  built on the host for the oracle and on the device for a batch, triangle
  storms run to T_sim = D; FP32 bit for bit, FP64 within 1e-12 m.
* Device-built against host-built: a batch opened from a resident base grid
  (fca_hip_batch_open_crops) ends every scenario bit for bit where a batch of
  host-built crop grids (fca_hip_batch_open) ends it -- h, hmax, tmax, the
  accumulators, the step count and the flooded counts -- and a crop run alone
  ends where it ends in a batch. This pins the device build of a scenario
  (window, roughness per class, storm depth, dry start, open-face count).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from floodca_amd import abi as A
from floodca_amd import ensemble as E
from floodca_amd import synth
from floodca_amd.hip import Batch, HipError

pytestmark = pytest.mark.gpu


def _base(prec, size=768):
    return synth.make_grid(size, size, seed=20261209, dtype=A.abi(prec).dtype)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_device_crops_equal_host_built_grids(prec):
    base = _base(prec)
    scens = E.draw(10, 7, base.shape, side=(24, 300), dur=(0.5, 1.5))
    scens[0].r0, scens[0].c0 = 0, 0                      # touches the base's corner
    scens[1].r0 = base.shape[0] - scens[1].side          # and its far edge
    print(f"case crops {prec}", flush=True)
    cases = [E.build(base, s, prec) for s in scens]
    ref = Batch([g for g, _ in cases], [c for _, c in cases],
                [E.acc_start(prec, g, c) for g, c in cases], prec=prec)
    ref.run()
    with E.device_base(base, prec) as db:
        got = Batch.from_crops(db, [E.crop_of(s, prec) for s in scens],
                               [E.config_of(s, base.dx, prec) for s in scens], prec=prec)
        got.run()
        for i, (g, _) in enumerate(cases):
            acc_r = ref.read(i)
            out = g.copy()
            acc_g = got.read(i, out)
            for k in ("h", "hmax", "tmax"):
                assert np.array_equal(getattr(g, k), getattr(out, k)), (i, k)
            assert ref.steps(i) == got.steps(i) > 0
            assert ref.flooded(i) == got.flooded(i)
            for k in ("t", "step", "V", "V_rain", "V_sea", "V_bnd", "clamp_err", "cap_active",
                      "cap_total", "dt_min_hits"):
                assert getattr(acc_r, k) == getattr(acc_g, k), (i, k)
        got.close()
    ref.close()


def test_crop_alone_equals_crop_in_batch():
    base = _base("f32", 512)
    scens = E.draw(5, 11, base.shape, side=(64, 256), dur=(0.5, 1.0))
    print("case crops alone vs batch", flush=True)
    with E.device_base(base) as db:
        _, batched = E.run_batched_crops(db, scens, base.dx)
        _, alone = E.run_alone_crops(db, scens, base.dx)
    assert batched == alone


def test_crop_window_outside_the_base_is_an_error():
    base = _base("f32", 128)
    sc = E.draw(1, 3, base.shape, side=(64, 64))[0]
    crop = E.crop_of(sc)
    crop["r0"] = 100
    with E.device_base(base) as db:
        with pytest.raises(HipError, match="outside"):
            Batch.from_crops(db, [crop], [E.config_of(sc, base.dx)])


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_day_scenarios_match_the_oracle(prec):
    """The reference-shaped ensemble's scenarios (bench.py): the whole,
    non-square base grid as one crop with its own event rain and storm,
    batched, against the oracle's run of the same grid built on the host
    (dry start, C = 1, Rday = the crop's depth), over the whole 24 h."""
    from oracle import Oracle
    a = A.abi(prec)
    base = synth.make_grid(70, 53, seed=20261211, dtype=a.dtype)
    scens = E.draw_days(3, 4)
    crops = [E.day_crop(base.shape, s, prec) for s in scens]
    cfgs = [E.day_config(s, base.dx, prec) for s in scens]
    print(f"case day scenarios {prec}", flush=True)
    with E.device_base(base, prec) as db:
        b = Batch.from_crops(db, crops, cfgs, prec=prec)
        b.run()
        o = Oracle(prec)
        for i, (cr, cfg) in enumerate(zip(crops, cfgs)):
            g = base.copy()
            g.h[...] = 0
            g.hmax[...] = 0
            g.tmax[...] = 0
            g.C[...] = 1
            g.Rday[...] = a.dtype.type(cr["rday"])
            acc_o, _ = o.run(g, cfg)
            out = g.copy()
            acc_g = b.read(i, out)
            assert acc_g.step == acc_o.step == b.steps(i) > 0
            if prec == "f32":
                for k in ("h", "hmax", "tmax"):
                    assert np.array_equal(getattr(g, k), getattr(out, k)), (i, k)
            else:
                assert np.abs(g.h - out.h).max() <= 1e-12
                assert np.abs(g.hmax - out.hmax).max() <= 1e-12
        b.close()


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_config3_scenarios_match_the_oracle(prec):
    """The batched ensemble's scenarios (E.draw with the config's
    distributions: storm duration U[2, 12] h with both ends taken, storm
    factor lognormal(0, 0.5), each Manning class x U[0.8, 1.2]; sides
    U[256, 300] so the oracle finishes in seconds) on crops of a 1024^2 base of
    the benchmark generator, one in its west corner (all SEA) and one in its
    east corner:
    the oracle runs E.build's host grid of each to T_sim = D, the GPU runs
    the device-built crops in one batch."""
    from oracle import Oracle

    a = A.abi(prec)
    base = synth.make_grid(1024, 1024, seed=20261213, dtype=a.dtype)
    n = 5 if prec == "f32" else 3
    scens = E.draw(n, 31, base.shape, side=(256, 300), dur=(2.0, 12.0))
    scens[0].dur_h = 2.0
    scens[1].dur_h = 12.0
    scens[0].r0, scens[0].c0 = 0, 0
    scens[-1].r0 = base.shape[0] - scens[-1].side
    scens[-1].c0 = base.shape[1] - scens[-1].side
    assert min(s.factor for s in scens) < 1.0 < max(s.factor for s in scens)
    assert all(np.any(np.abs(s.nscale - 1.0) > 0.05) for s in scens)
    o = Oracle(prec)
    cases = [E.build(base, s, prec) for s in scens]

    def oracle_run(case):
        g, cfg = case
        go = g.copy()
        acc, log = o.run(go, cfg, log_cap=1 << 16)
        return go, acc, log

    print(f"case config-3 scenarios vs oracle {prec}: sides {[s.side for s in scens]}, "
          f"D {[round(s.dur_h, 2) for s in scens]} h", flush=True)
    with ThreadPoolExecutor(len(cases)) as ex:  # the oracle releases the GIL
        refs = list(ex.map(oracle_run, cases))
    with E.device_base(base, prec) as db:
        b = Batch.from_crops(db, [E.crop_of(s, prec) for s in scens],
                             [E.config_of(s, base.dx, prec) for s in scens], prec=prec,
                             record_steps=1 << 16)
        b.run()
        for i, ((g, cfg), (go, acc_o, log_o)) in enumerate(zip(cases, refs)):
            out = g.copy()
            acc_g = b.read(i, out)
            log_g = b.steplog(i, 1 << 16)
            assert acc_o.t == pytest.approx(scens[i].dur_h * 3600.0)
            assert acc_g.step == acc_o.step == len(log_o) > 100, i
            if prec == "f32":
                for k in ("h", "hmax", "tmax"):
                    assert np.array_equal(getattr(go, k), getattr(out, k)), (i, k)
                assert [r.dt for r in log_g] == [r.dt for r in log_o], i
                assert [r.dt_base for r in log_g] == [r.dt_base for r in log_o], i
                for k in ("cap_active", "cap_total", "dt_min_hits"):
                    assert getattr(acc_g, k) == getattr(acc_o, k), (i, k)
            else:
                for k in ("h", "hmax"):
                    assert np.abs(getattr(go, k) - getattr(out, k)).max() <= 1e-12, (i, k)
            for k in ("V", "V_rain", "V_sea", "V_bnd", "clamp_err"):
                assert getattr(acc_g, k) == pytest.approx(getattr(acc_o, k), rel=1e-9,
                                                          abs=1e-6), (i, k)
            fl, md = b.flooded(i, E.THRESHOLDS)
            land = go.mask == A.FCA_LAND
            assert fl == [int(np.count_nonzero(land & (go.hmax >= t))) for t in E.THRESHOLDS]
            # the largest LAND hmax, at least 0 (ens_common.c:170-178); the
            # corner crop of the base is all SEA
            md_ref = max(0.0, float(go.hmax[land].max())) if land.any() else 0.0
            assert md == pytest.approx(md_ref, rel=1e-12 if prec == "f32" else 1e-9)
        b.close()


def test_tiles_per_wavefront_leave_results_unchanged():
    """The autotuner's second dimension: wavefronts walking 2 or 3 tiles per
    launch (a half or a third of the workgroups) end every scenario with the
    same steps, flooded counts and largest depth as one tile per wavefront,
    and with the same depths bit for bit."""
    base = _base("f32", 512)
    scens = E.draw(6, 13, base.shape, side=(64, 300), dur=(0.5, 1.0))
    print("case tiles per wavefront", flush=True)
    with E.device_base(base) as db:
        _, one = E.run_batched_crops(db, scens, base.dx, strip_rows=32)
        for k in (2, 3):
            _, many = E.run_batched_crops(db, scens, base.dx, strip_rows=32, tiles_per_wave=k)
            assert many == one, k
        crops = [E.crop_of(s) for s in scens]
        cfgs = [E.config_of(s, base.dx) for s in scens]
        b1 = Batch.from_crops(db, crops, cfgs, strip_rows=32)
        b3 = Batch.from_crops(db, crops, cfgs, strip_rows=32, tiles_per_wave=3)
        b1.run()
        b3.run()
        for i, s in enumerate(scens):
            g1 = synth.make_grid(s.side, s.side, seed=1)
            g3 = g1.copy()
            b1.read(i, g1)
            b3.read(i, g3)
            for k in ("h", "hmax", "tmax"):
                assert np.array_equal(getattr(g1, k), getattr(g3, k)), (i, k)
        b1.close()
        b3.close()
