"""Crop scenarios built on the device against the same scenarios built on the
host (GPU): a batch opened from a resident base grid (fca_hip_batch_open_crops)
ends every scenario bit for bit where a batch of host-built crop grids
(fca_hip_batch_open) ends it -- h, hmax, tmax, the accumulators, the step
count and the flooded counts -- and a crop run alone ends where it ends in a
batch. The host-built grids are themselves checked against the oracle by the
batch tests, so this pins the device build of a scenario (window, roughness
per class, storm depth, dry start, open-face count)."""
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
