"""The reference-style ensemble session (GPU): fca_hip_ens_open/run/counts/
close, which replace fca_cuda_ens_* (src/cuda/cuda_backend.h:19-32,
kernels.cu:836-1043). One grid stays resident; each run resets the state from
the scenario's grid (warm start or dry), takes its own forcing (rain scaled
as the reference's --lambda does, hyetograph, horizon) and adds its flooded
cells into per-cell counts kept on the device. Each scenario must end as the
oracle ends it (accumulators bit for bit in FP32), its flooded counts and
largest depth must equal the oracle's count rule on the oracle's hmax, and the
count rasters (uint16 and uint32) must equal the sum of the oracle's
per-scenario rasters."""
import numpy as np
import pytest

from floodca_amd import abi as A
from floodca_amd import synth
from floodca_amd.hip import Ensemble
from oracle import Oracle

pytestmark = pytest.mark.gpu

THR = (0.10, 0.25, 0.50)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_ensemble_session_matches_the_oracle_scenario_by_scenario(prec):
    a = A.abi(prec)
    base = synth.make_grid(96, 131, seed=31, dtype=a.dtype, rday=3.0)
    cfg0 = A.default_cfg(a)
    A.set_dx(a, cfg0, base.dx)
    o = Oracle(prec)
    print(f"case ensemble session {prec}", flush=True)
    ens = Ensemble(base, cfg0, THR, prec=prec)
    total = np.zeros((len(THR), base.z.size), dtype=np.uint64)
    try:
        for k, (lam, hyeto, D_h, T) in enumerate([(1.0, A.FCA_HYETO_TRIANGLE, 2.0, 5400.0),
                                                  (2.5, A.FCA_HYETO_TRIANGLE, 1.0, 3600.0),
                                                  (0.7, A.FCA_HYETO_UNIFORM, 24.0, 4000.0),
                                                  (4.0, A.FCA_HYETO_TRIANGLE, 0.5, 2700.0)]):
            g = base.copy()
            g.Rday = (base.Rday * a.dtype.type(lam)).astype(a.dtype)
            if k == 1:  # a warm start: hmax = h0
                g.h[g.mask == A.FCA_LAND] = a.dtype.type(0.08)
                g.hmax[...] = g.h
            cfg = A.default_cfg(a)
            A.set_dx(a, cfg, g.dx)
            cfg.hyeto, cfg.D_h, cfg.phi, cfg.T_sim = hyeto, D_h, 0.25, T
            go = g.copy()
            acc_o, _ = o.run(go, cfg)
            acc = o.acc_init(g, cfg)
            fl, md, _ = ens.run(g, cfg, acc)
            assert acc.step == acc_o.step and acc.t == pytest.approx(acc_o.t, rel=1e-15)
            if prec == "f32":
                for f in ("V", "V_rain", "V_sea", "V_bnd", "clamp_err"):
                    ref = getattr(acc_o, f)
                    assert getattr(acc, f) == pytest.approx(ref, rel=1e-9, abs=1e-9), f
                assert acc.cap_active == acc_o.cap_active
            counts = np.zeros(len(THR) * base.z.size, dtype=np.uint32)
            fl_o, md_o = o.ens_count(go, THR, counts)
            assert fl == fl_o and md == pytest.approx(md_o, abs=1e-12 if prec == "f64" else 0)
            total += counts.reshape(len(THR), -1)
        c32 = ens.counts(u32=True)
        c16 = ens.counts(u32=False)
    finally:
        ens.close()
    assert np.array_equal(c32.astype(np.uint64), total)
    assert np.array_equal(c16.astype(np.uint64), np.minimum(total, 65535))
    assert total.max() >= 2


def test_poll_runs_after_every_poll_every_steps():
    """fca_hip_ens_run calls poll(pctx) after every poll_every steps, as the
    reference's loop does (kernels.cu:967-968), and polling leaves the
    results unchanged."""
    a = A.abi("f32")
    base = synth.make_grid(64, 77, seed=33, dtype=a.dtype, rday=3.0)
    cfg = A.default_cfg(a)
    A.set_dx(a, cfg, base.dx)
    cfg.T_sim = 5400.0
    o = Oracle("f32")
    print("case ensemble poll", flush=True)
    ens = Ensemble(base, cfg, THR)
    try:
        g1 = base.copy()
        acc1 = o.acc_init(g1, cfg)
        r1 = ens.run(g1, cfg, acc1)[:2]
        seen = []
        g2 = base.copy()
        acc2 = o.acc_init(g2, cfg)
        r2 = ens.run(g2, cfg, acc2, poll=lambda ctx: seen.append(1), poll_every=7)[:2]
    finally:
        ens.close()
    assert acc1.step == acc2.step > 20
    assert len(seen) == acc2.step // 7
    assert r1 == r2
    for f in ("t", "V", "V_rain", "V_sea", "V_bnd", "cap_total"):
        assert getattr(acc1, f) == getattr(acc2, f), f


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_ensemble_session_reloads_forcing_days(prec, tmp_path):
    """A scenario with day files (model.c:338-358) run in the session: the day
    reloads happen per scenario and the result is the oracle's; a second
    scenario without forcing on the same session is unaffected by them."""
    a = A.abi(prec)
    base = synth.make_grid(48, 61, seed=35, dtype=a.dtype)
    cfg0 = A.default_cfg(a)
    A.set_dx(a, cfg0, base.dx)
    rng = np.random.default_rng(6)
    for d in range(2):
        (rng.random(base.shape) * (0.05 + 0.1 * d)).astype(np.float32).tofile(tmp_path / f"r{d}.bin")
    o = Oracle(prec)
    print(f"case ensemble forcing {prec}", flush=True)
    ens = Ensemble(base, cfg0, THR, prec=prec)
    try:
        for forced in (True, False):
            cfg = A.default_cfg(a)
            A.set_dx(a, cfg, base.dx)
            cfg.hyeto, cfg.T_sim = A.FCA_HYETO_UNIFORM, 86400.0 * 2 if forced else 7200.0
            if forced:
                cfg.lambda_ = 1.1
                cfg.n_force_days = 2
                for d in range(2):
                    cfg.force_files[d].value = str(tmp_path / f"r{d}.bin").encode()
            g, go = base.copy(), base.copy()
            acc_o, _ = o.run(go, cfg)
            acc = o.acc_init(g, cfg)
            fl, md, _ = ens.run(g, cfg, acc)
            assert acc.step == acc_o.step and acc.t == acc_o.t
            assert acc.force_day == acc_o.force_day == (1 if forced else -1)
            for f in ("V", "V_rain", "V_sea", "V_bnd", "clamp_err"):
                ref = getattr(acc_o, f)
                assert getattr(acc, f) == pytest.approx(ref, rel=1e-9, abs=1e-9), f
            counts = np.zeros(len(THR) * base.z.size, dtype=np.uint32)
            fl_o, md_o = o.ens_count(go, THR, counts)
            assert fl == fl_o
            assert md == pytest.approx(md_o, abs=1e-12 if prec == "f64" else 0)
    finally:
        ens.close()
