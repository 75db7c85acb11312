"""fca_hip_run's contract as a drop-in for fca_cuda_run (GPU), from
src/cuda/cuda_backend.h:12-14 and kernels.cu:617-776: the whole run on the
current device; h, hmax, tmax written back into the caller's grid; in acc the
balance terms V_rain, clamp_err, V_sea and V_bnd are added to what the caller
passed, while t, step, dt_base, the cap counts, dt_min_hits and the final V
are set; the timers the backend measures are filled and the per-phase ones it
cannot separate are -1; an empty cfg->reduce_impl gets the backend's default.
Also fca_hip_device_name and fca_hip_select_device (local rank modulo the
visible devices)."""
import numpy as np
import pytest

from floodca_amd.hip import backend
from oracle import Oracle

from conftest import make_case

pytestmark = pytest.mark.gpu

OFFSETS = {"V_rain": 5.0e3, "clamp_err": 2.5, "V_sea": 7.0e2, "V_bnd": 1.25e2}


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_run_writes_back_and_accumulates(prec):
    a, g, cfg = make_case(47, 61, prec=prec, T_sim=2400.0, seed=17, walls=True)
    o = Oracle(prec)
    go, gh = g.copy(), g.copy()
    acc_o, _ = o.run(go, cfg)
    acc_h = o.acc_init(gh, cfg)
    for k, v in OFFSETS.items():
        setattr(acc_h, k, v)
    cfg.reduce_impl = b""
    print(f"case fca_hip_run {prec}", flush=True)
    tm = backend(prec).run(gh, cfg, acc_h)
    if prec == "f32":
        for k in ("h", "hmax", "tmax"):
            assert np.array_equal(getattr(go, k), getattr(gh, k)), k
    else:
        assert np.abs(go.h - gh.h).max() <= 1e-12
    for k, v in OFFSETS.items():
        ref = getattr(acc_o, k) + v
        assert getattr(acc_h, k) == pytest.approx(ref, rel=1e-9, abs=1e-9), k
    for k in ("t", "step", "cap_total", "dt_min_hits"):
        assert getattr(acc_h, k) == getattr(acc_o, k), k
    assert acc_h.dt_base == pytest.approx(acc_o.dt_base, rel=1e-6)
    assert acc_h.V == pytest.approx(acc_o.V, rel=1e-9)
    assert acc_h.V0 == acc_o.V0
    assert tm.wall_total > 0 and tm.t_d2h >= 0 and tm.t_h2d >= 0 and tm.t_init >= 0
    assert tm.t_flux == tm.t_limiter == tm.t_update == tm.t_dt == -1
    assert cfg.reduce_impl == b"wave"


def test_device_name_and_selection():
    b = backend("f32")
    name = b.device_name()
    assert name and "gfx" in name
    assert b.lib.fca_hip_select_device(0) == 0
    n = b.lib.fca_hip_select_device(1 << 20)
    assert n >= 0
    assert b.lib.fca_hip_select_device(0) == 0
