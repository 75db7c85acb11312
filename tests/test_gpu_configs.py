"""The plain method at the benchmark configurations' own sizes, HIP vs the
oracle on identical seeded inputs (GPU).

* 512^2 at 90 m, the full 6 h of 50 mm/h uniform rain from a dry start: FP32
  bit-exact h, hmax, tmax and time-step sequence (K=4 and K=8); FP64
  max|dh| <= 1e-12 m.
  At this width most rows take the fused walk's LAND form, so this is the
  device arithmetic of the fast path checked over a whole flood.
* 4096^2, 12 steps from a 5 cm warm start (every LAND face carries water
  from the first step): the same bars, plus the relative volume drift of the
  run against the oracle's within 1e-6.
* Sparse start (one inland inflow cell, triangular hydrograph, no rain) on
  1024^2 for the first hour: quiet-tile skipping against the oracle, bit for bit.

The oracle is the C restatement of the reference's step, pinned against the
reference's own build by the CPU suite.
"""
import numpy as np
import pytest

from floodca_amd import abi as A
from floodca_amd import synth

from test_gpu_parity import assert_parity, run_both

pytestmark = pytest.mark.gpu


def _case(size, prec, h0=0.0, T_sim=21600.0, steps=0):
    a = A.abi(prec)
    g = synth.make_grid(size, size, seed=20261203, dtype=a.dtype, h0=h0)
    cfg = A.default_cfg(a)
    A.set_dx(a, cfg, g.dx)
    cfg.hyeto = A.FCA_HYETO_UNIFORM
    cfg.T_sim = T_sim
    cfg.steps = steps
    return a, g, cfg


@pytest.mark.parametrize("prec,K", [("f32", 4), ("f64", 4), ("f32", 8)])
def test_512_full_six_hours(prec, K):
    a, g, cfg = _case(512, prec)
    cfg.phys.K = cfg.neigh = K
    print(f"case 512^2 6 h {prec} K={K}", flush=True)
    go, acc_o, log_o, gh, acc_g, log_h = run_both(a, g, cfg, prec)
    assert acc_o.t == pytest.approx(21600.0)
    assert len(log_o) > 500
    assert_parity(prec, go, acc_o, log_o, gh, acc_g, log_h)


def _rel_drift(acc):
    return (acc.V + acc.V_sea + acc.V_bnd - acc.V_rain - acc.V0 - acc.clamp_err) / \
        max(acc.V_rain + acc.V0, 1e-30)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_4096_warm_start(prec):
    a, g, cfg = _case(4096, prec, h0=0.05, T_sim=1e9, steps=12)
    print(f"case 4096^2 12 steps {prec}", flush=True)
    go, acc_o, log_o, gh, acc_g, log_h = run_both(a, g, cfg, prec, cap=64)
    assert acc_g.step == 12
    assert_parity(prec, go, acc_o, log_o, gh, acc_g, log_h)
    assert abs(_rel_drift(acc_g) - _rel_drift(acc_o)) <= 1e-6


def test_sparse_start_skipping_matches_the_oracle():
    a = A.abi("f32")
    g = synth.make_grid(1024, 1024, seed=20261205, dtype=a.dtype)
    synth.point_inflow(g, peak_m3s=500.0, dur_h=2.0)
    cfg = A.default_cfg(a)
    A.set_dx(a, cfg, g.dx)
    cfg.hyeto, cfg.D_h, cfg.phi = A.FCA_HYETO_TRIANGLE, 2.0, 0.25
    cfg.T_sim = 3600.0
    print("case sparse start 1024^2 1 h", flush=True)
    go, acc_o, log_o, gh, acc_g, log_h = run_both(a, g, cfg, "f32")
    assert_parity("f32", go, acc_o, log_o, gh, acc_g, log_h)
    wet = np.count_nonzero(gh.hmax >= 1e-3)
    assert 0 < wet < 0.05 * g.z.size
