"""Multi-day forcing on the device against the oracle (GPU): at each day
boundary the kernel pauses, the host reloads that day's rain field (scaled by
lambda) and rebuilds cr = C*Rday on the device, and the run resumes
(model.c:338-358); a missing day file ends the run at T_sim. The oracle's
forcing path is itself checked against the reference's own CSV loader and
run loop (tests/test_oracle_vs_ref.py::test_two_day_forcing_csv). FP32 bit
for bit (h, hmax, tmax, dt sequence, accumulators), FP64 within 1e-12 m."""
import numpy as np
import pytest

from floodca_amd import abi as A

from conftest import make_case
from test_gpu_parity import assert_parity, run_both

pytestmark = pytest.mark.gpu


def _forcing(cfg, tmp_path, shape, days_present, days_listed, seed=1):
    rng = np.random.default_rng(seed)
    for d in range(days_present):
        (rng.random(shape) * (0.3 + d)).astype(np.float32).tofile(tmp_path / f"r{d}.bin")
    cfg.n_force_days = days_listed
    for d in range(days_listed):
        cfg.force_files[d].value = str(tmp_path / f"r{d}.bin").encode()


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("K", [4, 8])
def test_day_reloads_match_the_oracle(prec, K, tmp_path):
    a, g, cfg = make_case(29, 23, prec=prec, K=K, T_sim=86400.0 * 3, seed=13)
    cfg.hyeto = A.FCA_HYETO_UNIFORM
    cfg.lambda_ = 0.8
    _forcing(cfg, tmp_path, g.shape, days_present=3, days_listed=3)
    print(f"case forcing 3 days K={K} {prec}", flush=True)
    go, acc_o, log_o, gh, acc_g, log_h = run_both(a, g, cfg, prec)
    assert_parity(prec, go, acc_o, log_o, gh, acc_g, log_h)
    assert acc_o.force_day == acc_g.force_day == 2
    assert acc_g.t == 86400.0 * 3


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_missing_day_ends_the_run_at_t_sim(prec, tmp_path):
    a, g, cfg = make_case(31, 17, prec=prec, T_sim=86400.0 * 3, seed=14)
    cfg.hyeto = A.FCA_HYETO_UNIFORM
    _forcing(cfg, tmp_path, g.shape, days_present=2, days_listed=3)
    print(f"case forcing missing day {prec}", flush=True)
    go, acc_o, log_o, gh, acc_g, log_h = run_both(a, g, cfg, prec)
    assert_parity(prec, go, acc_o, log_o, gh, acc_g, log_h)
    assert acc_o.force_day == acc_g.force_day == 1
    assert acc_o.t == acc_g.t == 86400.0 * 3
