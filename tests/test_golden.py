"""The oracle and the host emulation of the fused walk against the reference's
own outputs (tests/golden, written by the reference build): bit-exact h, hmax,
tmax and time-step sequence in both precisions."""
import numpy as np
import pytest

from conftest import GOLDEN_CASES, emu_run, load_golden
from oracle import Oracle

REL = 1e-9


def check(g, acc, t_seq, dtb_seq, exp, prec):
    for k in ("h", "hmax", "tmax"):
        assert np.array_equal(getattr(g, k), exp[k]), f"{k}: max|d| = " \
            f"{np.abs(getattr(g, k).astype(float) - exp[k]).max()}"
    log = exp["log"]
    assert acc.step == exp["acc"]["step"] == len(log)
    assert list(t_seq) == list(log[:, 0])
    dtype = np.float32 if prec == "f32" else np.float64
    assert np.array_equal(np.asarray(dtb_seq, dtype=dtype), log[:, 2].astype(dtype))
    e = exp["acc"]
    assert acc.t == e["t"]
    assert acc.cap_active == e["cap_active"] and acc.cap_total == e["cap_total"]
    assert acc.dt_min_hits == e["dt_min_hits"]
    for k in ("V", "V_rain", "V_sea", "V_bnd", "clamp_err"):
        assert getattr(acc, k) == pytest.approx(e[k], rel=REL, abs=REL * max(1.0, abs(e[k])))


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_oracle_reproduces_reference(name, prec):
    a, g, cfg, exp = load_golden(name, prec)
    o = Oracle(prec)
    acc, log = o.run(g, cfg, log_cap=100000)
    check(g, acc, [r.t for r in log], [r.dt_base for r in log], exp, prec)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("strip", [0, 7])
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_fused_walk_emulation_reproduces_reference(name, prec, strip):
    """The fused kernel's dataflow (fca_walk.h), run on 64 emulated lanes."""
    a, g, cfg, exp = load_golden(name, prec)
    acc = Oracle(prec).acc_init(g, cfg)
    acc, log = emu_run(g, cfg, acc, prec, strip_rows=strip)
    check(g, acc, [r.t for r in log], [r.dt_base for r in log], exp, prec)
