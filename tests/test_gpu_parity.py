"""HIP step vs the oracle on identical seeded inputs (GPU).

FP32: bit-exact h, hmax, tmax and dt sequence (the kernels keep the reference's
operation order with contraction off). FP64: max|dh| <= 1e-12 m, the bar of the
reference's own double-precision comparison. Balance terms (FP64 sums in a
different order) within 1e-9 relative; cap counts exact.
"""
import numpy as np
import pytest

from floodca_amd import abi as A
from floodca_amd.hip import Sim
from oracle import Oracle

from conftest import make_case

CASES = [
    # nr, nc, K, boundary, strip, kernel, prec
    (61, 47, 4, A.FCA_BND_OPEN, 0, A.FCA_HIP_FUSED, "f32"),
    (61, 47, 8, A.FCA_BND_OPEN, 0, A.FCA_HIP_FUSED, "f32"),
    (61, 47, 4, A.FCA_BND_CLOSED, 0, A.FCA_HIP_FUSED, "f32"),
    (61, 130, 8, A.FCA_BND_CLOSED, 16, A.FCA_HIP_FUSED, "f32"),
    (97, 257, 4, A.FCA_BND_OPEN, 8, A.FCA_HIP_FUSED, "f32"),
    (7, 5, 4, A.FCA_BND_OPEN, 3, A.FCA_HIP_FUSED, "f32"),
    (1, 9, 8, A.FCA_BND_OPEN, 4, A.FCA_HIP_FUSED, "f32"),
    (40, 1, 4, A.FCA_BND_OPEN, 7, A.FCA_HIP_FUSED, "f32"),
    (61, 47, 4, A.FCA_BND_OPEN, 0, A.FCA_HIP_SPLIT, "f32"),
    (61, 130, 8, A.FCA_BND_OPEN, 0, A.FCA_HIP_SPLIT, "f32"),
    (61, 47, 4, A.FCA_BND_OPEN, 0, A.FCA_HIP_FUSED, "f64"),
    (61, 130, 8, A.FCA_BND_CLOSED, 16, A.FCA_HIP_FUSED, "f64"),
    (61, 47, 8, A.FCA_BND_OPEN, 0, A.FCA_HIP_SPLIT, "f64"),
]


def _ids(c):
    nr, nc, K, b, strip, kern, prec = c
    return f"{nr}x{nc}-K{K}-{'open' if b == 0 else 'closed'}-s{strip}-" \
           f"{'fused' if kern == 0 else 'split'}-{prec}"


def run_both(a, g, cfg, prec, kernel=A.FCA_HIP_FUSED, strip=0, cap=100000, cols_per_lane=0):
    o = Oracle(prec)
    go, gh = g.copy(), g.copy()
    acc_o, log_o = o.run(go, cfg, log_cap=cap)
    acc_h = o.acc_init(gh, cfg)
    with Sim(gh, cfg, acc_h, prec=prec, kernel=kernel, strip_rows=strip,
             record_steps=cap, cols_per_lane=cols_per_lane) as s:
        s.run()
        acc_g = s.read()
        log_h = s.steplog(cap)
    return go, acc_o, log_o, gh, acc_g, log_h


def assert_parity(prec, go, acc_o, log_o, gh, acc_g, log_h):
    assert acc_g.step == acc_o.step
    assert len(log_h) == len(log_o)
    if prec == "f32":
        for k in ("h", "hmax", "tmax"):
            d = np.abs(getattr(go, k).astype(np.float64) - getattr(gh, k))
            assert np.array_equal(getattr(go, k), getattr(gh, k)), f"{k}: max|d| = {d.max()}"
        assert [r.dt for r in log_o] == [r.dt for r in log_h]
        assert [r.dt_base for r in log_o] == [r.dt_base for r in log_h]
        assert [r.t for r in log_o] == [r.t for r in log_h]
        assert acc_g.cap_active == acc_o.cap_active
    else:
        assert np.abs(go.h - gh.h).max() <= 1e-12
        assert np.abs(go.hmax - gh.hmax).max() <= 1e-12
    assert [r.hit_event for r in log_o] == [r.hit_event for r in log_h]
    assert acc_g.cap_total == acc_o.cap_total
    assert acc_g.dt_min_hits == acc_o.dt_min_hits
    assert acc_g.t == pytest.approx(acc_o.t, rel=1e-15)
    for k in ("V", "V_rain", "V_sea", "V_bnd", "clamp_err"):
        ref = getattr(acc_o, k)
        assert getattr(acc_g, k) == pytest.approx(ref, rel=1e-9, abs=1e-9 * max(abs(ref), 1.0))


@pytest.mark.gpu
@pytest.mark.parametrize("c", CASES, ids=_ids)
def test_step_parity(c):
    nr, nc, K, b, strip, kern, prec = c
    a, g, cfg = make_case(nr, nc, prec=prec, K=K, boundary=b, T_sim=3600.0)
    print(f"case {_ids(c)}", flush=True)
    assert_parity(prec, *run_both(a, g, cfg, prec, kern, strip))


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_fixed_steps_triangle_events(prec):
    """Fixed step count, triangle hyetograph, out_every events and U = 3."""
    a, g, cfg = make_case(53, 71, prec=prec, K=8, hyeto=A.FCA_HYETO_TRIANGLE, D_h=2.0,
                          phi=0.25, out_every=900.0, T_sim=1e9, steps=400, U=3, rday=40.0)
    print(f"case triangle/U3/{prec}", flush=True)
    assert_parity(prec, *run_both(a, g, cfg, prec, A.FCA_HIP_FUSED, 16))


# the walk with two columns per work-item (tiles of 122 columns): grids with
# inner column groups of that width (LAND form), narrow and degenerate ones
CPL2_CASES = [
    # nr, nc, K, boundary, strip
    (61, 47, 4, A.FCA_BND_OPEN, 0),
    (61, 300, 4, A.FCA_BND_CLOSED, 16),
    (97, 400, 8, A.FCA_BND_OPEN, 8),
    (45, 260, 8, A.FCA_BND_CLOSED, 0),
    (7, 5, 4, A.FCA_BND_OPEN, 3),
    (40, 1, 8, A.FCA_BND_OPEN, 7),
]


@pytest.mark.gpu
@pytest.mark.parametrize("c", CPL2_CASES, ids=lambda c: f"{c[0]}x{c[1]}-K{c[2]}-b{c[3]}-s{c[4]}")
def test_step_parity_two_columns_per_lane(c):
    nr, nc, K, b, strip = c
    a, g, cfg = make_case(nr, nc, prec="f32", K=K, boundary=b, T_sim=3600.0, land_frac=0.9,
                          h0=0.05)
    print(f"case cpl2 {nr}x{nc} K={K} b={b} strip={strip}", flush=True)
    assert_parity("f32", *run_both(a, g, cfg, "f32", A.FCA_HIP_FUSED, strip, cols_per_lane=2))
