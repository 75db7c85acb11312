"""The fused kernel's dataflow (floodca_amd/csrc/fca_walk.h) on 64 emulated
lanes (one column per work-item) and on 128 (two), against the oracle, bit
for bit (CPU): strips and column groups cut anywhere, degenerate grids, both
neighbourhoods and boundaries, both precisions. The GPU runs the same
template with DPP shifts in place of the emulated ones."""
import numpy as np
import pytest

from conftest import emu_run, make_case
from floodca_amd import abi as A
from oracle import Oracle

SHAPES = [(61, 130, 16), (97, 121, 5), (7, 5, 3), (1, 9, 4), (40, 1, 7), (2, 2, 1), (3, 64, 2)]


@pytest.mark.parametrize("width", [64, 128])
@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("K", [4, 8])
@pytest.mark.parametrize("bnd", [A.FCA_BND_OPEN, A.FCA_BND_CLOSED])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}s{s[2]}")
def test_emulated_walk_matches_oracle(prec, K, bnd, shape, width):
    nr, nc, strip = shape
    a, g, cfg = make_case(nr, nc, prec=prec, K=K, boundary=bnd, T_sim=2400.0, walls=nr > 10,
                          seed=nr * 7 + nc)
    o = Oracle(prec)
    go, ge = g.copy(), g.copy()
    acc_o, log_o = o.run(go, cfg, log_cap=100000)
    acc_e, log_e = emu_run(ge, cfg, o.acc_init(ge, cfg), prec, strip_rows=strip, width=width)
    for k in ("h", "hmax", "tmax"):
        assert np.array_equal(getattr(go, k), getattr(ge, k)), k
    assert [r.dt for r in log_o] == [r.dt for r in log_e]
    assert acc_o.cap_active == acc_e.cap_active and acc_o.cap_total == acc_e.cap_total
    assert acc_o.V_rain == pytest.approx(acc_e.V_rain, rel=1e-12)


# (wavefront width, shape): each wide enough for an inner column group of
# that width (the LAND form's condition)
LAND_SHAPES = [(64, (70, 250, 32)), (64, (45, 180, 9)), (128, (70, 250, 32)),
               (128, (45, 300, 9))]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("K", [4, 8])
@pytest.mark.parametrize("bnd", [A.FCA_BND_OPEN, A.FCA_BND_CLOSED])
@pytest.mark.parametrize("width,shape", LAND_SHAPES,
                         ids=lambda s: f"{s[0]}x{s[1]}s{s[2]}" if isinstance(s, tuple) else s)
def test_land_form_matches_oracle(prec, K, bnd, shape, width):
    """Mostly-LAND grids wide enough for inner column groups, so that rows take
    the walk's LAND form next to general rows (coast, walls, edges); the
    result is the oracle's bit for bit, and the same with the LAND form off."""
    nr, nc, strip = shape
    a, g, cfg = make_case(nr, nc, prec=prec, K=K, boundary=bnd, T_sim=1800.0, land_frac=0.93,
                          seed=nr + nc, h0=0.08)
    rng = np.random.default_rng(nc)
    w = (rng.random(g.shape) < 0.002) & (g.mask == A.FCA_LAND)
    g.mask[w] = A.FCA_WALL
    o = Oracle(prec)
    go, ge, gg = g.copy(), g.copy(), g.copy()
    acc_o, log_o = o.run(go, cfg, log_cap=100000)
    forms = []
    acc_e, log_e = emu_run(ge, cfg, o.acc_init(ge, cfg), prec, strip_rows=strip, forms=forms,
                           width=width)
    assert forms[1] > 0 and forms[0] > 0, forms
    acc_g, log_g = emu_run(gg, cfg, o.acc_init(gg, cfg), prec, strip_rows=strip,
                           land_form=False, width=width)
    for k in ("h", "hmax", "tmax"):
        assert np.array_equal(getattr(go, k), getattr(ge, k)), k
        assert np.array_equal(getattr(gg, k), getattr(ge, k)), k
    assert [r.dt for r in log_o] == [r.dt for r in log_e] == [r.dt for r in log_g]
    assert acc_o.cap_active == acc_e.cap_active == acc_g.cap_active
    assert acc_o.V_rain == pytest.approx(acc_e.V_rain, rel=1e-12)
    assert acc_o.V_sea == pytest.approx(acc_e.V_sea, rel=1e-12, abs=1e-9)
    assert acc_o.V_bnd == pytest.approx(acc_e.V_bnd, rel=1e-12, abs=1e-9)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("K", [4, 8])
def test_all_three_row_forms_match_oracle(prec, K):
    """Half sea, half land on a wide grid: rows walked in the LAND form, the
    no-LAND form (all-sea wavefronts) and the general form all occur, and the
    run is the oracle's bit for bit."""
    a, g, cfg = make_case(64, 300, prec=prec, K=K, T_sim=1500.0, land_frac=0.5, seed=5,
                          h0=0.06)
    o = Oracle(prec)
    go, ge = g.copy(), g.copy()
    acc_o, log_o = o.run(go, cfg, log_cap=100000)
    forms = []
    acc_e, log_e = emu_run(ge, cfg, o.acc_init(ge, cfg), prec, strip_rows=24, forms=forms)
    assert min(forms) > 0, forms
    for k in ("h", "hmax", "tmax"):
        assert np.array_equal(getattr(go, k), getattr(ge, k)), k
    assert [r.dt for r in log_o] == [r.dt for r in log_e]
    assert acc_o.cap_active == acc_e.cap_active
    assert acc_o.V_sea == pytest.approx(acc_e.V_sea, rel=1e-12, abs=1e-9)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("K", [4, 8])
@pytest.mark.parametrize("bnd", [A.FCA_BND_OPEN, A.FCA_BND_CLOSED])
def test_quiet_tiles_skip_without_changing_results(prec, K, bnd):
    """A sparse start (dry grid, one inflow cell, a second short rain): most
    tiles are quiet, are skipped, and the run is the oracle's bit for bit,
    and the same with skipping off."""
    from floodca_amd import synth
    a, g, cfg = make_case(90, 250, prec=prec, K=K, boundary=bnd, T_sim=3000.0, h0=0.0,
                          seed=17, land_frac=0.8)
    synth.point_inflow(g, peak_m3s=50.0, dur_h=1.0)
    cfg.hyeto, cfg.D_h, cfg.phi = A.FCA_HYETO_TRIANGLE, 1.0, 0.25
    o = Oracle(prec)
    go, ge, gn = g.copy(), g.copy(), g.copy()
    acc_o, log_o = o.run(go, cfg, log_cap=100000)
    sk = []
    acc_e, log_e = emu_run(ge, cfg, o.acc_init(ge, cfg), prec, strip_rows=8, skipped=sk)
    acc_n, log_n = emu_run(gn, cfg, o.acc_init(gn, cfg), prec, strip_rows=8, skip=False)
    assert sk[0] > 0
    for k in ("h", "hmax", "tmax"):
        assert np.array_equal(getattr(go, k), getattr(ge, k)), k
        assert np.array_equal(getattr(gn, k), getattr(ge, k)), k
    assert [r.dt for r in log_o] == [r.dt for r in log_e] == [r.dt for r in log_n]
    assert acc_o.cap_active == acc_e.cap_active
    assert acc_e.V_rain == acc_n.V_rain and acc_e.V_bnd == acc_n.V_bnd


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("days_present", [3, 2])
def test_forcing_days_match_oracle(prec, days_present, tmp_path):
    """Day-boundary Rday reloads through the emulated controller (the pause,
    the reload scaled by lambda, cr rebuilt) and a missing day file, which
    ends the run at T_sim."""
    a, g, cfg = make_case(29, 23, prec=prec, K=4, T_sim=86400.0 * 3, seed=13)
    cfg.hyeto = A.FCA_HYETO_UNIFORM
    cfg.lambda_ = 0.8
    rng = np.random.default_rng(1)
    for d in range(days_present):
        (rng.random(g.shape) * (0.3 + d)).astype(np.float32).tofile(tmp_path / f"r{d}.bin")
    cfg.n_force_days = 3
    for d in range(3):
        cfg.force_files[d].value = str(tmp_path / f"r{d}.bin").encode()
    o = Oracle(prec)
    go, ge = g.copy(), g.copy()
    acc_o, log_o = o.run(go, cfg, log_cap=100000)
    acc_e, log_e = emu_run(ge, cfg, o.acc_init(ge, cfg), prec, strip_rows=7)
    for k in ("h", "hmax", "tmax"):
        assert np.array_equal(getattr(go, k), getattr(ge, k)), k
    assert [r.dt for r in log_o] == [r.dt for r in log_e]
    assert acc_o.step == acc_e.step and acc_o.t == acc_e.t == 86400.0 * 3
    assert acc_o.force_day == acc_e.force_day == days_present - 1


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("hyeto", [A.FCA_HYETO_UNIFORM, A.FCA_HYETO_TRIANGLE])
def test_hmax_below_the_initial_depth(prec, hyeto):
    """The walk stores hmax and tmax only when a cell stops rising, and the
    first step after an upload stores none (a caller's h may exceed its
    hmax). Warm starts whose hmax lies below h on a third of the cells --
    wet and dry, on tiles that go quiet and tiles that stay busy -- and whose
    tmax is nonzero end with the oracle's hmax and tmax bit for bit, with and
    without quiet-tile skipping."""
    a, g, cfg = make_case(90, 250, prec=prec, hyeto=hyeto, T_sim=3000.0, seed=5, h0=0.0,
                          D_h=0.5, phi=0.3, rday=0.0005)
    rng = np.random.default_rng(9)
    land = g.mask == A.FCA_LAND
    cols = np.arange(g.shape[1])[None, :]
    wet = land & (rng.random(g.shape) < 0.3) & (cols >= 232)      # the fifth column group
    g.h[wet] = rng.uniform(0.0, 0.2, wet.sum()).astype(g.h.dtype)
    g.hmax[...] = g.h
    low = land & (rng.random(g.shape) < 0.33)
    g.hmax[low] = (g.h[low] * rng.uniform(0.0, 0.9, low.sum())).astype(g.h.dtype)
    g.tmax[land] = rng.uniform(0.0, 100.0, land.sum()).astype(g.h.dtype)
    # no rain on the second column group (columns 58-115, its neighbours dry),
    # and there depths below tau_h above their hmax: the reference sets
    # hmax = h and tmax on the first step although h does not move, and the
    # tile must not go quiet before the walk stores them
    zone = (cols >= 58) & (cols < 116) & np.ones(g.shape, dtype=bool)
    g.C[zone] = 0
    damp = zone & land & (rng.random(g.shape) < 0.2)
    assert damp.sum() > 100
    g.h[damp] = a.dtype.type(3e-4)
    g.hmax[damp] = a.dtype.type(1e-4)
    o = Oracle(prec)
    go = g.copy()
    acc_o, _ = o.run(go, cfg)
    for skip in (True, False):
        ge = g.copy()
        skipped = []
        emu_run(ge, cfg, o.acc_init(ge, cfg), prec, strip_rows=8, skip=skip, skipped=skipped)
        for k in ("h", "hmax", "tmax"):
            assert np.array_equal(getattr(go, k), getattr(ge, k)), (k, skip)
        if skip and hyeto == A.FCA_HYETO_TRIANGLE:
            assert skipped[0] > 0  # a 0.5 mm storm over 30 min: dry tiles go quiet


@pytest.mark.parametrize("K", [4, 8])
def test_out_of_range_rows_take_the_general_form(K):
    """An all-LAND grid whose middle rows stand on a 2^21 m plateau (|z + h|
    beyond the LAND form's 2^20) and hold one 2^15 m column of water (|h|
    beyond 2^14): those rows, and the four after each, take the general form
    (more general rows than on the same grid without them), and the result is
    the oracle's bit for bit."""
    rows = []
    for tall in (False, True):
        a, g, cfg = make_case(48, 150, K=K, steps=40, T_sim=1e9, seed=40 + K, land_frac=1.0)
        g.mask[...] = A.FCA_LAND
        if tall:
            g.z[20:24, :] += np.float32(2.0 ** 21)
            g.h[33, 70] = np.float32(2.0 ** 15)
            g.hmax[33, 70] = g.h[33, 70]
        o = Oracle("f32")
        go, ge = g.copy(), g.copy()
        acc_o, log_o = o.run(go, cfg, log_cap=100000)
        forms = []
        acc_e, log_e = emu_run(ge, cfg, o.acc_init(ge, cfg), "f32", strip_rows=24, forms=forms)
        for k in ("h", "hmax", "tmax"):
            assert np.array_equal(getattr(go, k), getattr(ge, k)), (tall, k)
        assert [r.dt for r in log_o] == [r.dt for r in log_e]
        assert acc_o.cap_active == acc_e.cap_active
        rows.append(forms)
    assert rows[1][0] > rows[0][0] and rows[1][1] < rows[0][1], rows
