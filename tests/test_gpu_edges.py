"""Edge cases of the step against the oracle (GPU), to the bar of the parity
tests: FP32 bit for bit (h, hmax, tmax, dt sequence), FP64 within 1e-12 m.

* widths around the wavefront tile (58 output columns per 64 lanes), so the
  last tile is full, one column over, or one column short;
* grids without a LAND cell (all SEA, all WALL, SEA and WALL): no face is
  live, no depth moves, and every step is dt_max (model.c:150-152);
* a single cell, with the open and the closed boundary;
* deep water (tens of metres) and depths straddling tau_h;
* an all-LAND grid with walls, where the LAND row form covers the interior;
* LAND rows beyond the LAND form's range (a 2^21 m plateau, a 2^15 m column
  of water), which the walk sends through the general form.
"""
import numpy as np
import pytest

from floodca_amd import abi as A

from conftest import make_case
from test_gpu_parity import assert_parity, run_both

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nc", [57, 58, 59, 116, 117, 175])
@pytest.mark.parametrize("K", [4, 8])
def test_widths_around_the_tile(nc, K):
    a, g, cfg = make_case(23, nc, K=K, T_sim=1800.0, seed=nc + K, walls=True)
    print(f"case width nc={nc} K={K}", flush=True)
    assert_parity("f32", *run_both(a, g, cfg, "f32", strip=9))


@pytest.mark.parametrize("fill", ["sea", "wall", "sea+wall"])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_grid_without_land(fill, prec):
    a, g, cfg = make_case(19, 70, prec=prec, steps=12, T_sim=1e9, seed=3)
    if fill == "sea":
        g.mask[...] = A.FCA_SEA
    elif fill == "wall":
        g.mask[...] = A.FCA_WALL
    else:
        g.mask[...] = A.FCA_SEA
        g.mask[:, ::3] = A.FCA_WALL
    h_in = g.h.copy()
    print(f"case no land {fill} {prec}", flush=True)
    go, acc_o, log_o, gh, acc_g, log_h = run_both(a, g, cfg, prec)
    assert_parity(prec, go, acc_o, log_o, gh, acc_g, log_h)
    assert acc_g.step == 12
    assert np.array_equal(gh.h, h_in)
    assert all(r.dt == cfg.phys.dt_max for r in log_h)


@pytest.mark.parametrize("boundary", [A.FCA_BND_OPEN, A.FCA_BND_CLOSED])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_single_cell(boundary, prec):
    a, g, cfg = make_case(1, 1, prec=prec, boundary=boundary, T_sim=3600.0, seed=5, h0=0.2)
    print(f"case 1x1 boundary={boundary} {prec}", flush=True)
    go, acc_o, log_o, gh, acc_g, log_h = run_both(a, g, cfg, prec)
    assert_parity(prec, go, acc_o, log_o, gh, acc_g, log_h)
    assert acc_g.step > 0


@pytest.mark.parametrize("h0", [25.0, 1.0e-3, 0.9e-3, 1.1e-3])
def test_deep_and_threshold_depths(h0):
    a, g, cfg = make_case(37, 131, T_sim=900.0, seed=9, h0=h0, rday=0.3)
    print(f"case h0={h0}", flush=True)
    assert_parity("f32", *run_both(a, g, cfg, "f32"))


@pytest.mark.parametrize("K", [4, 8])
def test_all_land_with_walls(K):
    a, g, cfg = make_case(71, 300, K=K, T_sim=1200.0, seed=21, land_frac=1.0, walls=True)
    assert (g.mask == A.FCA_LAND).mean() > 0.9
    print(f"case all land K={K}", flush=True)
    assert_parity("f32", *run_both(a, g, cfg, "f32", strip=16))


@pytest.mark.parametrize("K", [4, 8])
def test_rows_beyond_the_land_form_range(K):
    """Levels past 2^20 m and a depth past 2^14 m on an all-LAND grid: the
    LAND form's fast face arithmetic assumes neither (its row test sends such
    rows to the general form), and the step stays the oracle's bit for bit."""
    a, g, cfg = make_case(48, 150, K=K, steps=40, T_sim=1e9, seed=40 + K, land_frac=1.0)
    g.mask[...] = A.FCA_LAND
    g.z[20:24, :] += np.float32(2.0 ** 21)
    g.h[33, 70] = np.float32(2.0 ** 15)
    g.hmax[33, 70] = g.h[33, 70]
    print(f"case beyond LAND range K={K}", flush=True)
    assert_parity("f32", *run_both(a, g, cfg, "f32", strip=24))
