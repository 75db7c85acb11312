"""The benchmark's own 16384^2 grid on the GPU (FP32, K=4 and K=8; 20 steps
of 50 mm/h rain from a dry start, the first bench steps), where the oracle
would take minutes a step, checked through properties that do not depend on
size:

* the same input gives the same bits across tile shapes: the fused walk with
  strips of 96 and of 64 rows, and with two columns per work-item (tiles of
  122 columns), ends with identical h, hmax, tmax and time-step sequences
  (no tile reads what another writes in the same step);
* the split kernels, the independent second implementation (face arrays, one
  launch per phase), end with the same bits too;
* every depth stays >= 0 and the mass-balance identity (diag.c:5-14) holds
  within 1e-6 of the water added.
"""
import types

import numpy as np
import pytest

from floodca_amd import abi as A
from floodca_amd import synth
from floodca_amd.hip import Sim

pytestmark = pytest.mark.gpu

SIZE, STEPS = 16384, 20


@pytest.fixture(scope="module")
def grid():
    a = A.abi("f32")
    print(f"case headline {SIZE}^2: generating the grid", flush=True)
    return a, synth.make_grid(SIZE, SIZE, seed=20261203, dtype=a.dtype)


def _run(a, g, K, **kw):
    cfg = A.default_cfg(a)
    A.set_dx(a, cfg, g.dx)
    cfg.phys.K = cfg.neigh = K
    cfg.hyeto = A.FCA_HYETO_UNIFORM
    cfg.T_sim = 1e9
    cfg.steps = STEPS
    acc = a.FcaAcc()
    acc.force_day = -1
    out = types.SimpleNamespace(h=np.empty_like(g.h), hmax=np.empty_like(g.h),
                                tmax=np.empty_like(g.h))
    with Sim(g, cfg, acc, prec="f32", record_steps=STEPS, **kw) as s:
        s.run()
        acc = s.read(out)
        log = s.steplog(STEPS)
    return out, acc, [(r.dt, r.dt_base) for r in log]


@pytest.mark.parametrize("K", [4, 8])
def test_headline_grid_same_bits_across_tiles_and_kernels(grid, K):
    a, g = grid
    print(f"case headline {SIZE}^2 K={K}: fused strips of 96", flush=True)
    o96, acc96, log96 = _run(a, g, K, strip_rows=96)
    print(f"case headline {SIZE}^2 K={K}: fused strips of 64", flush=True)
    o64, acc64, log64 = _run(a, g, K, strip_rows=64)
    print(f"case headline {SIZE}^2 K={K}: fused, two columns per work-item", flush=True)
    o2, acc2, log2 = _run(a, g, K, cols_per_lane=2)
    print(f"case headline {SIZE}^2 K={K}: split kernels", flush=True)
    osp, accsp, logsp = _run(a, g, K, kernel=A.FCA_HIP_SPLIT)
    assert acc96.step == acc64.step == acc2.step == accsp.step == STEPS
    assert log96 == log64 == log2 == logsp
    for k in ("h", "hmax", "tmax"):
        assert np.array_equal(getattr(o96, k), getattr(o64, k)), k
        assert np.array_equal(getattr(o96, k), getattr(o2, k)), k
        assert np.array_equal(getattr(o96, k), getattr(osp, k)), k
    land = g.mask == A.FCA_LAND
    assert o96.h.min() >= 0.0
    assert np.all(o96.h[~land] == g.h[~land])   # SEA and WALL depths never change
    assert np.count_nonzero(o96.h[land] >= 1e-3) > 0.99 * land.sum()
    for acc in (acc96, acc64, acc2, accsp):
        e = acc.V + acc.V_sea + acc.V_bnd - acc.V_rain - acc.V0 - acc.clamp_err
        assert abs(e) <= 1e-6 * (acc.V_rain + acc.V0)
        assert acc.cap_active == acc96.cap_active and acc.cap_total == acc96.cap_total
    for k in ("V", "V_rain", "V_sea", "V_bnd", "clamp_err"):
        assert getattr(acc64, k) == pytest.approx(getattr(acc96, k), rel=1e-12, abs=1e-6), k
        assert getattr(acc2, k) == pytest.approx(getattr(acc96, k), rel=1e-12, abs=1e-6), k
        assert getattr(accsp, k) == pytest.approx(getattr(acc96, k), rel=1e-9, abs=1e-6), k
    # the faces carried water: under uniform rain alone every LAND depth
    # would be the same
    hl = o96.h[land]
    assert np.count_nonzero(hl != np.median(hl)) > 0.1 * hl.size
    assert acc96.V_sea > 0.0
