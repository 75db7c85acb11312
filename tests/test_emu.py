"""The fused kernel's dataflow (floodca_amd/csrc/fca_walk.h) on 64 emulated
lanes against the oracle, bit for bit (CPU): strips and column groups cut
anywhere, degenerate grids, both neighbourhoods and boundaries, both
precisions. The GPU runs the same template with DPP shifts in place of the
emulated ones."""
import numpy as np
import pytest

from conftest import emu_run, make_case
from floodca_amd import abi as A
from oracle import Oracle

SHAPES = [(61, 130, 16), (97, 121, 5), (7, 5, 3), (1, 9, 4), (40, 1, 7), (2, 2, 1), (3, 64, 2)]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("K", [4, 8])
@pytest.mark.parametrize("bnd", [A.FCA_BND_OPEN, A.FCA_BND_CLOSED])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}s{s[2]}")
def test_emulated_walk_matches_oracle(prec, K, bnd, shape):
    nr, nc, strip = shape
    a, g, cfg = make_case(nr, nc, prec=prec, K=K, boundary=bnd, T_sim=2400.0, walls=nr > 10,
                          seed=nr * 7 + nc)
    o = Oracle(prec)
    go, ge = g.copy(), g.copy()
    acc_o, log_o = o.run(go, cfg, log_cap=100000)
    acc_e, log_e = emu_run(ge, cfg, o.acc_init(ge, cfg), prec, strip_rows=strip)
    for k in ("h", "hmax", "tmax"):
        assert np.array_equal(getattr(go, k), getattr(ge, k)), k
    assert [r.dt for r in log_o] == [r.dt for r in log_e]
    assert acc_o.cap_active == acc_e.cap_active and acc_o.cap_total == acc_e.cap_total
    assert acc_o.V_rain == pytest.approx(acc_e.V_rain, rel=1e-12)
