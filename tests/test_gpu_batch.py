"""The batched session against single-grid runs (GPU): every scenario of a
batch ends bit for bit where it ends when run alone (h, hmax, tmax, the step
log, the accumulators), whatever its size, configuration and length."""
import numpy as np
import pytest

from conftest import make_case
from floodca_amd import abi as A
from floodca_amd.hip import Batch, Sim
from oracle import Oracle

CAP = 4000


def scenarios(prec, K):
    """Uneven grids and configurations, all with neighbourhood K."""
    specs = [
        dict(nr=61, nc=130, T_sim=3000.0, seed=1),
        dict(nr=7, nc=5, T_sim=1800.0, seed=2, h0=0.2),
        dict(nr=200, nc=333, T_sim=2400.0, seed=3, hyeto=A.FCA_HYETO_TRIANGLE, D_h=0.5,
             phi=0.25, rday=40.0),
        dict(nr=97, nc=64, steps=120, T_sim=1e9, seed=4, U=3, walls=True),
        dict(nr=1, nc=70, T_sim=1200.0, seed=5),
        dict(nr=150, nc=150, T_sim=5400.0, seed=6, boundary=A.FCA_BND_CLOSED, out_every=900.0),
    ]
    out = []
    for sp in specs:
        sp = dict(sp)
        nr, nc = sp.pop("nr"), sp.pop("nc")
        out.append(make_case(nr, nc, prec=prec, K=K, **sp)[1:])
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("prec,K,cpl", [("f32", 4, 1), ("f32", 8, 1), ("f64", 4, 1),
                                        ("f64", 8, 1), ("f32", 4, 2), ("f32", 8, 2)])
def test_batch_equals_each_scenario_alone(prec, K, cpl):
    """cpl: columns per work-item of the walk, in the batch and alone."""
    o = Oracle(prec)
    cases = scenarios(prec, K)
    alone = []
    for g, cfg in cases:
        ga = g.copy()
        with Sim(ga, cfg, o.acc_init(ga, cfg), prec=prec, strip_rows=16, record_steps=CAP,
                 cols_per_lane=cpl) as s:
            s.run()
            acc = s.read()
            log = s.steplog(CAP)
        alone.append((ga, acc, log))
    grids = [g.copy() for g, _ in cases]
    accs = [o.acc_init(g, cfg) for g, (_, cfg) in zip(grids, cases)]
    with Batch(grids, [cfg for _, cfg in cases], accs, prec=prec, strip_rows=16,
               record_steps=CAP, cols_per_lane=cpl) as b:
        b.run()
        for i, (ga, acc_a, log_a) in enumerate(alone):
            acc_b = b.read(i)
            log_b = b.steplog(i, CAP)
            gb = grids[i]
            for k in ("h", "hmax", "tmax"):
                assert np.array_equal(getattr(ga, k), getattr(gb, k)), (i, k)
            assert len(log_a) == len(log_b) == acc_a.step, i
            assert [(r.t, r.dt) for r in log_a] == [(r.t, r.dt) for r in log_b], i
            for f in ("t", "step", "dt_base", "V", "V_rain", "V_sea", "V_bnd", "clamp_err",
                      "cap_active", "cap_total", "dt_min_hits"):
                assert getattr(acc_a, f) == getattr(acc_b, f), (i, f)
            fl, md = b.flooded(i)
            land = gb.mask == A.FCA_LAND
            hm = gb.hmax.astype(np.float64)
            assert fl == [int(np.sum(land & (hm >= t))) for t in (0.10, 0.25, 0.50)], i
            assert md == max(0.0, float(hm[land].max()) if land.any() else 0.0), i


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_graph_replays_equal_launches_issued_one_by_one(prec):
    """Runs replay captured graphs of 30 step launches between controller
    read-backs; the same runs with every launch issued on its own
    (graph_batch=1) end identically, single grid and batch alike."""
    o = Oracle(prec)
    cases = scenarios(prec, 4)[:4]
    res = {}
    for gb in (1, 0, 12):
        out = []
        for g, cfg in cases:
            ga = g.copy()
            with Sim(ga, cfg, o.acc_init(ga, cfg), prec=prec, strip_rows=16,
                     graph_batch=gb) as s:
                s.run()
                acc = s.read()
            out.append((ga.h.copy(), ga.hmax.copy(), acc.step, acc.t, acc.V_rain))
        grids = [g.copy() for g, _ in cases]
        with Batch(grids, [c for _, c in cases], [o.acc_init(g, c) for g, (_, c) in
                                                   zip(grids, cases)],
                   prec=prec, strip_rows=16, graph_batch=gb) as b:
            b.run()
            for i in range(len(cases)):
                acc = b.read(i)
                out.append((grids[i].h.copy(), grids[i].hmax.copy(), acc.step, acc.t,
                            acc.V_rain))
        res[gb] = out
    for gb in (0, 12):
        for x, y in zip(res[1], res[gb]):
            assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])
            assert x[2:] == y[2:]
