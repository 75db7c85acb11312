"""The oracle restatement against the reference's own code compiled here
(oracle/_ref, same flags): bit for bit on random and special face inputs, on
the forcing time line, on per-cell time steps, and on whole runs including a
two-day forcing CSV. Skipped where the reference tree was not available at
build time (the golden fixtures in tests/golden carry its results there)."""

import numpy as np
import pytest

from conftest import make_case
from floodca_amd import abi as A
from oracle import Oracle, Reference, ref_available

pytestmark = pytest.mark.skipif(not ref_available("f32"), reason="oracle/_ref not built")

PRECS = ["f32", "f64"]


def random_faces(rng, n, dtype):
    """Face inputs covering every branch: walls, sea on either side, dead band,
    dry, upwind sea, tiny roughness, supercritical slopes, cap-binding dt."""
    z = rng.uniform(-2, 30, (n, 2))
    h = np.where(rng.random((n, 2)) < 0.2, 0.0, rng.exponential(0.3, (n, 2)))
    # near-flat levels for the tau_l dead band
    flat = rng.random(n) < 0.1
    z[flat, 1] = z[flat, 0] + h[flat, 0] - h[flat, 1] + rng.uniform(-2e-5, 2e-5, flat.sum())
    nn = np.where(rng.random((n, 2)) < 0.02, 1e-14, rng.uniform(0.01, 0.2, (n, 2)))
    m = rng.choice([0, 0, 0, 0, 1, 2], size=(n, 2)).astype(np.uint8)
    d = rng.choice([30.0, 90.0, 127.27922], size=n)
    w = rng.choice([90.0, 37.27922], size=n)
    dt = np.where(rng.random(n) < 0.3, rng.uniform(0.01, 1.0, n), rng.uniform(1, 60, n))
    return [x.astype(dtype) if x.dtype.kind == "f" else x for x in (z, h, nn, m, d, w, dt)]


@pytest.mark.parametrize("prec", PRECS)
def test_face_flux_random(prec):
    o, r = Oracle(prec), Reference(prec)
    a = A.abi(prec)
    rng = np.random.default_rng(7)
    z, h, nn, m, d, w, dt = random_faces(rng, 20000, a.dtype)
    cfg = A.default_cfg(a)
    p = cfg.phys
    for i in range(len(d)):
        K = 8 if i % 3 == 0 else 4
        args = (float(z[i, 0]), float(h[i, 0]), float(nn[i, 0]), int(m[i, 0]), float(z[i, 1]),
                float(h[i, 1]), float(nn[i, 1]), int(m[i, 1]), float(d[i]), float(w[i]),
                float(dt[i]), K, 8100.0, p)
        qo, co = o.flux(*args)
        qr, cr = r.flux(*args)
        assert (qo == qr or (qo != qo and qr != qr)) and co == cr, (i, args, qo, qr)


@pytest.mark.parametrize("prec", PRECS)
def test_scalar_helpers_random(prec):
    o, r = Oracle(prec), Reference(prec)
    rng = np.random.default_rng(3)
    for x in np.concatenate([rng.exponential(1.0, 3000), [0.0, 1e-30, 1e-3, 1.0, 1e6]]):
        assert o.cbrt_sq(float(x)) == r.cbrt_sq(float(x))
        hv, av = float(rng.exponential(0.5)), 8100.0
        assert o.lib.or_limiter_s(float(x), hv, av) == r.lib.ref_limiter_s(float(x), hv, av)


@pytest.mark.parametrize("prec", PRECS)
def test_forcing_time_line_random(prec):
    o, r = Oracle(prec), Reference(prec)
    a = A.abi(prec)
    rng = np.random.default_rng(5)
    for i in range(3000):
        cfg = A.default_cfg(a)
        cfg.hyeto = A.FCA_HYETO_UNIFORM if i % 4 == 0 else A.FCA_HYETO_TRIANGLE
        cfg.D_h = float(rng.choice([0.0, 1.0, 2.5, 8.0, 24.0]))
        cfg.phi = float(rng.choice([0.0, 0.25, 0.5, 1.0]))
        cfg.out_every = float(rng.choice([0.0, 900.0, 3600.0]))
        cfg.T_sim = float(rng.choice([3600.0, 21600.0, 86400.0 * 3]))
        t = float(rng.uniform(-100, 200000)) if i % 5 else float(rng.choice([0, 86400, 900]))
        dt = float(rng.uniform(-1, 120))
        assert o.lib.or_hyeto_integral(cfg, t, dt) == r.lib.fca_hyeto_integral(cfg, t, dt)
        assert o.lib.or_next_event_time(cfg, t) == r.lib.fca_next_event_time(cfg, t)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("K", [4, 8])
def test_cell_dt_every_cell(prec, K):
    a, g, cfg = make_case(23, 19, prec=prec, K=K, walls=True, h0=0.02)
    rng = np.random.default_rng(9)
    g.h[...] = np.where(rng.random(g.shape) < 0.3, 0.0, rng.exponential(0.2, g.shape))
    o, r = Oracle(prec), Reference(prec)
    for bnd in (A.FCA_BND_OPEN, A.FCA_BND_CLOSED):
        cfg.phys.boundary = bnd
        for i in range(g.nrows):
            for j in range(g.ncols):
                assert o.cell_dt(g, i, j, cfg.phys) == r.cell_dt(g, i, j, cfg.phys)


RUNS = [
    dict(nr=45, nc=38, K=4, boundary=A.FCA_BND_OPEN, hyeto=A.FCA_HYETO_UNIFORM, T_sim=7200.0),
    dict(nr=45, nc=38, K=8, boundary=A.FCA_BND_CLOSED, hyeto=A.FCA_HYETO_TRIANGLE, D_h=1.0,
         phi=0.5, rday=50.0, T_sim=5400.0, walls=True),
    dict(nr=31, nc=64, K=8, U=4, steps=300, T_sim=1e9, out_every=600.0),
    dict(nr=1, nc=40, K=4, T_sim=3600.0),
    dict(nr=40, nc=2, K=8, T_sim=3600.0),
]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("run", range(len(RUNS)))
def test_whole_runs_bitwise(prec, run):
    kw = dict(RUNS[run])
    nr, nc = kw.pop("nr"), kw.pop("nc")
    a, g, cfg = make_case(nr, nc, prec=prec, **kw)
    go, gr = g.copy(), g.copy()
    acc_o, log_o = Oracle(prec).run(go, cfg, log_cap=100000)
    acc_r, log_r = Reference(prec).run(gr, cfg, log_cap=100000)
    for k in ("h", "hmax", "tmax"):
        assert np.array_equal(getattr(go, k), getattr(gr, k)), k
    assert acc_o.step == acc_r.step == len(log_r)
    assert [x.t for x in log_o] == list(log_r[:, 0])
    for k in ("V", "V_rain", "V_sea", "V_bnd", "clamp_err", "t", "cap_active", "cap_total",
              "dt_min_hits", "dt_base"):
        assert getattr(acc_o, k) == getattr(acc_r, k), k


@pytest.mark.parametrize("prec", PRECS)
def test_two_day_forcing_csv(prec, tmp_path):
    """Day-boundary Rday reloads (model.c:338-358) through the reference's own
    CSV loader (rain.c:120-200), lambda scaling, and a missing third day."""
    a, g, cfg = make_case(29, 23, prec=prec, K=4, T_sim=86400.0 * 3)
    cfg.hyeto = A.FCA_HYETO_UNIFORM
    cfg.lambda_ = 0.8
    rng = np.random.default_rng(1)
    for d in range(2):
        (rng.random(g.shape) * (0.3 + d)).astype(np.float32).tofile(tmp_path / f"r{d}.bin")
    (tmp_path / "forcing.csv").write_text(
        "day_index,date,file\n0,2020-06-01,r0.bin\n1,2020-06-02,r1.bin\n2,2020-06-03,r2.bin\n")
    r = Reference(prec)
    assert r.lib.fca_forcing_load_csv(cfg, str(tmp_path / "forcing.csv").encode()) == 0
    assert cfg.n_force_days == 3
    go, gr = g.copy(), g.copy()
    acc_o, _ = Oracle(prec).run(go, cfg)
    acc_r, _ = r.run(gr, cfg)
    assert acc_o.force_day == acc_r.force_day == 1
    assert acc_o.t == acc_r.t == 86400.0 * 3  # day 3 missing: the run ends at T_sim
    assert acc_o.step == acc_r.step
    for k in ("h", "hmax", "tmax", "Rday"):
        assert np.array_equal(getattr(go, k), getattr(gr, k)), k
