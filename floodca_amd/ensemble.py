"""Scenario ensembles on one GPU: the batched session against one run per
scenario.

The workload follows the reference's master-worker ensemble
(src/hybrid/main_ens_mpi_cuda.c:71-120, ens_common.c:166-258): every scenario
is a full run from a dry start with its own forcing, and reports its flooded
LAND cells at the thresholds {0.10, 0.25, 0.50} m and its largest depth. Here
scenarios also differ in grid: square crops of one large synthetic DEM with
uneven sides, so a batch mixes grid sizes and time steps.

Scenario draw (seeded):
  * crop: side uniform in [side_lo, side_hi], position uniform in the big grid;
  * rain: a triangular hyetograph (rain.c:13-28) of duration D uniform in
    [dur_lo, dur_hi] hours with its peak at mid-storm (phi = 0.5), mean
    intensity 50 mm/h times a lognormal factor (median 1, sigma 0.5), so the
    storm depth Rday = 0.05 m/h * factor * D (C = 1); the run lasts the storm,
    T_sim = D;
  * roughness: each of the 8 land-cover classes scaled by U[0.8, 1.2].
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np

from . import abi as A
from .grid import Grid
from .hip import Batch
from .synth import MANNING_CLASSES

THRESHOLDS = (0.10, 0.25, 0.50)


@dataclass
class Scenario:
    idx: int
    r0: int
    c0: int
    side: int
    factor: float
    dur_h: float
    nscale: np.ndarray


def draw(n: int, seed: int, big: tuple, side=(256, 2048), dur=(2.0, 12.0), sigma: float = 0.5):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        s = int(rng.integers(side[0], side[1] + 1))
        s = min(s, big[0], big[1])
        out.append(Scenario(idx=i, r0=int(rng.integers(0, big[0] - s + 1)),
                            c0=int(rng.integers(0, big[1] - s + 1)), side=s,
                            factor=float(rng.lognormal(0.0, sigma)),
                            dur_h=float(rng.uniform(dur[0], dur[1])),
                            nscale=rng.uniform(0.8, 1.2, MANNING_CLASSES.size)))
    return out


def build(base: Grid, sc: Scenario, prec: str = "f32"):
    """The scenario's grid (a crop of base with scaled roughness) and config."""
    a = A.abi(prec)
    dt = a.dtype
    sl = (slice(sc.r0, sc.r0 + sc.side), slice(sc.c0, sc.c0 + sc.side))
    mask = np.ascontiguousarray(base.mask[sl])
    n0 = base.n[sl]
    cls = np.searchsorted(MANNING_CLASSES.astype(n0.dtype), n0).clip(0, MANNING_CLASSES.size - 1)
    n = (MANNING_CLASSES[cls] * sc.nscale[cls]).astype(dt)
    shape = mask.shape
    depth = 0.05 * sc.factor * sc.dur_h
    g = Grid(z=np.ascontiguousarray(base.z[sl], dtype=dt), mask=mask, n=n,
             C=np.ones(shape, dtype=dt), Rday=np.full(shape, depth, dtype=dt),
             h=np.zeros(shape, dtype=dt), dx=base.dx)
    cfg = A.default_cfg(a)
    A.set_dx(a, cfg, g.dx)
    cfg.hyeto = A.FCA_HYETO_TRIANGLE
    cfg.D_h, cfg.phi = sc.dur_h, 0.5
    cfg.T_sim = sc.dur_h * 3600.0
    cfg.steps = 0
    return g, cfg


def classes(base: Grid) -> np.ndarray:
    """Land-cover class of every cell of base (its roughness's index in
    MANNING_CLASSES), as build() recovers it."""
    cls = np.searchsorted(MANNING_CLASSES.astype(base.n.dtype), base.n)
    return cls.clip(0, MANNING_CLASSES.size - 1).astype(np.uint8)


def crop_of(sc: Scenario, prec: str = "f32") -> dict:
    """The scenario as an FcaHipCrop: the same window, per-class roughness
    and storm depth as build() gives its grid, so a batch of crops built on
    the device reproduces a batch of built grids bit for bit."""
    dt = A.abi(prec).dtype
    return dict(r0=sc.r0, c0=sc.c0, nrows=sc.side, ncols=sc.side,
                rday=float(dt.type(0.05 * sc.factor * sc.dur_h)),
                n_class=[float(v) for v in (MANNING_CLASSES * sc.nscale).astype(dt)])


def config_of(sc: Scenario, dx: float, prec: str = "f32"):
    """The scenario's run configuration (as build())."""
    a = A.abi(prec)
    cfg = A.default_cfg(a)
    A.set_dx(a, cfg, dx)
    cfg.hyeto = A.FCA_HYETO_TRIANGLE
    cfg.D_h, cfg.phi = sc.dur_h, 0.5
    cfg.T_sim = sc.dur_h * 3600.0
    cfg.steps = 0
    return cfg


def device_base(base: Grid, prec: str = "f32"):
    """base's bed, mask and classes resident on the device."""
    from .hip import DeviceBase
    return DeviceBase(base, classes(base), prec=prec)


def run_batched_crops(dbase, scens, dx: float, prec="f32", strip_rows=0, phases=None,
                      run_wall=None, t_sim_cap=None, tiles_per_wave=0, cols_per_lane=0):
    """One batch of crop scenarios built on the device from dbase (a
    DeviceBase): returns (seconds, per-scenario (flooded, max depth, steps)),
    the time covering the configs, the scenario table, the run, the
    reductions and the release, as run_batched."""
    t0 = time.perf_counter()
    cfgs = [config_of(s, dx, prec) for s in scens]
    if t_sim_cap is not None:
        cfgs = [_with_tsim(c, min(float(c.T_sim), t_sim_cap)) for c in cfgs]
    crops = [crop_of(s, prec) for s in scens]
    t_cfg = time.perf_counter() - t0
    secs, res = run_crop_batch(dbase, crops, cfgs, prec, strip_rows, phases, run_wall,
                               tiles_per_wave, cols_per_lane)
    return t_cfg + secs, res


# The reference's own ensemble benchmark restated on synthetic inputs: 100
# scenarios on one 90 m grid of 1230 x 1206 cells, 24 h each, forcing drawn
# as its scenario generator draws it (results/e10_gpu100.csv:2,
# prep/floodca_prep/scenarios.py:34-53, src/hybrid/main_ens_mpi_cuda.c:69-120).
DAY_GRID = (1230, 1206)
DAY_RDAY = 0.03   # event rain depth (m) before lambda; C = 1


@dataclass
class DayScenario:
    idx: int
    lam: float
    dur_h: int
    phi: float


def draw_days(n: int, seed: int):
    """Forcing tuples drawn with the distributions of the reference's
    scenario generator (scenarios.py:44-46): lambda U[0.8, 1.5], storm
    duration one of 6, 12, 24 h, peak position phi U[0.25, 0.75]."""
    rng = np.random.default_rng(seed)
    lam = rng.uniform(0.8, 1.5, size=n)
    dur = rng.choice(np.array([6, 12, 24]), size=n)
    phi = rng.uniform(0.25, 0.75, size=n)
    return [DayScenario(i, float(lam[i]), int(dur[i]), float(phi[i])) for i in range(n)]


SCENARIO_FIELDS = ("scenario_id", "base_date", "lambda", "D_h", "phi", "seed")


def write_scenario_csv(path: str, dates, n: int, seed: int) -> list[dict]:
    """A scenario table in the reference's columns (scenarios.py:18): base
    date drawn from `dates`, lambda U[0.8, 1.5], D_h one of 6, 12, 24 h, phi
    U[0.25, 0.75] and a per-scenario seed, with lambda and phi written with 17
    significant digits as the reference's generator writes them
    (scenarios.py:34-53). Returns the rows."""
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, len(dates), size=n)
    lam = rng.uniform(0.8, 1.5, size=n)
    dur = rng.choice(np.array([6, 12, 24]), size=n)
    phi = rng.uniform(0.25, 0.75, size=n)
    seeds = rng.integers(0, 2**31 - 1, size=n, dtype=np.int64)
    rows = [{"scenario_id": k, "base_date": dates[int(idx[k])], "lambda": f"{float(lam[k]):.17g}",
             "D_h": int(dur[k]), "phi": f"{float(phi[k]):.17g}", "seed": int(seeds[k])}
            for k in range(n)]
    with open(path, "w", encoding="utf-8") as fp:
        fp.write(",".join(SCENARIO_FIELDS) + "\n")
        for r in rows:
            fp.write(",".join(str(r[f]) for f in SCENARIO_FIELDS) + "\n")
    return rows


def day_crop(shape, sc: DayScenario, prec: str = "f32") -> dict:
    """The whole base grid with the scenario's rain depth DAY_RDAY x lambda
    and the base roughness."""
    dt = A.abi(prec).dtype
    return dict(r0=0, c0=0, nrows=shape[0], ncols=shape[1],
                rday=float(dt.type(DAY_RDAY * sc.lam)),
                n_class=[float(v) for v in MANNING_CLASSES.astype(dt)])


def day_config(sc: DayScenario, dx: float, prec: str = "f32"):
    """A 24 h run under the scenario's triangular storm."""
    a = A.abi(prec)
    cfg = A.default_cfg(a)
    A.set_dx(a, cfg, dx)
    cfg.hyeto = A.FCA_HYETO_TRIANGLE
    cfg.D_h, cfg.phi = float(sc.dur_h), sc.phi
    cfg.T_sim = 86400.0
    cfg.steps = 0
    return cfg


def run_crop_batch(dbase, crops, cfgs, prec="f32", strip_rows=0, phases=None, run_wall=None,
                   tiles_per_wave=0, cols_per_lane=0):
    """One batch of crops (dicts as crop_of gives) with their configs, built
    on the device from dbase: returns (seconds, per-scenario (flooded, max
    depth, steps)), the time covering the scenario table, the run, the
    reductions and the release; run_wall, if a list, receives the seconds of
    the run alone; phases, if a dict, the seconds of each phase."""
    t0 = time.perf_counter()
    res = []
    b = Batch.from_crops(dbase, crops, cfgs, prec=prec, strip_rows=strip_rows,
                         tiles_per_wave=tiles_per_wave, cols_per_lane=cols_per_lane)
    try:
        t1 = time.perf_counter()
        b.run()
        t2 = time.perf_counter()
        for i in range(len(crops)):
            fl, md = b.flooded(i, THRESHOLDS)
            res.append((fl, md, b.steps(i)))
        t3 = time.perf_counter()
        if run_wall is not None:
            run_wall.append(b.wall)
    finally:
        b.close()
    t4 = time.perf_counter()
    if phases is not None:
        phases.update(open=t1 - t0, run=t2 - t1, results=t3 - t2, close=t4 - t3)
    return t4 - t0, res


def run_alone_crops(dbase, scens, dx: float, prec="f32"):
    """Each crop scenario in a session of its own, one after another, built
    on the device as run_batched_crops builds them: returns (seconds,
    per-scenario results)."""
    t0 = time.perf_counter()
    res = []
    for s in scens:
        with Batch.from_crops(dbase, [crop_of(s, prec)], [config_of(s, dx, prec)], prec=prec,
                              strip_rows=strip_alone(s.side, s.side)) as b:
            b.run()
            fl, md = b.flooded(0, THRESHOLDS)
            res.append((fl, md, b.steps(0)))
    return time.perf_counter() - t0, res


def acc_start(prec: str, g: Grid, cfg):
    """Accumulators of a run from g's state (fca_acc_init, model.c:102-109):
    V0 = V = sum of LAND h*A in double, force_day = -1, the rest zero."""
    a = A.abi(prec)
    acc = a.FcaAcc()
    land = g.mask == A.FCA_LAND
    acc.V0 = acc.V = float(np.sum(g.h[land].astype(np.float64) * float(cfg.phys.A)))
    acc.force_day = -1
    return acc


def run_batched(cases, prec="f32", strip_rows=0, run_wall=None, phases=None):
    """One batch over all cases [(grid, cfg)]: returns (seconds, per-scenario
    (flooded, max depth, steps)). The time covers upload, the run, the
    per-scenario reductions and the release; run_wall, if a list, receives
    the seconds of the run alone; phases, if a dict, the seconds of each
    phase (open, run, results, close)."""
    t0 = time.perf_counter()
    accs = [acc_start(prec, g, c) for g, c in cases]
    res = []
    b = Batch([g for g, _ in cases], [c for _, c in cases], accs, prec=prec,
              strip_rows=strip_rows)
    try:
        t1 = time.perf_counter()
        b.run()
        t2 = time.perf_counter()
        for i in range(len(cases)):
            fl, md = b.flooded(i, THRESHOLDS)
            res.append((fl, md, b.steps(i)))
        t3 = time.perf_counter()
        if run_wall is not None:
            run_wall.append(b.wall)
    finally:
        b.close()
    t4 = time.perf_counter()
    if phases is not None:
        phases.update(open=t1 - t0, run=t2 - t1, results=t3 - t2, close=t4 - t3)
    return t4 - t0, res


def strip_alone(nr: int, nc: int) -> int:
    """Strip height for a grid run on its own: short enough that the grid
    yields ~4096 wavefront tiles (the GPU's resident wavefronts at this
    kernel's occupancy), at least 8 rows and at most 96."""
    groups = (nc + 57) // 58
    return int(max(8, min(96, nr * groups // 4096)))


def run_alone(cases, prec="f32", strip_rows=None):
    """Each case in a session of its own, one after another (the reference's
    per-scenario ensemble): returns (seconds, per-scenario results). With
    strip_rows None each grid takes strip_alone's height."""
    t0 = time.perf_counter()
    res = []
    for g, c in cases:
        s = strip_alone(*g.shape) if strip_rows is None else strip_rows
        with Batch([g], [c], [acc_start(prec, g, c)], prec=prec, strip_rows=s) as b:
            b.run()
            fl, md = b.flooded(0, THRESHOLDS)
            res.append((fl, md, b.steps(0)))
    return time.perf_counter() - t0, res


def _with_tsim(cfg, t_sim):
    c = type(cfg)()
    ctypes.memmove(ctypes.addressof(c), ctypes.addressof(cfg), ctypes.sizeof(cfg))
    c.T_sim = t_sim
    return c


def autotune_strip(cases, prec="f32", candidates=(32, 48, 64, 96, 128), sample=24, seed=0):
    """Strip height for a batch: the candidate with the shortest batched run
    of a seeded sample of the cases, each cut to its first 20 minutes of
    simulated time (the run only, not the upload). Results do not depend on
    the strip height (the walk's arithmetic is per cell), only the time does.
    Returns (best, {candidate: seconds})."""
    rng = np.random.default_rng(seed)
    idx = rng.choice(len(cases), size=min(sample, len(cases)), replace=False)
    sub = [(cases[i][0], _with_tsim(cases[i][1], min(float(cases[i][1].T_sim), 1200.0)))
           for i in idx]
    timings = {}
    for s in candidates:
        wall = []
        run_batched(sub, prec, strip_rows=s, run_wall=wall)
        timings[s] = wall[0]
    best = min(timings, key=timings.get)
    return best, timings


def autotune_tile_crops(dbase, scens, dx: float, prec="f32", strips=(32, 48, 64, 96, 128),
                        tiles=(1, 2, 4), cols=(1, 2), sample=512, seed=0):
    """The tile shape and cells per lane of a batch of crop scenarios built on
    the device: the strip height (rows a wavefront walks per tile), the tiles
    each wavefront walks per launch (k > 1 launches a k-th of the workgroups,
    every lane walking k strips in turn) and the columns each work-item walks
    (1: tiles 58 output columns wide; 2: 122 columns, float only). Each candidate runs the sample batched, every scenario
    cut to its first 2 h of simulated time (runs of 20 min took ~10 ms, too
    short to rank them); the heights first at one tile per wavefront, then
    the tile counts at the best height. The sample should be as large as the
    batch it tunes: how many tiles a launch has decides between short strips
    (fuller GPU) and tall ones (fewer halo rows), and a sample of 24 crops
    picked 32 rows where batches of 512 ran faster with 96. Results do not
    depend on the choice (the walk's arithmetic is per cell), only the time
    does. The heights are timed at one tile and one column per work-item,
    then the tile counts at the best height, then the column counts at the
    best of those. Returns ((strip_rows, tiles_per_wave, cols_per_lane),
    {"strip x tiles x cols": seconds})."""
    rng = np.random.default_rng(seed)
    idx = rng.choice(len(scens), size=min(sample, len(scens)), replace=False)
    sub = [scens[i] for i in idx]
    timings = {}

    def timed(strip, k, c):
        key = f"{strip}x{k}x{c}"
        if key not in timings:
            wall = []
            run_batched_crops(dbase, sub, dx, prec, strip_rows=strip, run_wall=wall,
                              t_sim_cap=7200.0, tiles_per_wave=k, cols_per_lane=c)
            timings[key] = wall[0]
        return timings[key]
    if prec != "f32":
        cols = (1,)
    best_strip = min(strips, key=lambda st: timed(st, 1, 1))
    best_k = min(tiles, key=lambda k: timed(best_strip, k, 1))
    best_c = min(cols, key=lambda c: timed(best_strip, best_k, c))
    return (best_strip, best_k, best_c), timings
