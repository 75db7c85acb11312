#!/usr/bin/env python3
"""Benchmark of the fused WCA step on one MI355X.

Workload (default): the single large grid of the plain method -- 16384^2
cells at 90 m, seeded fractal DEM on a 0->60 m plane with a west sea (~70%
LAND), 8-class Voronoi Manning n, C = 1, uniform 50 mm/h rain (Rday = 1.2 m,
uniform hyetograph), dry start, K = 4, U = 1, FP32, open boundary. T_sim is
set to 1e9 s so the run is bounded by the step count, as the reference's
benchmarks do (--steps N --tsim 1e9).

A "step" is one launch of the fused step kernel over the whole grid. The
flood first spins up for a fixed 300 steps (untimed; they wet every LAND cell,
so the timed window is the same mid-flood state whatever W and K are), then W
warmup steps run (untimed), then exactly K steps are timed with HIP events on
the session's stream, inputs resident in HBM.

Printed: one JSON line with value = LAND cell-updates/s (the reference's MCUPS
definition x 1e6, src/common/log.c:133-135), the HBM roofline of the step
kernel (algorithmic bytes per launch / average launch time vs 8 TB/s), and
the reference serial build's throughput on the host cores (a bounded sample
of the 512^2 case) as the CPU baseline. Nested in the same line, each with
its own workload: the 4096^2 grid in FP32 and FP64, the sparse start on 8192^2 (quiet-tile skipping against the
dense walk), a 96-scenario ensemble (batched against one session per
scenario), the reference's own ensemble benchmark in shape (100 scenarios of
1230 x 1206 cells for 24 h each, against its published V100 figure) and the
surrogate-screened ensemble (4096 CA runs of training data
and 20000 candidates by default; --sur-cand 100000 is the
full case).

``--calibrate-counters`` instead copies a buffer of known size (hipMemcpy
device to device) so that rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of that
dispatch can be set against the bytes it must move; the factors scale the
fused kernel's counters into HBM bytes (profiles/).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from floodca_amd import abi as A  # noqa: E402
from floodca_amd import synth  # noqa: E402
from floodca_amd.grid import write_dir  # noqa: E402
from floodca_amd.hip import Sim, backend  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "cell-updates/s & %HBM roofline; ensemble scenarios/h surrogate vs plain CA"


def bytes_per_step(n_land: int, n_all: int, prec: str) -> int:
    """Algorithmic HBM bytes of one fused step: each LAND cell reads z, n,
    cr (= C*Rday), h, hmax and mask and writes h; every other cell reads only
    its mask. The hmax/tmax writes of rising cells are not credited."""
    w = 4 if prec == "f32" else 8
    return n_land * (5 * w + 1 + w) + (n_all - n_land) * 1


class HipTimer:
    """HIP events on a given stream (the session's own)."""

    def __init__(self):
        try:
            self.rt = C.CDLL("libamdhip64.so")
        except OSError:
            self.rt = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.rt.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.rt.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.rt.hipEventSynchronize.argtypes = [C.c_void_p]
        self.rt.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.rt.hipEventDestroy.argtypes = [C.c_void_p]
        self.a, self.b = C.c_void_p(), C.c_void_p()
        for e in (self.a, self.b):
            if self.rt.hipEventCreate(C.byref(e)) != 0:
                raise RuntimeError("hipEventCreate failed")

    def start(self, stream):
        self.rt.hipEventRecord(self.a, stream)

    def stop(self, stream) -> float:
        self.rt.hipEventRecord(self.b, stream)
        self.rt.hipEventSynchronize(self.b)
        ms = C.c_float()
        self.rt.hipEventElapsedTime(C.byref(ms), self.a, self.b)
        return float(ms.value)

    def device_sync(self):
        self.rt.hipDeviceSynchronize()


def calibrate_counters(mib: int = 1024, reps: int = 3) -> int:
    """Copies mib MiB device to device reps times (each copy reads and writes
    exactly mib MiB), for the counter calibration under rocprofv3 --pmc."""
    rt = HipTimer().rt
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    if rt.hipMalloc(C.byref(a), C.c_size_t(n)) or rt.hipMalloc(C.byref(b), C.c_size_t(n)):
        print("hipMalloc failed", file=sys.stderr)
        return 1
    rt.hipMemset(a, 1, C.c_size_t(n))
    rt.hipDeviceSynchronize()
    for _ in range(reps):
        rt.hipMemcpy(b, a, C.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    rt.hipDeviceSynchronize()
    rt.hipFree(a)
    rt.hipFree(b)
    print(f"copied {mib} MiB x {reps}", flush=True)
    return 0


def load_traffic(workload: str):
    """Per-launch HBM bytes measured with rocprofv3 --pmc (FETCH_SIZE x2 +
    WRITE_SIZE, scaled by what the counters read for copies of known size)
    for this workload, as committed under profiles/ (None if not measured)."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p, encoding="utf-8") as fp:
            t = json.load(fp).get(workload)
    except (OSError, ValueError):
        return None
    return None if t is None else float(t["bytes"])


VALU_ISSUE_PER_S = 256 * 2.4e9  # wave64 VALU instructions per second: one per CU per clock


def load_valu(workload: str):
    """(wave64 VALU instructions per launch, shader clock in Hz during the
    launch) for this workload, as committed under profiles/ (rocprofv3 --pmc
    SQ_INSTS_VALU; GRBM_GUI_ACTIVE per XCD over the dispatch's duration);
    None for what was not measured."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p, encoding="utf-8") as fp:
            t = json.load(fp).get(workload)
    except (OSError, ValueError):
        return None, None
    if t is None:
        return None, None
    return (float(t["valu_insts"]) if "valu_insts" in t else None,
            float(t["clock_hz"]) if "clock_hz" in t else None)


def cpu_baseline(seconds: float, prec: str):
    """The reference serial build (oracle/_ref, compiled from the reference's
    sources with its stock -O3 flags) on the 512^2 case, one core, timed for
    about `seconds` of CPU work. Falls back to the oracle restatement
    ("port") when the reference build is absent."""
    exe = os.path.join(ROOT, "oracle", "_ref", f"floodca_serial_{prec}")
    g = synth.make_grid(512, 512, seed=20261203)
    n_land = g.n_land
    with tempfile.TemporaryDirectory() as td:
        write_dir(td, g)
        if os.path.exists(exe):
            def run(steps):
                out = subprocess.run([exe, "--grid", td, "--steps", str(steps), "--tsim", "1e9",
                                      "--hyeto", "uniform", "--no-output", "--csv", ""],
                                     capture_output=True, text=True, check=True).stdout
                m = re.search(r"wall=([0-9.eE+-]+)", out)
                return float(m.group(1))
            w0 = run(20)
            steps = max(20, int(seconds / max(w0 / 20, 1e-6)))
            wall = run(steps)
            kind = "reference"
        else:
            from oracle import Oracle
            a = A.abi(prec)
            o = Oracle(prec)
            cfg = A.default_cfg(a)
            A.set_dx(a, cfg, g.dx)
            cfg.hyeto, cfg.T_sim = A.FCA_HYETO_UNIFORM, 1e9
            gg = g.astype(a.dtype)
            cfg.steps = 10
            t = time.time()
            o.run(gg.copy(), cfg)
            w0 = time.time() - t
            steps = max(10, int(seconds / max(w0 / 10, 1e-6)))
            cfg.steps = steps
            t = time.time()
            o.run(gg, cfg)
            wall = time.time() - t
            kind = "port"
    out = {"value": n_land * steps / wall, "unit": "cell-updates/s", "cores": 1, "kind": kind,
           "sample": f"512^2 grid, {n_land} LAND cells, {steps} steps from a dry start with "
                     f"50 mm/h rain, FP32, serial, wall {wall:.1f} s"}
    try:
        out["omp"] = cpu_baseline_omp(seconds, prec)
    except (OSError, subprocess.SubprocessError, AttributeError, ValueError) as e:
        out["omp"] = {"error": f"{type(e).__name__}: {e}"}
    if kind != "reference":
        out["note"] = (f"{exe} is not built (oracle/build_oracle.sh needs the reference "
                       "sources): the oracle restatement was timed instead")
    return out


def cpu_baseline_omp(seconds: float, prec: str, size: int = 2048):
    """The reference OpenMP build (oracle/_ref/floodca_omp_*, stock -O3) on
    a size^2 grid of the same generator with every host core the process may
    use (OMP_NUM_THREADS, else the affinity mask), threads bound close to
    cores as the reference's job scripts set them (slurm/env.sh:35-36), timed
    for about `seconds`. None when the build is absent."""
    exe = os.path.join(ROOT, "oracle", "_ref", f"floodca_omp_{prec}")
    if not os.path.exists(exe):
        return None
    cores = int(os.environ.get("OMP_NUM_THREADS", 0)) or len(os.sched_getaffinity(0))
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    g = synth.make_grid(size, size, seed=20261203)
    with tempfile.TemporaryDirectory() as td:
        write_dir(td, g)

        def run(steps):
            out = subprocess.run([exe, "--grid", td, "--steps", str(steps), "--tsim", "1e9",
                                  "--hyeto", "uniform", "--threads", str(cores), "--no-output",
                                  "--csv", ""], capture_output=True, text=True, check=True,
                                 env=env).stdout
            return float(re.search(r"wall=([0-9.eE+-]+)", out).group(1))
        w0 = run(10)
        steps = max(10, int(seconds / max(w0 / 10, 1e-6)))
        wall = run(steps)
    return {"value": g.n_land * steps / wall, "unit": "cell-updates/s", "cores": cores,
            "kind": "reference",
            "sample": f"{size}^2 grid, {g.n_land} LAND cells, {steps} steps from a dry start "
                      f"with 50 mm/h rain, FP32, OpenMP on {cores} threads, wall {wall:.1f} s"}


def ensemble_bench(base, prec: str, n_scen: int, n_alone: int, seed: int = 20261204):
    """Scenario ensemble on crops of base (floodca_amd.ensemble): all n_scen
    scenarios in one batched session, and a seeded sample of n_alone of them
    run one session each (the per-scenario baseline) and batched again, whose
    results must be identical. Every scenario is built on the device from the
    resident base grid (uploaded once, outside the timing). Scenarios/h counts
    opening the sessions, the run, the per-scenario reductions and the
    release."""
    from floodca_amd import ensemble as E
    t0 = time.time()
    scen = E.draw(n_scen, seed, base.shape)
    dbase = E.device_base(base, prec)
    cells = sum(s.side * s.side for s in scen)
    t_base = time.time() - t0
    (strip, tpw, cpl), tune = E.autotune_tile_crops(dbase, scen, base.dx, prec)
    ph = {}
    t_b, r_b = E.run_batched_crops(dbase, scen, base.dx, prec, strip_rows=strip, phases=ph,
                                   tiles_per_wave=tpw, cols_per_lane=cpl)
    rng = np.random.default_rng(seed + 1)
    idx = sorted(rng.choice(n_scen, size=min(n_alone, n_scen), replace=False).tolist())
    sub = [scen[i] for i in idx]
    t_a, r_a = E.run_alone_crops(dbase, sub, base.dx, prec)
    t_s, r_s = E.run_batched_crops(dbase, sub, base.dx, prec, strip_rows=strip,
                                   tiles_per_wave=tpw, cols_per_lane=cpl)
    dbase.close()
    same = r_a == r_s == [r_b[i] for i in idx]
    steps = [r[2] for r in r_b]
    land = [int(np.count_nonzero(base.mask[sc.r0:sc.r0 + sc.side, sc.c0:sc.c0 + sc.side]
                                 == A.FCA_LAND)) for sc in scen]
    land_updates = float(sum(n * k for n, k in zip(land, steps)))
    print(f"[bench] ensemble: {n_scen} scenarios ({cells / 1e6:.0f} M cells), base on the "
          f"device in {t_base:.1f}s, strip {strip} x {tpw} tiles per wavefront x {cpl} "
          f"columns per work-item "
          f"({ {k: round(v, 3) for k, v in tune.items()} }), "
          f"batched {t_b:.2f}s ({ {k: round(v, 2) for k, v in ph.items()} }); "
          f"sample of {len(idx)}: alone {t_a:.2f}s, "
          f"batched {t_s:.2f}s, identical={same}", file=sys.stderr, flush=True)
    return {
        "scenarios_per_h": n_scen / t_b * 3600.0,
        "sample_alone_scenarios_per_h": len(idx) / t_a * 3600.0,
        "sample_batched_scenarios_per_h": len(idx) / t_s * 3600.0,
        # throughput of the whole batch against the sample run one session
        # each (the sample batched again is the identity check: at the
        # whole batch's strip height a batch of 24 is not tuned)
        "speedup_vs_alone": (t_a / len(idx)) / (t_b / n_scen),
        "sample_speedup_batched_vs_alone": t_a / t_s,
        "results_identical": bool(same),
        "batched_phases_s": {k: round(v, 3) for k, v in ph.items()},
        "land_cell_updates_per_s_batched_run": land_updates / ph["run"],
        "scenarios": n_scen, "sample": len(idx), "cells": cells,
        "steps_min_max": [min(steps), max(steps)], "strip_rows": strip,
        "tiles_per_wave": tpw, "cols_per_lane": cpl,
        "autotune_s": {k: round(v, 4) for k, v in tune.items()},
        "workload": "crops of the 16384^2 DEM, side U[256, 2048], built on the device; "
                    "triangular storm D U[2, 12] h, 50 mm/h x lognormal(0, 0.5); Manning x "
                    "U[0.8, 1.2] per class; T_sim = D; FP32, K=4",
    }


def day_ensemble_bench(prec: str, n_scen: int, n_alone: int, seed: int = 20261210):
    """The reference's own ensemble benchmark in shape (results/e10_gpu100.csv:2:
    100 scenarios, one 90 m grid of 1230 x 1206 cells, 24 h each, 609.3
    scenarios/h on one V100) on synthetic inputs: the same generator's DEM
    at that size, a uniform event rain of DAY_RDAY x lambda, forcing drawn
    with the reference generator's distributions. All scenarios in one
    batched session built on the device, and a seeded sample run one
    session each (the reference's per-scenario runs), whose results must be
    identical."""
    from floodca_amd import ensemble as E
    a = A.abi(prec)
    base = synth.make_grid(*E.DAY_GRID, seed=seed, dtype=a.dtype)
    scen = E.draw_days(n_scen, seed)
    crops = [E.day_crop(base.shape, s, prec) for s in scen]
    cfgs = [E.day_config(s, base.dx, prec) for s in scen]
    rng = np.random.default_rng(seed + 1)
    idx = sorted(rng.choice(n_scen, size=min(n_alone, n_scen), replace=False).tolist())
    with E.device_base(base, prec) as dbase:
        ph = {}
        t_b, r_b = E.run_crop_batch(dbase, crops, cfgs, prec, phases=ph)
        t0 = time.perf_counter()
        r_a = []
        for i in idx:
            r_a += E.run_crop_batch(dbase, [crops[i]], [cfgs[i]], prec,
                                    strip_rows=E.strip_alone(*base.shape))[1]
        t_a = time.perf_counter() - t0
    same = r_a == [r_b[i] for i in idx]
    steps = [r[2] for r in r_b]
    land = int(np.count_nonzero(base.mask == A.FCA_LAND))
    # the reference's figure and the median step count of its 100 runs
    # (results/e10_gpu100.csv:2; results/e10_gpu100/B4.csv, column steps): its
    # real terrain and rain need ~11x the steps of these synthetic days, so
    # the comparison is also given per step (same grid size)
    ref, ref_steps = 609.265, 16317.5
    sph = n_scen / t_b * 3600.0
    per_step = sph * float(np.median(steps)) / ref_steps / ref
    print(f"[bench] reference-shaped ensemble: {n_scen} scenarios of {base.shape[0]}x"
          f"{base.shape[1]} for 24 h, steps {min(steps)}..{max(steps)}, batched {t_b:.2f}s "
          f"({sph:.0f} scenarios/h; against the reference's V100 figure {sph / ref:.0f}x, "
          f"{per_step:.1f}x per step); sample of {len(idx)} alone {t_a:.2f}s, "
          f"identical={same}", file=sys.stderr, flush=True)
    return {
        "scenarios_per_h": sph,
        "alone_scenarios_per_h": len(idx) / t_a * 3600.0,
        "speedup_vs_alone": (t_a / len(idx)) / (t_b / n_scen),
        "results_identical": bool(same),
        "batched_phases_s": {k: round(v, 3) for k, v in ph.items()},
        "land_cell_updates_per_s_batched_run": land * float(sum(steps)) / ph["run"],
        "steps_min_median_max": [min(steps), int(np.median(steps)), max(steps)],
        "scenarios": n_scen, "sample": len(idx), "land_cells": land,
        "reference_scenarios_per_h": ref,
        "reference_median_steps": ref_steps,
        "reference": "results/e10_gpu100.csv:2, 1x V100, its own grid and daily rain fields",
        "vs_reference": sph / ref,
        "vs_reference_per_step": per_step,
        "workload": f"{n_scen} scenarios on one synthetic {base.shape[0]}x{base.shape[1]} grid "
                    f"at 90 m, 24 h each, triangular storm of {E.DAY_RDAY} m x lambda "
                    "(lambda U[0.8, 1.5], D_h in {6, 12, 24} h, phi U[0.25, 0.75]); FP32, K=4",
    }


def sparse_bench(prec: str, size: int = 8192, hours: float = 8.0):
    """Sparse-start flood front (no rain, one inland inflow cell with a
    triangular 500 m^3/s hydrograph peaking at 2 h over 8 h, dry start) on a
    size^2 grid of the same generator: the run to the end of the hydrograph
    with quiet-tile skipping, and again with every tile walked; the two must
    end identically."""
    a = A.abi(prec)
    g = synth.make_grid(size, size, seed=20261205, dtype=a.dtype)
    r, c = synth.point_inflow(g, peak_m3s=500.0, dur_h=hours)
    cfg = A.default_cfg(a)
    A.set_dx(a, cfg, g.dx)
    cfg.hyeto, cfg.D_h, cfg.phi = A.FCA_HYETO_TRIANGLE, hours, 0.25
    cfg.T_sim = hours * 3600.0
    out = {}
    for name, flags in (("skip", 0), ("dense", A.FCA_HIP_NO_SKIP)):
        gg = g.copy()
        acc = a.FcaAcc()
        acc.force_day = -1
        timer = HipTimer()
        with Sim(gg, cfg, acc, prec=prec, flags=flags) as sim:
            stream = C.c_void_p(sim.stream)
            timer.device_sync()
            timer.start(stream)
            sim.run()
            ms = timer.stop(stream)
            res = sim.read()
        out[name] = (ms, res.step, gg)
    same = all(np.array_equal(getattr(out["skip"][2], k), getattr(out["dense"][2], k))
               for k in ("h", "hmax", "tmax")) and out["skip"][1] == out["dense"][1]
    gs = out["skip"][2]
    wet = float(np.mean(gs.hmax[gs.mask == A.FCA_LAND] >= 1e-3))
    print(f"[bench] sparse start {size}^2: {out['skip'][1]} steps, skip {out['skip'][0]:.1f} ms, "
          f"dense {out['dense'][0]:.1f} ms, identical={same}, wet fraction {wet:.4f}",
          file=sys.stderr, flush=True)
    return {"grid": [size, size], "steps": out["skip"][1], "ms_skip": out["skip"][0],
            "ms_dense": out["dense"][0], "speedup": out["dense"][0] / out["skip"][0],
            "results_identical": bool(same), "final_wet_fraction_hmax_1mm": wet,
            "inflow_cell": [r, c],
            "workload": f"{size}^2 dry start, one inflow cell, triangular 500 m^3/s hydrograph "
                        f"peaking at 2 h over {hours:g} h, no rain, FP32, K=4"}


def sparse_sweep_bench(prec: str, size: int = 8192, fracs=(0.001, 0.01, 0.05, 0.2),
                       steps: int = 200, depth: float = 0.05):
    """Quiet-tile skipping against the wet fraction: the sparse-start grid (no
    rain but the inflow cell's hydrograph) from warm starts in which a disc of
    LAND around the inflow cell, holding `frac` of all LAND cells, is wet at
    `depth`; `steps` steps timed with skipping and with every tile walked
    (results must be identical). The hydrograph itself (7.2 Mm^3) cannot wet
    more than ~1.9% of this grid's LAND to 1 mm, so larger fractions need a
    warm start."""
    a = A.abi(prec)
    g0 = synth.make_grid(size, size, seed=20261205, dtype=a.dtype)
    r, c = synth.point_inflow(g0, peak_m3s=500.0, dur_h=8.0)
    land = g0.mask == A.FCA_LAND
    rr, cc = np.nonzero(land)
    d2 = (rr.astype(np.int64) - r) ** 2 + (cc.astype(np.int64) - c) ** 2
    order = np.argsort(d2, kind="stable")
    cfg = A.default_cfg(a)
    A.set_dx(a, cfg, g0.dx)
    cfg.hyeto, cfg.D_h, cfg.phi = A.FCA_HYETO_TRIANGLE, 8.0, 0.25
    cfg.T_sim = 1e9
    cfg.steps = steps
    res = []
    for f in fracs:
        k = int(round(f * rr.size))
        g = g0.copy()
        g.h[rr[order[:k]], cc[order[:k]]] = a.dtype.type(depth)
        g.hmax[...] = g.h
        out = {}
        for name, flags in (("skip", 0), ("dense", A.FCA_HIP_NO_SKIP)):
            gg = g.copy()
            acc = a.FcaAcc()
            acc.force_day = -1
            timer = HipTimer()
            with Sim(gg, cfg, acc, prec=prec, flags=flags) as sim:
                stream = C.c_void_p(sim.stream)
                timer.device_sync()
                timer.start(stream)
                sim.run()
                ms = timer.stop(stream)
                st = sim.read().step
            out[name] = (ms, st, gg)
        same = all(np.array_equal(getattr(out["skip"][2], q), getattr(out["dense"][2], q))
                   for q in ("h", "hmax", "tmax"))
        wet_end = float(np.mean(out["skip"][2].h[land] >= 1e-3))
        res.append({"wet_fraction_start": k / rr.size, "wet_fraction_end": wet_end,
                    "steps": out["skip"][1], "ms_skip": out["skip"][0],
                    "ms_dense": out["dense"][0], "speedup": out["dense"][0] / out["skip"][0],
                    "results_identical": bool(same)})
        print(f"[bench] sparse sweep {size}^2, {100 * k / rr.size:.2f}% of LAND wet: "
              f"{out['skip'][1]} steps, skip {out['skip'][0]:.1f} ms, dense "
              f"{out['dense'][0]:.1f} ms ({out['dense'][0] / out['skip'][0]:.2f}x), "
              f"identical={same}", file=sys.stderr, flush=True)
    return {"grid": [size, size], "runs": res,
            "workload": f"{size}^2, no rain but the inflow cell's hydrograph; warm start with a "
                        f"disc of LAND around the inflow cell {depth:g} m deep; {steps} steps "
                        f"with quiet-tile skipping and dense, FP32, K=4"}


def grid4096_bench(steps: int = 500, warmup: int = 300):
    """The plain method on the 4096^2 grid of the same generator in FP32 and
    in FP64 (the double-precision mode of the reference's comparison), timed
    as the main line: warmup steps from a dry start under 50 mm/h, then steps
    timed with HIP events on the session's stream."""
    out = {}
    for prec in ("f32", "f64"):
        a = A.abi(prec)
        g = synth.make_grid(4096, 4096, seed=20261203, dtype=a.dtype)
        cfg = A.default_cfg(a)
        A.set_dx(a, cfg, g.dx)
        cfg.hyeto = A.FCA_HYETO_UNIFORM
        cfg.T_sim = 1e9
        acc = a.FcaAcc()
        acc.force_day = -1
        timer = HipTimer()
        with Sim(g, cfg, acc, prec=prec) as sim:
            stream = C.c_void_p(sim.stream)
            sim.advance(warmup)
            sim.sync()
            timer.device_sync()
            timer.start(stream)
            sim.advance(steps)
            ms = timer.stop(stream)
            res = sim.read()
        bps = bytes_per_step(g.n_land, g.z.size, prec)
        out[prec] = {"ms_per_step": ms / steps,
                     "land_cell_updates_per_s": g.n_land * steps / (ms * 1e-3),
                     "algorithmic_gb_per_s": bps / (ms / steps * 1e-3) / 1e9,
                     "steps": steps, "warmup": warmup, "steps_taken": res.step}
        print(f"[bench] 4096^2 {prec}: {ms / steps:.4f} ms/step", file=sys.stderr, flush=True)
    out["workload"] = ("4096^2 at 90 m, same generator, 50 mm/h uniform rain from a dry start, "
                       "K=4, U=1, fused kernel; FP32 and FP64")
    return out


def forward_flops(cfg) -> int:
    """Multiply-adds x 2 of one sample's forward pass (every GEMM's M*N*K)."""
    from floodca_amd import nn
    total = 0
    for ly in nn.layers(cfg):
        total += ly.hout * ly.hout * ly.cout * ly.k
    return 2 * total


def surrogate_bench(base, args):
    """The surrogate-screened ensemble (floodca_amd.surrogate.run_pipeline)
    on crops of the benchmark DEM, and the forward pass timed alone with HIP
    events on its stream."""
    import torch
    from floodca_amd import nn
    from floodca_amd import surrogate as SG
    pc = SG.PipelineCfg(n_train=args.sur_train, n_val=max(64, args.sur_train // 16),
                        n_cand=args.sur_cand, epochs=args.sur_epochs, n_alone=args.sur_alone,
                        train_sigma=args.sur_train_sigma)
    res = SG.run_pipeline(base, pc, log=lambda m: print(f"[bench] {m}", file=sys.stderr,
                                                       flush=True))
    o = res.pop("_objects")
    net = o["net"]
    B = pc.screen_batch
    q = pc.R // 4
    X = torch.zeros((B, q, q, 64), dtype=torch.uint16, device="cuda")
    S = torch.zeros((B, 4), dtype=torch.float32, device="cuda")
    for _ in range(3):
        net.forward(X, S, B)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    reps = 20
    for _ in range(reps):
        net.forward(X, S, B)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    fl = forward_flops(net.cfg)
    tflops = fl * B / (ms * 1e-3) / 1e12
    res["forward_ms_per_batch"] = ms
    res["forward_batch"] = B
    res["forward_gflop_per_sample"] = fl / 1e9
    res["forward_tflops"] = tflops
    res["forward_mfma_frac_of_bf16_peak"] = tflops / 2500.0
    res["workload"] = (f"{pc.n_train} CA runs ({pc.n_val} held out) and {pc.n_cand} candidates "
                       f"from the ensemble generator on crops of the {base.shape[0]}^2 DEM, "
                       f"side U{list(pc.side)}, storm D U{list(pc.dur)} h; network "
                       f"{net.n} parameters at {pc.R}^2, BF16 on MFMA, FP32 accumulation; "
                       f"CA reruns of the top {pc.frac_area:.0%} by predicted flooded area "
                       f"+ {pc.frac_unc:.0%} by uncertainty, and of a seeded "
                       f"{pc.frac_audit:.0%} audit of the unflagged rest (counted in the "
                       f"time)")
    print(f"[bench] surrogate forward: {ms:.2f} ms per {B} samples, {tflops:.1f} TFLOP/s",
          file=sys.stderr, flush=True)
    return {k: (round(v, 6) if isinstance(v, float) else v) for k, v in res.items()}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--spinup", type=int, default=300,
                    help="untimed steps of the flood before the warmup (the timed state)")
    ap.add_argument("--size", type=int, default=16384)
    ap.add_argument("--prec", choices=("f32", "f64"), default="f32")
    ap.add_argument("--kernel", choices=("fused", "split"), default="fused")
    ap.add_argument("--strip", type=int, default=0, help="rows per wavefront strip (0 = default)")
    ap.add_argument("--cols-per-lane", type=int, default=0, choices=(0, 1, 2),
                    help="columns each work-item walks (0 = default 1; 2: float only)")
    ap.add_argument("--neigh", type=int, default=4, choices=(4, 8))
    ap.add_argument("--no-land-form", action="store_true",
                    help="fused walk in its general form only (same results)")
    ap.add_argument("--ens-scenarios", type=int, default=512,
                    help="scenarios of the ensemble measured after the main line (0 = skip)")
    ap.add_argument("--ens-alone", type=int, default=24,
                    help="of which this many are also run one session each")
    ap.add_argument("--day-scenarios", type=int, default=100,
                    help="scenarios of the reference-shaped ensemble (1230x1206, 24 h each; "
                         "0 = skip)")
    ap.add_argument("--day-alone", type=int, default=4,
                    help="of which this many are also run one session each")
    ap.add_argument("--sparse-size", type=int, default=8192,
                    help="grid side of the sparse-start measurement (0 = skip)")
    ap.add_argument("--surrogate", choices=("off", "on"), default="on",
                    help="run the surrogate-screened ensemble after the main line")
    ap.add_argument("--sur-train", type=int, default=4096,
                    help="CA runs of its dataset")
    ap.add_argument("--sur-cand", type=int, default=100000,
                    help="candidates it screens")
    ap.add_argument("--sur-epochs", type=int, default=30)
    ap.add_argument("--sur-train-sigma", type=float, default=0.5,
                    help="spread of the training storms' log intensity")
    ap.add_argument("--sur-alone", type=int, default=48,
                    help="candidates timed one session each (the all-CA baseline)")
    ap.add_argument("--no-4096", dest="grid4096", action="store_false",
                    help="skip the 4096^2 FP32/FP64 measurement")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-counters", action="store_true",
                    help="only copy 1 GiB device to device 3 times (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_counters:
        return calibrate_counters()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(one grid fits one MI355X's 288 GB, and nothing here spans GPUs)",
              file=sys.stderr)
        return 2

    prec = args.prec
    a = A.abi(prec)
    t0 = time.time()
    g = synth.make_grid(args.size, args.size, seed=20261203, dtype=a.dtype)
    t_gen = time.time() - t0
    n_all, n_land = g.z.size, g.n_land
    cfg = A.default_cfg(a)
    A.set_dx(a, cfg, g.dx)
    cfg.phys.K = cfg.neigh = args.neigh
    cfg.hyeto = A.FCA_HYETO_UNIFORM
    cfg.T_sim = 1e9
    acc = a.FcaAcc()
    acc.force_day = -1
    acc.V0 = acc.V = 0.0
    kern = A.FCA_HIP_FUSED if args.kernel == "fused" else A.FCA_HIP_SPLIT
    print(f"[bench] grid {args.size}^2 generated in {t_gen:.1f}s, LAND {n_land}, "
          f"device {backend(prec).device_name()}", file=sys.stderr, flush=True)

    timer = HipTimer()
    flags = A.FCA_HIP_NO_LAND_FORM if args.no_land_form else 0
    with Sim(g, cfg, acc, prec=prec, kernel=kern, strip_rows=args.strip, flags=flags,
             cols_per_lane=args.cols_per_lane) as sim:
        stream = C.c_void_p(sim.stream)
        sim.advance(args.spinup + args.warmup)
        sim.sync()
        timer.device_sync()
        timer.start(stream)
        sim.advance(args.steps)
        ms = timer.stop(stream)
        timer.device_sync()
        out = sim.read()
    steps_done = out.step
    first = args.spinup + args.warmup + 1
    if steps_done != first - 1 + args.steps:
        print(f"[bench] warning: {steps_done} steps taken of {first - 1 + args.steps}",
              file=sys.stderr)
    ms_per_step = ms / args.steps
    value = n_land * args.steps / (ms * 1e-3)
    bps = bytes_per_step(n_land, n_all, prec)
    achieved = bps / (ms_per_step * 1e-3) / 1e9
    workload = f"fused-{prec}-K{args.neigh}-{args.size}"
    # measured for the default kernel options only
    default_kernel = (kern == A.FCA_HIP_FUSED and args.strip == 0 and not args.no_land_form
                      and args.cols_per_lane in (0, 1))
    traffic = load_traffic(workload) if default_kernel else None
    valu, clock_hz = load_valu(workload) if default_kernel else (None, None)
    emass = abs(out.V + out.V_sea + out.V_bnd - out.V_rain - out.V0 - out.clamp_err) / \
        max(out.V_rain + out.V0, 1e-30)
    print(f"[bench] {ms_per_step:.4f} ms/step, {achieved:.0f} GB/s algorithmic, "
          f"t={out.t:.0f}s, E_mass={emass:.2e}", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": prec,
        "data": "synthetic (seeded fractal DEM, Voronoi land cover, uniform rain)",
        "config": {
            "workload": f"plain WCA step, single grid {args.size}^2 at 90 m, {prec.upper()}, "
                        f"K={args.neigh}, U=1, 50 mm/h uniform rain from a dry start, "
                        f"{args.kernel} kernel; timed steps {first}-{first - 1 + args.steps} "
                        f"of the flood",
            "timed_steps": [first, first - 1 + args.steps],
            "grid": [args.size, args.size], "land_cells": n_land, "all_cells": n_all,
            "kernel": args.kernel, "neigh": args.neigh, "sim_time_s": out.t,
            "cols_per_lane": args.cols_per_lane or 1,
            "all_cell_updates_per_s": n_all * args.steps / (ms * 1e-3),
            "mass_err_rel": emass,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": bps,
            "bytes_rule": "per LAND cell z,n,cr,h,hmax + mask read, h written; 1 B mask per other cell",
            # the instruction-issue ceiling beside the HBM one: wave64 VALU
            # instructions per launch (rocprofv3, profiles/) at one per CU per clock
            "issue": (None if valu is None else {
                "valu_insts_per_launch": valu, "issue_ms": valu / VALU_ISSUE_PER_S * 1e3,
                "frac": valu / VALU_ISSUE_PER_S / (ms_per_step * 1e-3),
                "rate_per_s": VALU_ISSUE_PER_S,
                # the same ceiling at the clock the counters measured during
                # the launch (256 CUs x clock)
                "clock_hz": clock_hz,
                "frac_at_clock": (None if clock_hz is None else
                                  valu / (256 * clock_hz) / (ms_per_step * 1e-3))}),
        },
    }
    if args.grid4096 and args.size != 4096 and prec == "f32" and args.neigh == 4:
        line["grid_4096"] = grid4096_bench()
    if args.sparse_size > 0 and prec == "f32" and args.neigh == 4:
        line["sparse_start"] = sparse_bench(prec, args.sparse_size)
        line["sparse_sweep"] = sparse_sweep_bench(prec, args.sparse_size)
    if args.ens_scenarios > 0 and prec == "f32" and args.neigh == 4:
        line["ensemble"] = ensemble_bench(g, prec, args.ens_scenarios, args.ens_alone)
    if args.day_scenarios > 0 and prec == "f32" and args.neigh == 4:
        line["reference_ensemble"] = day_ensemble_bench(prec, args.day_scenarios,
                                                        args.day_alone)
    if args.surrogate == "on" and prec == "f32" and args.neigh == 4:
        line["surrogate"] = surrogate_bench(g, args)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.cpu_seconds, prec)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
