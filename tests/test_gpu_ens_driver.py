"""The reference's GPU ensemble as one process on one MI355X (GPU):
oracle/_ref/floodca_ens_hip_{f32,f64} is floodca_amd/csrc/main_ens_hip.c
compiled with the reference's src/common, the counterpart of
src/hybrid/main_ens_mpi_cuda.c. Given a grid dir, a scenario table in the
reference's columns and one rain raster per base date, it must write the
count rasters count_010/025/050.bin with their JSON headers and one CSV row
per scenario in the reference's columns (ens_common.c:190-258), every field
equal to what the oracle computes for the same scenarios: each scenario the
grid's initial depth, rain_<date>.bin x lambda, a triangular storm of D_h and
phi, run to --tsim. Both modes (one ensemble session per scenario, and
batches of scenarios resident together) must write the same files. The
scenario reader's rules are checked on the way: a short row is skipped and
--first skips data rows. Built where the reference tree exists; skipped
otherwise."""
import json
import os
import subprocess

import numpy as np
import pytest

from floodca_amd import abi as A
from floodca_amd import ensemble as E
from floodca_amd import synth
from floodca_amd.grid import load_dir, write_bin, write_dir
from oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")
THR = (0.10, 0.25, 0.50)
TSIM = 4 * 3600.0

pytestmark = pytest.mark.gpu


def _inputs(tmp_path, prec):
    grid_dir, rain_dir = tmp_path / "grid", tmp_path / "rain"
    g = synth.make_grid(83, 67, seed=41, h0=0.01)
    write_dir(str(grid_dir), g)
    os.makedirs(rain_dir)
    dates = ["2005-09-22", "2005-09-23", "2005-09-25"]
    rng = np.random.default_rng(43)
    rains = {}
    for d in dates:
        r = (rng.uniform(0.02, 0.12) * (1.0 + 0.5 * rng.random(g.shape))).astype(np.float32)
        write_bin(str(rain_dir / f"rain_{d}.bin"), r, f"rain_{d}", g.dx)
        rains[d] = r
    csv = tmp_path / "scenarios.csv"
    rows = E.write_scenario_csv(str(csv), dates, 7, seed=20260915)
    with open(csv, "a", encoding="utf-8") as fp:
        fp.write("99,2005-09-22,1.0\n")  # a short row: not a scenario
    return grid_dir, rain_dir, csv, rows, rains


def _oracle(grid_dir, rows, rains, prec):
    a = A.abi(prec)
    o = Oracle(prec)
    base = load_dir(str(grid_dir), dtype=a.dtype)
    N = base.z.size
    counts = np.zeros(len(THR) * N, dtype=np.uint32)
    out = []
    for r in rows:
        g = base.copy()
        lam = a.dtype.type(float(r["lambda"]))
        g.Rday[...] = rains[r["base_date"]].astype(a.dtype)
        if lam != 1:
            g.Rday *= lam
        g.hmax[...] = g.h
        g.tmax[...] = 0
        cfg = A.default_cfg(a)
        A.set_dx(a, cfg, g.dx)
        cfg.hyeto = A.FCA_HYETO_TRIANGLE
        cfg.D_h = float(r["D_h"])
        cfg.phi = float(r["phi"])
        cfg.T_sim = TSIM
        acc, _ = o.run(g, cfg)
        fl, md = o.ens_count(g, THR, counts)
        out.append(dict(steps=acc.step, t=acc.t, V_rain=acc.V_rain, V_sea=acc.V_sea,
                        flooded=fl, max_depth=md, acc=acc))
    return out, counts.reshape(len(THR), *base.shape)


def _run(exe, grid_dir, rain_dir, csv, out_dir, ens_csv, extra=()):
    r = subprocess.run([exe, "--grid", str(grid_dir), "--scenario-file", str(csv), "--rain-dir",
                        str(rain_dir), "--nscen", "100", "--first", "1", "--tsim", str(TSIM),
                        "--out", str(out_dir), "--ens-csv", str(ens_csv), *extra],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    done = [ln for ln in r.stdout.splitlines() if ln.startswith("done backend=ens_hip")]
    assert done and "scen_per_hour=" in done[0], r.stdout[-2000:]
    return done[0]


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_ensemble_driver_matches_the_oracle(prec, tmp_path):
    exe = os.path.join(REF, f"floodca_ens_hip_{prec}")
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built (needs the reference tree at build time)")
    grid_dir, rain_dir, csv, rows, rains = _inputs(tmp_path, prec)
    rows = rows[1:]  # --first 1
    print(f"case ensemble driver {prec}: {len(rows)} scenarios", flush=True)
    ref, ref_counts = _oracle(grid_dir, rows, rains, prec)
    outs = {}
    for mode, extra in (("session", ()), ("batch", ("--batch", "4"))):
        out_dir, ens_csv = tmp_path / f"out_{mode}", tmp_path / f"ens_{mode}.csv"
        done = _run(exe, grid_dir, rain_dir, csv, out_dir, ens_csv, extra)
        assert f"scenarios={len(rows)} " in done
        lines = ens_csv.read_text().splitlines()
        assert lines[0] == ("scenario_id,base_date,lambda,D_h,phi,seed,backend,worker_rank,"
                            "device,steps,sim_time_s,wall_s,rain_volume_m3,sea_volume_m3,"
                            "mass_err_rel,flooded_cells_010,flooded_cells_025,"
                            "flooded_cells_050,max_depth_m")
        assert len(lines) == 1 + len(rows)
        fields = [ln.split(",") for ln in lines[1:]]
        for f, r, o in zip(fields, rows, ref):
            assert f[:6] == [str(r[k]) for k in E.SCENARIO_FIELDS]
            assert f[6:9] == ["ens_hip", "0", "gpu0"]
            assert int(f[9]) == o["steps"] > 10
            assert f[10] == f"{o['t']:.9g}" and float(f[10]) == TSIM
            assert float(f[11]) > 0.0
            assert float(f[12]) == pytest.approx(o["V_rain"], rel=1e-8)
            assert float(f[13]) == pytest.approx(o["V_sea"], rel=1e-8, abs=1e-6)
            assert abs(float(f[14])) < 1e-6
            assert [int(x) for x in f[15:18]] == list(o["flooded"])
            assert f[18] == f"{o['max_depth']:.6f}"
        for j, tag in enumerate(("010", "025", "050")):
            cnt = np.fromfile(out_dir / f"count_{tag}.bin", dtype=np.uint16).reshape(
                ref_counts.shape[1:])
            assert np.array_equal(cnt.astype(np.uint32), ref_counts[j]), tag
            hdr = json.loads((out_dir / f"count_{tag}.json").read_text())
            assert hdr["dtype"] == "uint16" and hdr["n_scenarios"] == len(rows)
            assert (hdr["nrows"], hdr["ncols"]) == cnt.shape
            assert hdr["threshold_m"] == THR[j] and hdr["name"] == f"count_{tag}"
        outs[mode] = [ln.split(",") for ln in lines[1:]]
    # the two modes agree on every field but the wall and the last digits of
    # the mass-balance residual (the FP64 balance terms are summed per
    # workgroup, and the batch walks strips of another height)
    for a_, b_ in zip(outs["session"], outs["batch"]):
        assert a_[:11] + a_[12:14] + a_[15:] == b_[:11] + b_[12:14] + b_[15:]
        assert abs(float(a_[14]) - float(b_[14])) < 1e-12
