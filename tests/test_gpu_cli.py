"""The library behind the reference's own command-line driver (GPU).

oracle/_ref/floodca_hip_{f32,f64} is floodca_amd/csrc/main_hip.c compiled
with the reference's src/common (CLI, grid-dir loading, warm start,
accumulators, CSV report) and linked against libfloodca_hip*.so -- the
executable a maintainer adds beside floodca_cuda. Run on a grid dir written
for the reference, it must leave the hmax and h_final rasters the oracle
computes from the same dir (FP32 bit for bit, FP64 within 1e-12 m), and
append a CSV row with the reference serial driver's columns and write the
reference's run manifest. Built by
__graft_entry__.build() where the reference tree exists; skipped otherwise.
"""
import os
import subprocess

import numpy as np
import pytest

from floodca_amd import abi as A
from floodca_amd import synth
from floodca_amd.grid import load_dir, read_bin, write_dir
from oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")

pytestmark = pytest.mark.gpu


def _exe(name):
    p = os.path.join(REF, name)
    if not os.path.exists(p):
        pytest.skip(f"{p} not built (needs the reference tree at build time)")
    return p


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_reference_driver_runs_the_hip_backend(prec, tmp_path):
    exe = _exe(f"floodca_hip_{prec}")
    grid_dir, out_dir = tmp_path / "grid", tmp_path / "out"
    csv = tmp_path / "runs.csv"
    g = synth.make_grid(193, 257, seed=21, h0=0.05)
    write_dir(str(grid_dir), g)
    print(f"case floodca_hip {prec}", flush=True)
    r = subprocess.run([exe, "--grid", str(grid_dir), "--tsim", "3600", "--hyeto", "uniform",
                        "--out", str(out_dir), "--csv", str(csv)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    done = [ln for ln in r.stdout.splitlines() if ln.startswith("done backend=hip")]
    assert done and "hbm_alg_gbps=" in done[0], r.stdout[-2000:]
    a = A.abi(prec)
    go = load_dir(str(grid_dir), dtype=a.dtype)
    cfg = A.default_cfg(a)
    A.set_dx(a, cfg, go.dx)
    cfg.hyeto = A.FCA_HYETO_UNIFORM
    cfg.T_sim = 3600.0
    acc, _ = Oracle(prec).run(go, cfg)
    assert f"steps={acc.step} " in done[0]
    hmax = read_bin(str(out_dir / "hmax.bin"))
    h = read_bin(str(out_dir / "h_final.bin"))
    if prec == "f32":
        assert np.array_equal(hmax, go.hmax) and np.array_equal(h, go.h)
    else:
        assert np.abs(hmax - go.hmax).max() <= 1e-12 and np.abs(h - go.h).max() <= 1e-12
    # the run manifest the reference's report writes (cli.c:124-151)
    import json
    man = json.loads((out_dir / "run_manifest.json").read_text())
    assert man["backend"] == "hip" and man["steps"] == acc.step and man["t"] == 3600.0
    assert abs(man["E_mass"]) < 1e-5 and "gfx" in man["gpu"]
    assert man["V_rain"] == pytest.approx(acc.V_rain, rel=1e-6)
    # the CSV row has the reference serial driver's columns
    csv_ref = tmp_path / "serial.csv"
    subprocess.run([_exe(f"floodca_serial_{prec}"), "--grid", str(grid_dir), "--steps", "2",
                    "--tsim", "1e9", "--out", str(tmp_path / "serial_out"), "--csv", str(csv_ref)],
                   capture_output=True, text=True, timeout=300, check=True)
    ours, theirs = csv.read_text().splitlines(), csv_ref.read_text().splitlines()
    assert ours[0] == theirs[0]
    assert len(ours[1].split(",")) == len(theirs[1].split(","))
    assert ",hip," in ours[1]


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_reference_driver_with_a_forcing_csv(prec, tmp_path):
    """--forcing through the reference's CSV loader (rain.c:120-200) and the
    library's day reloads: the same rasters as the oracle with the same day
    files."""
    exe = _exe(f"floodca_hip_{prec}")
    grid_dir, out_dir = tmp_path / "grid", tmp_path / "out"
    g = synth.make_grid(41, 53, seed=23)
    write_dir(str(grid_dir), g)
    rng = np.random.default_rng(3)
    for d in range(2):
        (rng.random(g.shape) * (0.05 + 0.1 * d)).astype(np.float32).tofile(tmp_path / f"r{d}.bin")
    (tmp_path / "forcing.csv").write_text(
        "day_index,date,file\n0,2021-07-01,r0.bin\n1,2021-07-02,r1.bin\n")
    print(f"case floodca_hip forcing {prec}", flush=True)
    r = subprocess.run([exe, "--grid", str(grid_dir), "--forcing", str(tmp_path / "forcing.csv"),
                        "--tsim", str(86400.0 * 2), "--hyeto", "uniform", "--out", str(out_dir)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    a = A.abi(prec)
    go = load_dir(str(grid_dir), dtype=a.dtype)
    cfg = A.default_cfg(a)
    A.set_dx(a, cfg, go.dx)
    cfg.hyeto = A.FCA_HYETO_UNIFORM
    cfg.T_sim = 86400.0 * 2
    cfg.n_force_days = 2
    for d in range(2):
        cfg.force_files[d].value = f"{tmp_path}/r{d}.bin".encode()
    acc, _ = Oracle(prec).run(go, cfg)
    assert acc.force_day == 1
    done = [ln for ln in r.stdout.splitlines() if ln.startswith("done backend=hip")]
    assert done and f"steps={acc.step} " in done[0], r.stdout[-2000:]
    hmax = read_bin(str(out_dir / "hmax.bin"))
    h = read_bin(str(out_dir / "h_final.bin"))
    if prec == "f32":
        assert np.array_equal(hmax, go.hmax) and np.array_equal(h, go.h)
    else:
        assert np.abs(hmax - go.hmax).max() <= 1e-12 and np.abs(h - go.h).max() <= 1e-12


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_reference_driver_warm_start_rain_lambda_k8(prec, tmp_path):
    """--warmstart (h and hmax from a float32 raster, cli.c:97-112), --rain
    with --lambda (grid.c:256-268), --neigh 8, --U 2, a closed boundary, a
    triangular storm and output events, all through the reference's CLI code,
    against the oracle on the same inputs."""
    exe = _exe(f"floodca_hip_{prec}")
    grid_dir, out_dir = tmp_path / "grid", tmp_path / "out"
    g = synth.make_grid(59, 71, seed=29)
    write_dir(str(grid_dir), g)
    rng = np.random.default_rng(4)
    h0 = np.where(g.mask == A.FCA_LAND, rng.random(g.shape) * 0.2, 0).astype(np.float32)
    h0.tofile(tmp_path / "warm.bin")
    rain = (rng.random(g.shape) * 0.4).astype(np.float32)
    rain.tofile(tmp_path / "rain.bin")
    args = ["--grid", str(grid_dir), "--warmstart", str(tmp_path / "warm.bin"),
            "--rain", str(tmp_path / "rain.bin"), "--lambda", "1.3", "--neigh", "8", "--U", "2",
            "--boundary", "closed", "--hyeto", "triangle", "--D_h", "1.5", "--phi", "0.3",
            "--out_every", "900", "--tsim", "5400", "--out", str(out_dir)]
    print(f"case floodca_hip warm start {prec}", flush=True)
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    a = A.abi(prec)
    go = load_dir(str(grid_dir), dtype=a.dtype)
    go.h[...] = h0.astype(a.dtype)
    go.hmax[...] = go.h
    go.Rday[...] = rain.astype(a.dtype) * a.dtype.type(1.3)
    cfg = A.default_cfg(a)
    A.set_dx(a, cfg, go.dx)
    cfg.phys.K = cfg.neigh = 8
    cfg.phys.U = 2
    cfg.phys.boundary = A.FCA_BND_CLOSED
    cfg.hyeto, cfg.D_h, cfg.phi = A.FCA_HYETO_TRIANGLE, 1.5, 0.3
    cfg.out_every, cfg.T_sim = 900.0, 5400.0
    cfg.lambda_ = 1.3
    acc, _ = Oracle(prec).run(go, cfg)
    done = [ln for ln in r.stdout.splitlines() if ln.startswith("done backend=hip")]
    assert done and f"steps={acc.step} " in done[0], r.stdout[-2000:]
    hmax = read_bin(str(out_dir / "hmax.bin"))
    h = read_bin(str(out_dir / "h_final.bin"))
    if prec == "f32":
        assert np.array_equal(hmax, go.hmax) and np.array_equal(h, go.h)
    else:
        assert np.abs(hmax - go.hmax).max() <= 1e-12 and np.abs(h - go.h).max() <= 1e-12
