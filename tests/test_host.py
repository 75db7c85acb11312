"""Host-side logic (CPU): grid-dir format, synthetic inputs, configuration
defaults against the reference's own, and the library's argument checks."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import make_case
from floodca_amd import abi as A
from floodca_amd import synth
from floodca_amd.grid import load_dir, read_bin, write_bin, write_dir
from floodca_amd.hip import HipError, Sim, backend
from oracle import Reference, ref_available


def test_grid_dir_roundtrip_and_defaults(tmp_path):
    g = synth.make_grid(37, 29, seed=4, h0=0.1)
    write_dir(str(tmp_path), g)
    back = load_dir(str(tmp_path))
    for k in ("z", "n", "C", "Rday", "h", "mask"):
        assert np.array_equal(getattr(back, k), getattr(g, k)), k
    assert np.array_equal(back.hmax, back.h) and not back.tmax.any()
    for f in ("n.bin", "C.bin", "Rday.bin", "h0.bin"):
        os.remove(tmp_path / f)
    d = load_dir(str(tmp_path))          # defaults of grid.c:197-223
    assert (d.n == np.float32(0.03)).all() and not d.C.any() and not d.Rday.any()
    assert not d.h.any()


def test_bin_json_header(tmp_path):
    arr = np.arange(12, dtype=np.float32).reshape(3, 4)
    write_bin(str(tmp_path / "x.bin"), arr, dx=90.0)
    assert np.array_equal(read_bin(str(tmp_path / "x.bin")), arr)


@pytest.mark.skipif(not ref_available("f32"), reason="oracle/_ref not built")
def test_reference_reads_our_grid_dir(tmp_path):
    """The reference's loader (grid.c:162-237) reads what write_dir writes."""
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "oracle", "_ref", "floodca_serial_f32")
    g = synth.make_grid(16, 12, seed=2)
    write_dir(str(tmp_path), g)
    import subprocess
    out = subprocess.run([exe, "--grid", str(tmp_path), "--steps", "3", "--tsim", "1e9",
                          "--no-output", "--csv", ""], capture_output=True, text=True,
                         check=True).stdout
    assert "done backend=serial steps=3" in out and "dx=90" in out


def test_synthetic_generator_properties():
    a = synth.make_grid(256, 256, seed=20261203)
    b = synth.make_grid(256, 256, seed=20261203)
    for k in ("z", "mask", "n"):
        assert np.array_equal(getattr(a, k), getattr(b, k))      # seeded
    land = (a.mask == A.FCA_LAND).mean()
    assert abs(land - 0.70) < 0.01
    from scipy import ndimage
    lab, _ = ndimage.label(a.mask == A.FCA_SEA)
    assert set(np.unique(lab[a.mask == A.FCA_SEA])) <= set(np.unique(lab[:, 0]))
    assert a.n.min() >= 0.02 - 1e-7 and a.n.max() <= 0.15 + 1e-7
    assert len(np.unique(a.n)) == 8 and (a.C == 1).all() and (a.Rday == np.float32(1.2)).all()
    assert 0.0 <= a.z.min() and a.z.max() <= 80.0 + 1e-3
    r, c = synth.point_inflow(a)
    assert a.Rday[r, c] == pytest.approx(0.5 * 500 * 8 * 3600 / 8100.0, rel=1e-6)
    assert np.count_nonzero(a.Rday) == 1


@pytest.mark.skipif(not ref_available("f32"), reason="oracle/_ref not built")
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_defaults_match_reference_config(prec):
    a = A.abi(prec)
    ref = Reference(prec)
    theirs = a.FcaRunCfg()
    ref.lib.fca_cfg_defaults.argtypes = [C.POINTER(a.FcaRunCfg)]
    ref.lib.fca_cfg_defaults(C.byref(theirs))
    ours = A.default_cfg(a)
    assert bytes(ours) == bytes(theirs)


@pytest.mark.parametrize("bad,msg", [
    (dict(K=5), "K must be 4 or 8"),
    (dict(alpha=0.0), "alpha*dx must be positive"),
])
def test_library_rejects_bad_configuration_without_a_device(bad, msg):
    a, g, cfg = make_case(8, 8)
    if "K" in bad:
        cfg.phys.K = bad["K"]
    if "alpha" in bad:
        cfg.phys.alpha = bad["alpha"]
    acc = a.FcaAcc()
    with pytest.raises(HipError, match=msg.replace("*", r"\*")):
        Sim(g, cfg, acc)


@pytest.mark.parametrize("prec,cpl,msg", [
    ("f32", 3, "cols_per_lane must be 0, 1 or 2"),
    ("f32", -1, "cols_per_lane must be 0, 1 or 2"),
    ("f64", 2, "cols_per_lane 2 is built for the float library only"),
])
def test_library_rejects_bad_columns_per_lane_without_a_device(prec, cpl, msg):
    a, g, cfg = make_case(8, 8, prec=prec)
    acc = a.FcaAcc()
    with pytest.raises(HipError, match=msg):
        Sim(g, cfg, acc, prec=prec, cols_per_lane=cpl)


def test_library_rejects_empty_grid():
    a = A.abi("f32")
    b = backend("f32")
    g = a.FcaGrid()
    cfg = A.default_cfg(a)
    acc = a.FcaAcc()
    assert not b.lib.fca_hip_sim_open(C.byref(g), C.byref(cfg), C.byref(acc), None)
    assert "empty grid" in b.error()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """No CPU fallback: without the built HIP library the binding raises
    instead of computing anything."""
    from floodca_amd import hip

    monkeypatch.setenv("FCA_HIP_LIB", str(tmp_path / "absent.so"))
    with pytest.raises(hip.HipError, match="not built"):
        hip.Backend("f32")


def test_noalloc_entry_point_rejects_before_touching_the_device():
    """fca_hip_run_noalloc checks its arguments before any device call: a grid
    beyond FCA_HIP_NOALLOC_MAX_CELLS, a NULL field and multi-day forcing are
    refused with a message (the addresses given are never dereferenced)."""
    fields = ("z", "h", "n", "C", "Rday", "hmax", "tmax", "mask")
    b = backend("f32")
    a = b.a
    cfg = A.default_cfg(a)
    A.set_dx(a, cfg, 90.0)
    acc = a.FcaAcc()
    with pytest.raises(HipError, match="exceeds the 4194304 cells"):
        b.run_noalloc({k: 1 << 20 for k in fields}, 2049, 2048, 90.0, cfg, acc)
    with pytest.raises(HipError, match="NULL"):
        b.run_noalloc({k: 0 for k in fields}, 16, 16, 90.0, cfg, acc)
    cfg.n_force_days = 2
    with pytest.raises(HipError, match="forcing"):
        b.run_noalloc({k: 1 << 20 for k in fields}, 16, 16, 90.0, cfg, acc)
    cfg.n_force_days = 0
    cfg.phys.K = 6
    with pytest.raises(HipError, match="K must be 4 or 8"):
        b.run_noalloc({k: 1 << 20 for k in fields}, 16, 16, 90.0, cfg, acc)
