"""Shared test setup: the `gpu` marker and small case builders."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from floodca_amd import abi as A  # noqa: E402
from floodca_amd import synth  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def make_case(nr, nc, prec="f32", K=4, boundary=A.FCA_BND_OPEN, hyeto=A.FCA_HYETO_UNIFORM,
              T_sim=7200.0, steps=0, U=1, seed=11, h0=0.05, land_frac=0.7, rday=1.2,
              D_h=24.0, phi=0.5, out_every=0.0, walls=False):
    """Synthetic grid + config with the reference defaults and dx = 90 m."""
    a = A.abi(prec)
    lf = land_frac if min(nr, nc) > 3 else 1.0
    g = synth.make_grid(nr, nc, seed=seed, dtype=a.dtype, land_frac=lf, h0=h0, rday=rday)
    if walls:
        rng = np.random.default_rng(seed + 5)
        w = rng.random(g.shape) < 0.03
        g.mask[w & (g.mask == A.FCA_LAND)] = A.FCA_WALL
    cfg = A.default_cfg(a)
    A.set_dx(a, cfg, g.dx)
    cfg.phys.K = K
    cfg.neigh = K
    cfg.phys.U = U
    cfg.phys.boundary = boundary
    cfg.hyeto = hyeto
    cfg.T_sim = T_sim
    cfg.steps = steps
    cfg.D_h, cfg.phi, cfg.out_every = D_h, phi, out_every
    return a, g, cfg


@pytest.fixture
def case():
    return make_case


GOLDEN = os.path.join(ROOT, "tests", "golden")
GOLDEN_CASES = ["g33x17_k4_open_uniform", "g33x17_k8_closed_uniform",
                "g57x41_k4_open_triangle_walls", "g57x41_k8_open_steps_u3", "g7x5_k4_open_dry",
                "g128x96_k4_open_uniform"]


def load_golden(name, prec):
    """A reference run from tests/golden: (abi, input grid, cfg, expected dict)."""
    import json

    from floodca_amd.grid import Grid

    a = A.abi(prec)
    f = np.load(os.path.join(GOLDEN, f"{name}_{prec}.npz"), allow_pickle=False)
    g = Grid(z=f["z"], mask=f["mask"], n=f["n"], C=f["C"], Rday=f["Rday"], h=f["h0"].copy(),
             hmax=f["hmax0"].copy(), tmax=f["tmax0"].copy(), dx=float(f["dx"]))
    cfg = A.default_cfg(a)
    for k, v in json.loads(str(f["cfg"])).items():
        if k.startswith("phys."):
            setattr(cfg.phys, k[5:], v)
        else:
            setattr(cfg, k, v)
    exp = {"h": f["h"], "hmax": f["hmax"], "tmax": f["tmax"], "log": f["log"],
           "acc": json.loads(str(f["acc"]))}
    return a, g, cfg, exp


def emu_lib(prec):
    """ctypes handle of the host emulation of the fused walk (tests/emu)."""
    import ctypes as C

    a = A.abi(prec)
    L = C.CDLL(os.path.join(ROOT, "tests", "emu", f"libwalk_emu_{prec}.so"))
    L.emu_run.restype = C.c_int
    L.emu_run.argtypes = [C.POINTER(a.FcaGrid), C.POINTER(a.FcaRunCfg), C.POINTER(a.FcaAcc),
                          C.c_int, C.POINTER(a.FcaHipStepRec), C.c_int, C.POINTER(C.c_int)]
    return L


def emu_run(grid, cfg, acc, prec, strip_rows=0, cap=100000, land_form=True, forms=None,
            skip=True, skipped=None, width=64):
    """Runs the emulated fused walk; land_form=False keeps every row in the
    general form, skip=False walks every tile, width=128 emulates two columns
    per work-item (tiles of 122 columns). forms, if a list, receives the rows
    walked in the general, the LAND and the no-LAND form; skipped, if a list,
    the number of quiet tiles skipped."""
    import ctypes as C

    a = A.abi(prec)
    L = emu_lib(prec)
    g = grid.fca_grid(a)
    rec = (a.FcaHipStepRec * cap)()
    n = C.c_int()
    L.emu_set_land_form(-1 if land_form else 0)
    L.emu_set_skip(1 if skip else 0)
    L.emu_set_width(width)
    try:
        assert L.emu_run(C.byref(g), C.byref(cfg), C.byref(acc), strip_rows, rec, cap,
                         C.byref(n)) == 0
    finally:
        L.emu_set_land_form(-1)
        L.emu_set_skip(1)
        L.emu_set_width(64)
    if skipped is not None:
        L.emu_skipped_tiles.restype = C.c_long
        skipped[:] = [L.emu_skipped_tiles()]
    if forms is not None:
        gen, land, noland = C.c_long(), C.c_long(), C.c_long()
        L.emu_form_rows(C.byref(gen), C.byref(land), C.byref(noland))
        forms[:] = [gen.value, land.value, noland.value]
    return acc, list(rec)[:min(n.value, cap)]
