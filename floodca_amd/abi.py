"""ctypes mirrors of the FloodCA C ABI (include/floodca_hip.h).

The layouts restate ``src/common/fca_types.h:34-159`` of the reference field by
field; ``real`` is float32 or float64 depending on the build. tests/test_abi.py
checks ``sizeof``/``offsetof`` of every field against probes compiled from the
header (both precisions) and against the reference's own header layout
recorded in tests/golden/abi_layout.json.

``Abi(prec)`` bundles the structure classes of one precision.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

FCA_LAND, FCA_SEA, FCA_WALL = 0, 1, 2
FCA_BND_OPEN, FCA_BND_CLOSED = 0, 1
FCA_HYETO_TRIANGLE, FCA_HYETO_UNIFORM = 0, 1
FCA_MAX_FORCE_DAYS = 64
FCA_HIP_FUSED, FCA_HIP_SPLIT = 0, 1
FCA_HIP_NO_LAND_FORM, FCA_HIP_NO_SKIP = 1, 2
FCA_HIP_MAX_CLASSES = 16


def _make(real):
    class FcaPhys(C.Structure):
        _fields_ = [(n, real) for n in ("dx", "A", "g", "alpha", "tau_h", "tau_l", "z_sea",
                                         "dt_max", "dt_min", "S_bnd", "h_flood")] + [
            ("U", C.c_int), ("K", C.c_int), ("boundary", C.c_int)]

    class FcaRunCfg(C.Structure):
        _fields_ = [
            ("phys", FcaPhys), ("T_sim", real), ("out_every", real),
            ("diag_every", C.c_int), ("progress_every", C.c_int), ("hyeto", C.c_int),
            ("D_h", real), ("phi", real), ("lambda_", real),
            ("steps", C.c_int), ("no_output", C.c_int), ("neigh", C.c_int),
            ("config_path", C.c_char * 512), ("grid_dir", C.c_char * 512),
            ("scenario", C.c_char * 128), ("scenario_file", C.c_char * 512),
            ("warmstart", C.c_char * 512), ("rain_file", C.c_char * 512),
            ("out_dir", C.c_char * 512), ("csv_path", C.c_char * 512),
            ("exp_id", C.c_char * 32),
            ("repeat", C.c_int), ("threads", C.c_int), ("ranks", C.c_int), ("gpus", C.c_int),
            ("block_x", C.c_int), ("block_y", C.c_int),
            ("halo_mode", C.c_char * 16), ("reduce_impl", C.c_char * 16),
            ("gpu_model", C.c_char * 64), ("backend", C.c_char * 32),
            ("forcing_csv", C.c_char * 512), ("n_force_days", C.c_int),
            ("force_files", (C.c_char * 768) * FCA_MAX_FORCE_DAYS),
        ]

    class FcaComm(C.Structure):
        _fields_ = [("ctx", C.c_void_p), ("rank", C.c_int), ("halo", C.c_void_p),
                    ("min_real", C.c_void_p), ("sum_double", C.c_void_p),
                    ("load_rday", C.c_void_p)]

    P = C.POINTER(real)

    class FcaGrid(C.Structure):
        _fields_ = [("nrows", C.c_int), ("ncols", C.c_int),
                    ("r0", C.c_int), ("r1", C.c_int), ("c0", C.c_int), ("c1", C.c_int),
                    ("edge_n", C.c_int), ("edge_s", C.c_int), ("edge_w", C.c_int),
                    ("edge_e", C.c_int), ("comm", C.POINTER(FcaComm)),
                    ("dx", real), ("x0", real), ("y0", real),
                    ("z", P), ("h", P), ("n", P), ("C", P), ("mask", C.POINTER(C.c_uint8)),
                    ("Rday", P), ("hmax", P), ("tmax", P), ("s", P), ("out", P),
                    ("Qx", P), ("Qy", P), ("Qd1", P), ("Qd2", P), ("Qopen", P)]

    class FcaAcc(C.Structure):
        _fields_ = [("V", C.c_double), ("V0", C.c_double), ("V_rain", C.c_double),
                    ("V_sea", C.c_double), ("V_bnd", C.c_double), ("clamp_err", C.c_double),
                    ("cap_active", C.c_long), ("cap_total", C.c_long),
                    ("dt_min_hits", C.c_long), ("t", C.c_double), ("step", C.c_int),
                    ("dt_base", real), ("force_day", C.c_int)]

    class FcaTimers(C.Structure):
        _fields_ = [(n, C.c_double) for n in (
            "wall_total", "t_init", "t_flux", "t_limiter", "t_update", "t_dt", "t_io",
            "t_h2d", "t_d2h", "t_halo", "t_allreduce", "bytes_per_step")] + [
            ("active_cells", C.c_long)]

    class FcaHipOpts(C.Structure):
        _fields_ = [("kernel", C.c_int), ("strip_rows", C.c_int), ("graph_batch", C.c_int),
                    ("record_steps", C.c_int), ("max_blocks", C.c_int),
                    ("flags", C.c_int), ("tiles_per_wave", C.c_int), ("cols_per_lane", C.c_int)]

    class FcaHipStepRec(C.Structure):
        _fields_ = [("t", C.c_double), ("F", C.c_double), ("dt_base", real), ("dt", real),
                    ("step", C.c_int), ("hit_event", C.c_int)]

    class FcaHipCrop(C.Structure):
        _fields_ = [("r0", C.c_int), ("c0", C.c_int), ("nrows", C.c_int), ("ncols", C.c_int),
                    ("rday", real), ("n_class", real * FCA_HIP_MAX_CLASSES)]

    return dict(FcaPhys=FcaPhys, FcaRunCfg=FcaRunCfg, FcaComm=FcaComm, FcaGrid=FcaGrid,
                FcaAcc=FcaAcc, FcaTimers=FcaTimers, FcaHipOpts=FcaHipOpts,
                FcaHipStepRec=FcaHipStepRec, FcaHipCrop=FcaHipCrop)


@dataclass(frozen=True)
class Abi:
    """Structure classes and dtypes of one precision ("f32" or "f64")."""

    prec: str

    def __post_init__(self):
        real = C.c_float if self.prec == "f32" else C.c_double
        object.__setattr__(self, "real", real)
        object.__setattr__(self, "dtype", np.dtype(np.float32 if self.prec == "f32"
                                                   else np.float64))
        for k, v in _make(real).items():
            object.__setattr__(self, k, v)


_ABIS: dict[str, Abi] = {}


def abi(prec: str = "f32") -> Abi:
    if prec not in ("f32", "f64"):
        raise ValueError(f"precision must be 'f32' or 'f64', not {prec!r}")
    if prec not in _ABIS:
        _ABIS[prec] = Abi(prec)
    return _ABIS[prec]


def default_cfg(a: Abi):
    """FcaRunCfg with the reference defaults (src/common/config.c:8-45)."""
    cfg = a.FcaRunCfg()
    p = cfg.phys
    p.dx = 30.0
    p.A = p.dx * p.dx
    p.g, p.alpha, p.tau_h, p.tau_l = 9.81, 0.5, 1e-3, 1e-5
    p.z_sea, p.dt_max, p.dt_min, p.S_bnd, p.h_flood = 0.0, 60.0, 0.01, 0.001, 0.25
    p.U, p.K, p.boundary = 1, 4, FCA_BND_OPEN
    cfg.T_sim, cfg.out_every, cfg.diag_every = 86400.0, 0.0, 100
    cfg.hyeto, cfg.D_h, cfg.phi, cfg.lambda_ = FCA_HYETO_TRIANGLE, 24.0, 0.5, 1.0
    cfg.steps, cfg.no_output, cfg.neigh = 0, 0, 4
    cfg.threads, cfg.block_x, cfg.block_y = 1, 16, 16
    cfg.backend = b"serial"
    cfg.exp_id = b"none"
    cfg.csv_path = b"results/bench.csv"
    cfg.out_dir = b"results/run"
    return cfg


def set_dx(a: Abi, cfg, dx: float) -> None:
    """Sets dx and A = dx*dx, the product rounded in the build precision
    (src/common/config.c:61-63, cli.c:400-401)."""
    d = a.dtype.type(dx)
    cfg.phys.dx = float(d)
    cfg.phys.A = float(a.dtype.type(d * d))
