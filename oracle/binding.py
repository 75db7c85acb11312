"""ctypes bindings of the oracle libraries (see oracle/__init__.py)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from floodca_amd.abi import abi

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))


def ref_available(prec: str = "f32") -> bool:
    return os.path.exists(os.path.join(ORACLE_DIR, "_ref", f"libfca_ref_{prec}.so"))


def _face_sig(a):
    r = a.real
    return [r, r, r, C.c_uint8, r, r, r, C.c_uint8, r, r, r, C.c_int, r,
            C.POINTER(a.FcaPhys), C.POINTER(C.c_int)]


class _Lib:
    def __init__(self, path: str, prec: str):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run oracle/build_oracle.sh)")
        self.a = abi(prec)
        self.prec = prec
        self.lib = C.CDLL(path)

    def face_flux(self, fn, za, ha, na, ma, zb, hb, nb, mb, d, w, dt, K, A, phys):
        cap = C.c_int(0)
        q = fn(za, ha, na, ma, zb, hb, nb, mb, d, w, dt, K, A, C.byref(phys), C.byref(cap))
        return q, cap.value


class Oracle(_Lib):
    """The C restatement of the step (oracle/fca_oracle.c)."""

    def __init__(self, prec: str = "f32", path: str | None = None):
        """path: another build of fca_oracle.c (the tolerance study's variants);
        default the oracle itself."""
        super().__init__(path or os.path.join(ORACLE_DIR, f"liboracle_{prec}.so"), prec)
        a, L, r = self.a, self.lib, self.a.real
        L.or_face_flux.restype = r
        L.or_face_flux.argtypes = _face_sig(a)
        L.or_cbrt_sq.restype = r
        L.or_cbrt_sq.argtypes = [r]
        L.or_limiter_s.restype = r
        L.or_limiter_s.argtypes = [r, r, r]
        L.or_cell_dt.restype = r
        L.or_cell_dt.argtypes = [C.c_void_p] * 4 + [C.c_int] * 4 + [C.POINTER(a.FcaPhys)]
        L.or_compute_dt_base.restype = r
        L.or_compute_dt_base.argtypes = [C.POINTER(a.FcaGrid), C.POINTER(a.FcaPhys),
                                         C.POINTER(a.FcaAcc)]
        L.or_hyeto_integral.restype = C.c_double
        L.or_hyeto_integral.argtypes = [C.POINTER(a.FcaRunCfg), C.c_double, C.c_double]
        L.or_next_event_time.restype = C.c_double
        L.or_next_event_time.argtypes = [C.POINTER(a.FcaRunCfg), C.c_double]
        L.or_volume_land.restype = C.c_double
        L.or_volume_land.argtypes = [C.POINTER(a.FcaGrid), r]
        L.or_acc_init.restype = None
        L.or_acc_init.argtypes = [C.POINTER(a.FcaAcc), C.POINTER(a.FcaGrid), r]
        L.or_mass_error.restype = C.c_double
        L.or_mass_error.argtypes = [C.POINTER(a.FcaAcc)]
        L.or_step.restype = C.c_int
        L.or_step.argtypes = [C.POINTER(a.FcaGrid), C.POINTER(a.FcaRunCfg),
                              C.POINTER(a.FcaAcc), C.POINTER(a.FcaHipStepRec)]
        L.or_run.restype = C.c_int
        L.or_run.argtypes = [C.POINTER(a.FcaGrid), C.POINTER(a.FcaRunCfg), C.POINTER(a.FcaAcc),
                             C.POINTER(a.FcaHipStepRec), C.c_int, C.POINTER(C.c_int)]
        L.or_libm_range.restype = None
        L.or_libm_range.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.c_float, C.c_void_p]
        L.or_cbrt_sq_cr_mismatches.restype = C.c_long
        L.or_cbrt_sq_cr_mismatches.argtypes = [C.c_uint32, C.c_uint32]
        L.or_ens_count.restype = None
        L.or_ens_count.argtypes = [C.POINTER(a.FcaGrid), C.c_int, C.POINTER(C.c_double),
                                   C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_double)]

    def flux(self, *args):
        return self.face_flux(self.lib.or_face_flux, *args)

    def libm_range(self, kind: int, first: int, count: int, g: float = 9.81):
        """Reference float-mode libm values for the floats with bit patterns
        first..first+count-1 (kind 0 cbrt_sq, 1 sqrt, 2 critical speed)."""
        out = np.empty(count, dtype=np.float32)
        self.lib.or_libm_range(kind, first, count, g, out.ctypes.data)
        return out

    def cbrt_sq(self, hf):
        return self.lib.or_cbrt_sq(hf)

    def cell_dt(self, grid, r, c, phys):
        return self.lib.or_cell_dt(grid.z.ctypes.data, grid.h.ctypes.data, grid.n.ctypes.data,
                                   grid.mask.ctypes.data, r, c, grid.nrows, grid.ncols,
                                   C.byref(phys))

    def acc_init(self, grid, cfg):
        acc = self.a.FcaAcc()
        g = grid.fca_grid(self.a)
        self.lib.or_acc_init(C.byref(acc), C.byref(g), cfg.phys.A)
        return acc

    def run(self, grid, cfg, acc=None, log_cap: int = 0):
        """Runs or_run on grid in place; returns (acc, step records)."""
        if acc is None:
            acc = self.acc_init(grid, cfg)
        g = grid.fca_grid(self.a)
        log = (self.a.FcaHipStepRec * max(log_cap, 1))()
        n = C.c_int(0)
        rc = self.lib.or_run(C.byref(g), C.byref(cfg), C.byref(acc), log, log_cap, C.byref(n))
        if rc != 0:
            raise MemoryError("or_run failed")
        return acc, list(log)[:min(n.value, log_cap)]

    def ens_count(self, grid, thr, counts):
        g = grid.fca_grid(self.a)
        t = (C.c_double * len(thr))(*thr)
        fl = (C.c_long * len(thr))()
        md = C.c_double(0)
        self.lib.or_ens_count(C.byref(g), len(thr), t, counts.ctypes.data, fl, C.byref(md))
        return list(fl), md.value


class Reference(_Lib):
    """The reference's own src/common compiled here (parity build) + ref_shim.c."""

    def __init__(self, prec: str = "f32"):
        super().__init__(os.path.join(ORACLE_DIR, "_ref", f"libfca_ref_{prec}.so"), prec)
        a, L, r = self.a, self.lib, self.a.real
        L.ref_face_flux.restype = r
        L.ref_face_flux.argtypes = _face_sig(a)
        L.ref_cbrt_sq.restype = r
        L.ref_cbrt_sq.argtypes = [r]
        L.ref_limiter_s.restype = r
        L.ref_limiter_s.argtypes = [r, r, r]
        L.ref_cell_dt.restype = r
        L.ref_cell_dt.argtypes = [C.c_void_p] * 4 + [C.c_int] * 4 + [C.POINTER(a.FcaPhys)]
        L.fca_hyeto_integral.restype = C.c_double
        L.fca_hyeto_integral.argtypes = [C.POINTER(a.FcaRunCfg), C.c_double, C.c_double]
        L.fca_next_event_time.restype = C.c_double
        L.fca_next_event_time.argtypes = [C.POINTER(a.FcaRunCfg), C.c_double]
        L.fca_compute_dt_base.restype = r
        L.fca_compute_dt_base.argtypes = [C.POINTER(a.FcaGrid), C.POINTER(a.FcaPhys),
                                          C.POINTER(a.FcaAcc)]
        L.fca_acc_init.restype = None
        L.fca_acc_init.argtypes = [C.POINTER(a.FcaAcc), C.POINTER(a.FcaGrid), r]
        L.fca_forcing_load_csv.restype = C.c_int
        L.fca_forcing_load_csv.argtypes = [C.POINTER(a.FcaRunCfg), C.c_char_p]
        L.ref_run_logged.restype = C.c_int
        L.ref_run_logged.argtypes = [C.POINTER(a.FcaGrid), C.POINTER(a.FcaRunCfg),
                                     C.POINTER(a.FcaAcc), C.POINTER(C.c_double), C.c_int]

    def flux(self, *args):
        return self.face_flux(self.lib.ref_face_flux, *args)

    def cbrt_sq(self, hf):
        return self.lib.ref_cbrt_sq(hf)

    def cell_dt(self, grid, r, c, phys):
        return self.lib.ref_cell_dt(grid.z.ctypes.data, grid.h.ctypes.data, grid.n.ctypes.data,
                                    grid.mask.ctypes.data, r, c, grid.nrows, grid.ncols,
                                    C.byref(phys))

    def acc_init(self, grid, cfg):
        acc = self.a.FcaAcc()
        g = grid.fca_grid(self.a, scratch=True)
        self.lib.fca_acc_init(C.byref(acc), C.byref(g), cfg.phys.A)
        return acc

    def run(self, grid, cfg, acc=None, log_cap: int = 0):
        """Runs the reference's fca_step loop on grid in place.
        Returns (acc, log) with log rows (t_before, t_after, dt_base)."""
        if acc is None:
            acc = self.acc_init(grid, cfg)
        g = grid.fca_grid(self.a, scratch=True)  # after acc_init: it replaces the scratch
        log = np.zeros((max(log_cap, 1), 3))
        n = self.lib.ref_run_logged(C.byref(g), C.byref(cfg), C.byref(acc),
                                    log.ctypes.data_as(C.POINTER(C.c_double)), log_cap)
        return acc, log[:min(n, log_cap)]
