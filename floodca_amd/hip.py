"""ctypes binding of the HIP backend (include/floodca_hip.h).

Loads the in-tree ``libfloodca_hip.so`` (float) or ``libfloodca_hip_f64.so``
(double). There is no fallback: if the library or a GPU is missing, calls
raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .abi import FCA_HIP_FUSED, abi
from .build import lib_path, load_hip_library

_LIBS: dict[str, "Backend"] = {}


class HipError(RuntimeError):
    pass


class Backend:
    """The C-ABI of one precision."""

    def __init__(self, prec: str = "f32"):
        path = lib_path(prec)
        if not os.path.exists(path):
            raise HipError(f"{path} is not built: run python -m floodca_amd.build")
        self.prec = prec
        self.a = a = abi(prec)
        self.lib = L = load_hip_library(path)
        G, Cf, Ac, Tm = (C.POINTER(a.FcaGrid), C.POINTER(a.FcaRunCfg), C.POINTER(a.FcaAcc),
                         C.POINTER(a.FcaTimers))
        P = C.POINTER(a.real)
        sig = {
            "fca_hip_run": (C.c_int, [G, Cf, Ac, Tm]),
            "fca_hip_device_name": (None, [C.c_char_p, C.c_size_t]),
            "fca_hip_select_device": (C.c_int, [C.c_int]),
            "fca_hip_last_error": (C.c_char_p, []),
            "fca_hip_sim_open": (C.c_void_p, [G, Cf, Ac, C.POINTER(a.FcaHipOpts)]),
            "fca_hip_sim_advance": (C.c_int, [C.c_void_p, C.c_int]),
            "fca_hip_sim_sync": (C.c_int, [C.c_void_p]),
            "fca_hip_sim_read": (C.c_int, [C.c_void_p, P, P, P, Ac]),
            "fca_hip_sim_steplog": (C.c_int, [C.c_void_p, C.POINTER(a.FcaHipStepRec), C.c_int]),
            "fca_hip_sim_run": (C.c_int, [C.c_void_p]),
            "fca_hip_sim_stream": (C.c_void_p, [C.c_void_p]),
            "fca_hip_sim_close": (None, [C.c_void_p]),
            "fca_hip_ens_open": (C.c_void_p, [G, Cf, C.c_int, C.POINTER(C.c_double)]),
            "fca_hip_ens_run": (C.c_int, [C.c_void_p, G, Cf, Ac, Tm, C.POINTER(C.c_long),
                                          C.POINTER(C.c_double), C.c_void_p, C.c_void_p,
                                          C.c_int]),
            "fca_hip_ens_counts": (C.c_int, [C.c_void_p, C.c_void_p]),
            "fca_hip_ens_counts_u32": (C.c_int, [C.c_void_p, C.c_void_p]),
            "fca_hip_ens_close": (None, [C.c_void_p]),
            "fca_hip_batch_open": (C.c_void_p, [C.c_int, C.POINTER(G), C.POINTER(Cf), Ac,
                                                C.POINTER(a.FcaHipOpts)]),
            "fca_hip_batch_nscen": (C.c_int, [C.c_void_p]),
            "fca_hip_batch_run": (C.c_int, [C.c_void_p]),
            "fca_hip_batch_read": (C.c_int, [C.c_void_p, C.c_int, P, P, P, Ac]),
            "fca_hip_batch_steplog": (C.c_int, [C.c_void_p, C.c_int,
                                                C.POINTER(a.FcaHipStepRec), C.c_int]),
            "fca_hip_batch_flooded": (C.c_int, [C.c_void_p, C.c_int, C.c_int,
                                                C.POINTER(C.c_double), C.POINTER(C.c_long),
                                                C.POINTER(C.c_double)]),
            "fca_hip_batch_wall": (C.c_double, [C.c_void_p]),
            "fca_hip_batch_device_fields": (C.c_int, [C.c_void_p, C.c_int,
                                                      C.POINTER(C.c_void_p),
                                                      C.POINTER(C.c_void_p),
                                                      C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "fca_hip_batch_status": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                               C.POINTER(C.c_double)]),
            "fca_hip_batch_close": (None, [C.c_void_p]),
            "fca_hip_base_open": (C.c_void_p, [G, C.c_void_p]),
            "fca_hip_base_close": (None, [C.c_void_p]),
            "fca_hip_batch_open_crops": (C.c_void_p, [C.c_void_p, C.c_int,
                                                      C.POINTER(a.FcaHipCrop), C.POINTER(Cf),
                                                      C.POINTER(a.FcaHipOpts)]),
            "fca_hip_run_noalloc": (C.c_int, [G, Cf, Ac, Tm, C.c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args

    def error(self) -> str:
        return (self.lib.fca_hip_last_error() or b"").decode(errors="replace")

    def check(self, rc, what):
        if rc is None or (isinstance(rc, int) and rc < 0):
            raise HipError(f"{what}: {self.error() or 'failed'} (rc={rc})")
        return rc

    def device_name(self) -> str:
        buf = C.create_string_buffer(128)
        self.lib.fca_hip_device_name(buf, 128)
        return buf.value.decode()

    def run(self, grid, cfg, acc):
        """fca_hip_run on grid in place; returns the FcaTimers."""
        tm = self.a.FcaTimers()
        tm.t_h2d = tm.t_d2h = tm.t_halo = tm.t_allreduce = tm.bytes_per_step = -1
        g = grid.fca_grid(self.a)
        self.check(self.lib.fca_hip_run(C.byref(g), C.byref(cfg), C.byref(acc), C.byref(tm)),
                   "fca_hip_run")
        return tm


    def run_noalloc(self, fields: dict, nrows: int, ncols: int, dx: float, cfg, acc,
                    stream=None):
        """fca_hip_run_noalloc on fields already in device memory: ``fields``
        maps z, h, n, C, mask, Rday, hmax, tmax to device addresses (ints, e.g.
        a tensor's data_ptr()). h, hmax and tmax receive the results in place;
        returns the FcaTimers."""
        a = self.a
        g = a.FcaGrid()
        g.nrows, g.ncols, g.dx = nrows, ncols, dx
        g.r0, g.r1, g.c0, g.c1 = 0, nrows, 0, ncols
        P = C.POINTER(a.real)
        for k in ("z", "h", "n", "C", "Rday", "hmax", "tmax"):
            setattr(g, k, C.cast(C.c_void_p(fields[k]), P))
        g.mask = C.cast(C.c_void_p(fields["mask"]), C.POINTER(C.c_uint8))
        tm = self.a.FcaTimers()
        tm.t_h2d = tm.t_d2h = tm.t_halo = tm.t_allreduce = tm.bytes_per_step = -1
        self.check(self.lib.fca_hip_run_noalloc(C.byref(g), C.byref(cfg), C.byref(acc),
                                                C.byref(tm), stream), "fca_hip_run_noalloc")
        return tm


def backend(prec: str = "f32") -> Backend:
    if prec not in _LIBS:
        _LIBS[prec] = Backend(prec)
    return _LIBS[prec]


def opts(prec="f32", kernel=FCA_HIP_FUSED, strip_rows=0, graph_batch=0, record_steps=0,
         max_blocks=0, flags=0, tiles_per_wave=0, cols_per_lane=0):
    o = abi(prec).FcaHipOpts()
    o.kernel, o.strip_rows, o.graph_batch, o.record_steps = (kernel, strip_rows, graph_batch,
                                                             record_steps)
    o.max_blocks, o.flags, o.tiles_per_wave = max_blocks, flags, tiles_per_wave
    o.cols_per_lane = cols_per_lane
    return o


class Sim:
    """A grid resident in HBM (fca_hip_sim_*)."""

    def __init__(self, grid, cfg, acc, prec="f32", **kw):
        self.b = backend(prec)
        self.a = self.b.a
        self.grid, self.cfg = grid, cfg
        self.opts = opts(prec, **kw)
        g = grid.fca_grid(self.a)
        self.h = self.b.lib.fca_hip_sim_open(C.byref(g), C.byref(cfg), C.byref(acc),
                                             C.byref(self.opts))
        if not self.h:
            raise HipError(f"fca_hip_sim_open: {self.b.error()}")

    def advance(self, n: int) -> int:
        return self.b.check(self.b.lib.fca_hip_sim_advance(self.h, n), "advance")

    def sync(self):
        self.b.check(self.b.lib.fca_hip_sim_sync(self.h), "sync")

    def run(self):
        self.b.check(self.b.lib.fca_hip_sim_run(self.h), "run")

    def read(self, grid=None):
        """Copies h, hmax, tmax into grid (default: the grid given at open) and
        returns the FcaAcc."""
        grid = grid or self.grid
        acc = self.a.FcaAcc()
        P = C.POINTER(self.a.real)
        self.b.check(self.b.lib.fca_hip_sim_read(self.h, grid.h.ctypes.data_as(P),
                                                 grid.hmax.ctypes.data_as(P),
                                                 grid.tmax.ctypes.data_as(P), C.byref(acc)),
                     "read")
        return acc

    def steplog(self, cap: int):
        buf = (self.a.FcaHipStepRec * max(cap, 1))()
        n = self.b.check(self.b.lib.fca_hip_sim_steplog(self.h, buf, cap), "steplog")
        return list(buf)[:n]

    @property
    def stream(self):
        return self.b.lib.fca_hip_sim_stream(self.h)

    def close(self):
        if self.h:
            self.b.lib.fca_hip_sim_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


POLL_FN = C.CFUNCTYPE(None, C.c_void_p)


class Ensemble:
    """fca_hip_ens_* session: static fields resident, one scenario per run()."""

    def __init__(self, grid, cfg, thresholds=(0.10, 0.25, 0.50), prec="f32"):
        self.b = backend(prec)
        self.a = self.b.a
        self.nthr = len(thresholds)
        self.N = grid.nrows * grid.ncols
        thr = (C.c_double * self.nthr)(*thresholds)
        g = grid.fca_grid(self.a)
        self.h = self.b.lib.fca_hip_ens_open(C.byref(g), C.byref(cfg), self.nthr, thr)
        if not self.h:
            raise HipError(f"fca_hip_ens_open: {self.b.error()}")

    def run(self, grid, cfg, acc, poll=None, poll_every=0):
        tm = self.a.FcaTimers()
        fl = (C.c_long * self.nthr)()
        md = C.c_double(0)
        g = grid.fca_grid(self.a)
        cb = POLL_FN(poll) if poll else None
        self.b.check(self.b.lib.fca_hip_ens_run(self.h, C.byref(g), C.byref(cfg), C.byref(acc),
                                                C.byref(tm), fl, C.byref(md),
                                                C.cast(cb, C.c_void_p) if cb else None, None,
                                                poll_every), "ens_run")
        return list(fl), md.value, tm

    def counts(self, u32=True):
        if u32:
            out = np.zeros(self.nthr * self.N, dtype=np.uint32)
            self.b.check(self.b.lib.fca_hip_ens_counts_u32(self.h, out.ctypes.data), "counts")
        else:
            out = np.zeros(self.nthr * self.N, dtype=np.uint16)
            self.b.check(self.b.lib.fca_hip_ens_counts(self.h, out.ctypes.data), "counts")
        return out.reshape(self.nthr, -1)

    def close(self):
        if self.h:
            self.b.lib.fca_hip_ens_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBase:
    """fca_hip_base_*: a grid's bed, mask and land-cover classes resident on
    the device, the source of crop scenarios (Batch.from_crops)."""

    def __init__(self, grid, cls, prec="f32"):
        self.b = backend(prec)
        a = self.b.a
        self.cls = np.ascontiguousarray(cls, dtype=np.uint8)
        if self.cls.shape != grid.shape:
            raise ValueError("one class per cell")
        self._keep = grid if grid.z.dtype == a.dtype else grid.astype(a.dtype)
        self._g = self._keep.fca_grid(a)
        self.h = self.b.lib.fca_hip_base_open(C.byref(self._g), self.cls.ctypes.data)
        if not self.h:
            raise HipError(f"fca_hip_base_open: {self.b.error()}")
        self.shape = grid.shape

    def close(self):
        if self.h:
            self.b.lib.fca_hip_base_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """fca_hip_batch_*: scenarios resident at once, advanced together."""

    def __init__(self, grids, cfgs, accs, prec="f32", **kw):
        self.b = backend(prec)
        self.a = a = self.b.a
        n = len(grids)
        if not (n == len(cfgs) == len(accs)) or n == 0:
            raise ValueError("need one cfg and one acc per grid")
        self.grids = list(grids)
        self.opts = opts(prec, **kw)
        self._g = [g.fca_grid(a) for g in grids]
        self._gp = (C.POINTER(a.FcaGrid) * n)(*[C.pointer(x) for x in self._g])
        self._c = list(cfgs)
        self._cp = (C.POINTER(a.FcaRunCfg) * n)(*[C.pointer(x) for x in self._c])
        self._acc = (a.FcaAcc * n)(*accs)
        self.h = self.b.lib.fca_hip_batch_open(n, self._gp, self._cp, self._acc,
                                               C.byref(self.opts))
        if not self.h:
            raise HipError(f"fca_hip_batch_open: {self.b.error()}")
        self.n = n

    @classmethod
    def from_crops(cls, base: "DeviceBase", crops, cfgs, prec="f32", **kw):
        """A batch of crop scenarios built on the device from base: crops is
        a list of FcaHipCrop (or dicts with r0, c0, nrows, ncols, rday,
        n_class), cfgs one FcaRunCfg each. The scenarios hold no host grids,
        so read() needs a grid to copy into."""
        self = cls.__new__(cls)
        self.b = backend(prec)
        self.a = a = self.b.a
        n = len(crops)
        if n == 0 or n != len(cfgs):
            raise ValueError("need one cfg per crop")
        self.grids = [None] * n
        self.opts = opts(prec, **kw)
        cr = (a.FcaHipCrop * n)()
        for i, c in enumerate(crops):
            if isinstance(c, dict):
                cr[i].r0, cr[i].c0 = c["r0"], c["c0"]
                cr[i].nrows, cr[i].ncols = c["nrows"], c["ncols"]
                cr[i].rday = c["rday"]
                for k, v in enumerate(c["n_class"]):
                    cr[i].n_class[k] = v
            else:
                cr[i] = c
        self._crops = cr
        self._base = base
        self._c = list(cfgs)
        self._cp = (C.POINTER(a.FcaRunCfg) * n)(*[C.pointer(x) for x in self._c])
        self.h = self.b.lib.fca_hip_batch_open_crops(base.h, n, cr, self._cp,
                                                     C.byref(self.opts))
        if not self.h:
            raise HipError(f"fca_hip_batch_open_crops: {self.b.error()}")
        self.n = n
        return self

    def run(self):
        self.b.check(self.b.lib.fca_hip_batch_run(self.h), "batch run")

    @property
    def wall(self) -> float:
        return self.b.lib.fca_hip_batch_wall(self.h)

    def read(self, i, grid=None):
        """Copies scenario i's h, hmax, tmax into grid (default: its own) and
        returns its FcaAcc."""
        grid = grid or self.grids[i]
        if grid is None:
            raise ValueError("a crop scenario has no host grid: pass one to read into")
        acc = self.a.FcaAcc()
        P = C.POINTER(self.a.real)
        self.b.check(self.b.lib.fca_hip_batch_read(self.h, i, grid.h.ctypes.data_as(P),
                                                   grid.hmax.ctypes.data_as(P),
                                                   grid.tmax.ctypes.data_as(P), C.byref(acc)),
                     "batch read")
        return acc

    def steplog(self, i, cap):
        buf = (self.a.FcaHipStepRec * max(cap, 1))()
        n = self.b.check(self.b.lib.fca_hip_batch_steplog(self.h, i, buf, cap), "steplog")
        return list(buf)[:n]

    def steps(self, i) -> int:
        """Steps scenario i has taken (its controller record)."""
        st = C.c_int()
        t = C.c_double()
        self.b.check(self.b.lib.fca_hip_batch_status(self.h, i, C.byref(st), C.byref(t)),
                     "status")
        return st.value

    def device_fields(self, i):
        """(hmax, mask, nr, nc): device addresses of scenario i's fields."""
        hm, mk, nr, nc = C.c_void_p(), C.c_void_p(), C.c_int(), C.c_int()
        self.b.check(self.b.lib.fca_hip_batch_device_fields(self.h, i, C.byref(hm), C.byref(mk),
                                                            C.byref(nr), C.byref(nc)),
                     "device fields")
        return hm.value, mk.value, nr.value, nc.value

    def flooded(self, i, thresholds=(0.10, 0.25, 0.50)):
        thr = (C.c_double * len(thresholds))(*thresholds)
        fl = (C.c_long * len(thresholds))()
        md = C.c_double(0)
        self.b.check(self.b.lib.fca_hip_batch_flooded(self.h, i, len(thresholds), thr, fl,
                                                      C.byref(md)), "flooded")
        return list(fl), md.value

    def close(self):
        if self.h:
            self.b.lib.fca_hip_batch_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
