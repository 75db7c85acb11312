"""Grid directories and host-side grids.

A *grid dir* holds ``dem.bin`` + ``dem.json`` (float32 raster, row-major,
``nrows``/``ncols``/``dx`` in the JSON header), ``mask.bin`` (uint8: 0 LAND,
1 SEA, 2 WALL) and optionally ``n.bin``, ``C.bin``, ``Rday.bin``, ``h.bin`` or
``h0.bin`` (float32). Missing files take the reference defaults of
``fca_grid_load_dir`` (src/common/grid.c:162-237): n = 0.03, C = 0, Rday = 0,
h = 0; then hmax = h and tmax = 0. The JSON header follows
``fca_write_json_header`` (src/common/io.c:34-54), the raster files are raw
little-endian arrays as in ``prep/floodca_prep/io_bin.py:20-45``.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass, field

import numpy as np

from .abi import Abi, FCA_LAND

_FIELDS = ("z", "h", "n", "C", "Rday", "hmax", "tmax")


def write_bin(path: str, arr: np.ndarray, name: str | None = None, dx: float = 30.0,
              x0: float = 0.0, y0: float = 0.0, extra: dict | None = None) -> None:
    """Writes ``<path>`` raw and ``<path minus .bin>.json`` with the header keys the
    reference reader parses (io.c:90-144)."""
    arr = np.ascontiguousarray(arr)
    arr.tofile(path)
    hdr = {"name": name or os.path.splitext(os.path.basename(path))[0],
           "nrows": int(arr.shape[0]), "ncols": int(arr.shape[1]), "dx": float(dx),
           "dtype": str(arr.dtype), "crs": "EPSG:32643", "x0": float(x0), "y0": float(y0),
           "nodata": -9999, "source": "floodca_amd"}
    if extra:
        hdr.update(extra)
    with open(os.path.splitext(path)[0] + ".json", "w", encoding="utf-8") as fp:
        json.dump(hdr, fp)
        fp.write("\n")


def read_header(json_path: str) -> dict:
    with open(json_path, encoding="utf-8") as fp:
        return json.load(fp)


def read_bin(path: str, shape=None, dtype=None) -> np.ndarray:
    """Reads a raster; shape/dtype default to the sibling JSON header."""
    if shape is None or dtype is None:
        hdr = read_header(os.path.splitext(path)[0] + ".json")
        shape = shape or (int(hdr["nrows"]), int(hdr["ncols"]))
        dtype = dtype or np.dtype(hdr.get("dtype", "float32"))
    return np.fromfile(path, dtype=dtype).reshape(shape)


@dataclass
class Grid:
    """Host arrays of one grid in the build precision (row-major)."""

    z: np.ndarray
    mask: np.ndarray
    n: np.ndarray
    C: np.ndarray
    Rday: np.ndarray
    h: np.ndarray
    hmax: np.ndarray = None
    tmax: np.ndarray = None
    dx: float = 30.0
    x0: float = 0.0
    y0: float = 0.0
    _keep: list = field(default_factory=list, repr=False)

    def __post_init__(self):
        if self.hmax is None:
            self.hmax = self.h.copy()
        if self.tmax is None:
            self.tmax = np.zeros_like(self.h)

    @property
    def shape(self):
        return self.z.shape

    @property
    def nrows(self):
        return self.z.shape[0]

    @property
    def ncols(self):
        return self.z.shape[1]

    @property
    def n_land(self) -> int:
        return int(np.count_nonzero(self.mask == FCA_LAND))

    def astype(self, dtype) -> "Grid":
        f = {k: np.ascontiguousarray(getattr(self, k), dtype=dtype) for k in _FIELDS}
        return Grid(mask=np.ascontiguousarray(self.mask, dtype=np.uint8), dx=self.dx,
                    x0=self.x0, y0=self.y0, **f)

    def copy(self) -> "Grid":
        return Grid(mask=self.mask.copy(), dx=self.dx, x0=self.x0, y0=self.y0,
                    **{k: getattr(self, k).copy() for k in _FIELDS})

    def fca_grid(self, a: Abi, scratch: bool = False):
        """FcaGrid pointing at these arrays (which must already be ``a.dtype``).
        scratch=True also allocates the CPU backends' s/out/face arrays, with
        Qopen zeroed, as fca_grid_alloc does (grid.c:60-94)."""
        nr, nc = self.shape
        for k in _FIELDS:
            arr = getattr(self, k)
            if arr.dtype != a.dtype or not arr.flags.c_contiguous or arr.shape != (nr, nc):
                raise ValueError(f"field {k}: need C-contiguous {a.dtype} of shape {(nr, nc)}")
        if self.mask.dtype != np.uint8 or not self.mask.flags.c_contiguous:
            raise ValueError("mask must be C-contiguous uint8")
        g = a.FcaGrid()
        g.nrows, g.ncols = nr, nc
        g.r0, g.r1, g.c0, g.c1 = 0, nr, 0, nc
        g.edge_n = g.edge_s = g.edge_w = g.edge_e = 1
        g.dx, g.x0, g.y0 = self.dx, self.x0, self.y0
        P = C.POINTER(a.real)
        for k in _FIELDS:
            setattr(g, k, getattr(self, k).ctypes.data_as(P))
        g.mask = self.mask.ctypes.data_as(C.POINTER(C.c_uint8))
        keep = []
        if scratch:
            def buf(n, fill=0.0):
                b = np.full(max(int(n), 1), fill, dtype=a.dtype)
                keep.append(b)
                return b.ctypes.data_as(P)
            N = nr * nc
            g.s, g.out = buf(N, 1.0), buf(N)
            g.Qx, g.Qy = buf(nr * (nc - 1)), buf((nr - 1) * nc)
            g.Qd1, g.Qd2 = buf((nr - 1) * (nc - 1)), buf((nr - 1) * (nc - 1))
            g.Qopen = buf(4 * N)
        self._keep = keep
        return g


def load_dir(path: str, dtype=np.float32) -> Grid:
    """Loads a grid dir with the reference defaults (grid.c:162-237)."""
    hdr = read_header(os.path.join(path, "dem.json"))
    nr, nc = int(hdr["nrows"]), int(hdr["ncols"])
    dx = float(hdr.get("dx", 30.0))

    def f32(name, default):
        p = os.path.join(path, name)
        if os.path.exists(p) and os.path.getsize(p) >= nr * nc * 4:
            return np.fromfile(p, dtype=np.float32, count=nr * nc).reshape(nr, nc).astype(dtype)
        return np.full((nr, nc), default, dtype=dtype)

    z = np.fromfile(os.path.join(path, "dem.bin"), dtype=np.float32,
                    count=nr * nc).reshape(nr, nc).astype(dtype)
    mask = np.fromfile(os.path.join(path, "mask.bin"), dtype=np.uint8,
                       count=nr * nc).reshape(nr, nc)
    hp = os.path.join(path, "h.bin")
    h = f32("h.bin", 0.0) if os.path.exists(hp) else f32("h0.bin", 0.0)
    return Grid(z=z, mask=mask, n=f32("n.bin", 0.03), C=f32("C.bin", 0.0),
                Rday=f32("Rday.bin", 0.0), h=h, dx=dx,
                x0=float(hdr.get("x0", 0.0)), y0=float(hdr.get("y0", 0.0)))


def write_dir(path: str, g: Grid, with_h: bool = True) -> None:
    """Writes a grid dir the reference loader reads (always an explicit C.bin:
    a missing C.bin means C = 0 and no rain, grid.c:203-208)."""
    os.makedirs(path, exist_ok=True)
    write_bin(os.path.join(path, "dem.bin"), g.z.astype(np.float32), "dem", g.dx, g.x0, g.y0)
    g.mask.astype(np.uint8).tofile(os.path.join(path, "mask.bin"))
    g.n.astype(np.float32).tofile(os.path.join(path, "n.bin"))
    g.C.astype(np.float32).tofile(os.path.join(path, "C.bin"))
    g.Rday.astype(np.float32).tofile(os.path.join(path, "Rday.bin"))
    if with_h:
        g.h.astype(np.float32).tofile(os.path.join(path, "h0.bin"))
