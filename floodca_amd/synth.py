"""Seeded synthetic inputs for the FloodCA workloads.

One generator serves every size, so the small cases the oracle can finish and
the 16384^2 benchmark grid come from the same rules:

* DEM: diamond-square fractal (Hurst exponent H = 0.8) on a (2^k+1)^2 lattice,
  cropped to the grid, rescaled to ``relief`` metres (default 20 m) and added
  to a plane rising from 0 m at column 0 to 60 m at the last column.
* SEA: cells 4-connected to column 0 with z <= a threshold, the threshold
  chosen so the LAND fraction is ``land_frac`` (default 0.70), the same rule
  as the reference's west-edge sea mask (prep/floodca_prep/dem.py:132-148).
  SEA cells are the fixed-level open boundary.
* Land cover: 8 classes on seeded Voronoi patches (one seed per ~64x64 cells),
  mapped to Manning n in [0.02, 0.15]; C = 1 everywhere (an explicit C field:
  a missing C.bin would mean no rain, src/common/grid.c:203-208).
* Forcing: uniform 50 mm/h is ``Rday = 1.2`` m with ``hyeto = uniform``
  (F = dt/86400, src/common/rain.c:37-40).
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage
from scipy.spatial import cKDTree

from .abi import FCA_LAND, FCA_SEA
from .grid import Grid

MANNING_CLASSES = np.array([0.020, 0.025, 0.035, 0.045, 0.060, 0.080, 0.110, 0.150],
                           dtype=np.float64)
RDAY_50MM_H = 1.2  # m/day at 50 mm/h


def diamond_square(k: int, H: float, rng: np.random.Generator) -> np.ndarray:
    """Fractal surface on a (2^k+1)^2 lattice (float32, unnormalised)."""
    n = (1 << k) + 1
    z = np.zeros((n, n), dtype=np.float32)
    z[0::n - 1, 0::n - 1] = rng.standard_normal((2, 2), dtype=np.float32)
    step, scale, f = n - 1, 1.0, np.float32(2.0 ** (-H))
    while step > 1:
        half = step // 2
        # diamond: square centres from the four corners
        c = (z[0:-1:step, 0:-1:step] + z[0:-1:step, step::step] +
             z[step::step, 0:-1:step] + z[step::step, step::step]) * np.float32(0.25)
        z[half::step, half::step] = c + np.float32(scale) * rng.standard_normal(
            c.shape, dtype=np.float32)
        # square: edge midpoints from their (up to four) diamond neighbours
        for r0, c0 in ((0, half), (half, 0)):
            rows = np.arange(r0, n, step)
            cols = np.arange(c0, n, step)
            acc = np.zeros((rows.size, cols.size), dtype=np.float32)
            cnt = np.zeros((rows.size, cols.size), dtype=np.float32)
            for dr, dc in ((-half, 0), (half, 0), (0, -half), (0, half)):
                rr, cc = rows + dr, cols + dc
                okr, okc = (rr >= 0) & (rr < n), (cc >= 0) & (cc < n)
                sub = z[np.ix_(rr[okr], cc[okc])]
                acc[np.ix_(okr, okc)] += sub
                cnt[np.ix_(okr, okc)] += 1
            z[np.ix_(rows, cols)] = acc / cnt + np.float32(scale) * rng.standard_normal(
                acc.shape, dtype=np.float32)
        step, scale = half, scale * f
    return z


def _sea_from_west(z: np.ndarray, thr: float) -> np.ndarray:
    lab, _ = ndimage.label(z <= thr)
    ids = np.unique(lab[:, 0])
    ids = ids[ids != 0]
    return np.isin(lab, ids) if ids.size else np.zeros(z.shape, dtype=bool)


def sea_mask(z: np.ndarray, land_frac: float, tol: float = 0.004) -> np.ndarray:
    """West-connected sea with the threshold searched for LAND ~= land_frac."""
    target = 1.0 - land_frac
    stride = max(1, int(np.ceil(max(z.shape) / 1024)))
    zc = z[::stride, ::stride]
    lo, hi = float(z.min()) - 1.0, float(z.max())
    for _ in range(40):  # bisection on a coarse copy
        mid = 0.5 * (lo + hi)
        if _sea_from_west(zc, mid).mean() < target:
            lo = mid
        else:
            hi = mid
    thr = 0.5 * (lo + hi)
    sea = _sea_from_west(z, thr)
    if stride > 1:  # secant refinement at full resolution
        t0, f0 = thr, sea.mean() - target
        t1 = thr + 0.05 * (1 if f0 < 0 else -1)
        for _ in range(6):
            if abs(f0) <= tol:
                break
            s1 = _sea_from_west(z, t1)
            f1 = s1.mean() - target
            if abs(f1) <= abs(f0):
                sea = s1
            if f1 == f0:
                break
            t0, t1, f0 = t1, t1 - f1 * (t1 - t0) / (f1 - f0), f1
    return sea


def voronoi_classes(nrows: int, ncols: int, nclass: int, rng: np.random.Generator,
                    cells_per_seed: int = 4096) -> np.ndarray:
    """Class index per cell from seeded Voronoi patches (uint8)."""
    nseed = max(nclass, int(np.ceil(nrows * ncols / cells_per_seed)))
    pts = rng.random((nseed, 2)) * np.array([nrows, ncols])
    cls = rng.integers(0, nclass, nseed).astype(np.uint8)
    stride = max(1, int(np.ceil(max(nrows, ncols) / 2048)))
    r = np.arange(0, nrows, stride) + 0.5 * stride
    c = np.arange(0, ncols, stride) + 0.5 * stride
    rr, cc = np.meshgrid(r, c, indexing="ij")
    _, idx = cKDTree(pts).query(np.column_stack([rr.ravel(), cc.ravel()]), workers=-1)
    coarse = cls[idx].reshape(rr.shape)
    if stride > 1:
        coarse = np.repeat(np.repeat(coarse, stride, 0), stride, 1)
    return np.ascontiguousarray(coarse[:nrows, :ncols])


def make_dem(nrows: int, ncols: int, seed: int, H: float = 0.8, relief: float = 20.0,
             rise: float = 60.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    k = max(1, int(np.ceil(np.log2(max(nrows, ncols) - 1 if max(nrows, ncols) > 2 else 2))))
    f = diamond_square(k, H, rng)[:nrows, :ncols]
    lo, hi = float(f.min()), float(f.max())
    f = (f - lo) * np.float32(relief / (hi - lo if hi > lo else 1.0))
    plane = np.linspace(0.0, rise, ncols, dtype=np.float32)[None, :]
    return np.ascontiguousarray(f + plane, dtype=np.float32)


def make_grid(nrows: int, ncols: int, seed: int = 20261203, dx: float = 90.0,
              land_frac: float = 0.70, rday: float = RDAY_50MM_H, h0: float = 0.0,
              dtype=np.float32, relief: float = 20.0) -> Grid:
    """Synthetic grid: fractal DEM, west sea, Voronoi roughness, C = 1, uniform Rday,
    uniform initial depth h0 on LAND (0 = dry start)."""
    z = make_dem(nrows, ncols, seed, relief=relief)
    sea = sea_mask(z, land_frac) if land_frac < 1.0 else np.zeros(z.shape, dtype=bool)
    mask = np.where(sea, FCA_SEA, FCA_LAND).astype(np.uint8)
    rng = np.random.default_rng(seed + 1)
    cls = voronoi_classes(nrows, ncols, MANNING_CLASSES.size, rng)
    n = MANNING_CLASSES[cls].astype(dtype)
    h = np.where(mask == FCA_LAND, h0, 0.0).astype(dtype)
    return Grid(z=z.astype(dtype), mask=mask, n=n, C=np.ones((nrows, ncols), dtype=dtype),
                Rday=np.full((nrows, ncols), rday, dtype=dtype), h=h, dx=dx)


def point_inflow(g: Grid, peak_m3s: float = 500.0, dur_h: float = 8.0, seed: int = 7):
    """Turns g into the sparse-start case: no rain except one inland LAND cell whose
    Rday delivers a triangular hydrograph of the given peak over dur_h hours with
    hyeto = triangle (D_h = dur_h, phi = 0.25): the triangle's peak rate is 2/T,
    so V = peak*T/2 and Rday = V/(A*C) (src/common/rain.c:13-28).
    Returns (row, col)."""
    rng = np.random.default_rng(seed)
    land = np.argwhere(g.mask == FCA_LAND)
    nr, nc = g.shape
    inner = land[(land[:, 0] > nr // 4) & (land[:, 0] < 3 * nr // 4) &
                 (land[:, 1] > nc // 2)]
    r, c = inner[rng.integers(0, len(inner))] if len(inner) else land[len(land) // 2]
    T = dur_h * 3600.0
    g.Rday[...] = 0
    g.Rday[r, c] = 0.5 * peak_m3s * T / (g.dx * g.dx * float(g.C[r, c]))
    return int(r), int(c)
