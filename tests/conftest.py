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
