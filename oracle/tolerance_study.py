"""Tolerance study: does a float step that is not bit-exact stay within the
FP32 bars of the rebuilt path (per-cell max|dh| <= 1e-5 m, relative
domain-volume drift <= 1e-6, time steps equal to 1 ulp)?

Each variant is this directory's oracle (fca_oracle.c, a restatement of
src/kernels/physics.h and src/common/model.c) compiled with one change:

* ``fma``: contraction on (-ffp-contract=fast, FMA hardware), the difference
  the bars themselves name as the only expected one;
* ``cbrtf``: hf^(2/3) as cbrtf(hf*hf) in float (glibc, within 1 ulp) instead
  of (float)cbrt((double)hf*hf) (physics.h:19-24);
* ``exp2log2``: hf^(2/3) as exp2f(log2f(hf)*2/3), the shape of a hardware
  log/exp evaluation;
* ``exp2log2_fma``: that, the slope root as sqrtf, and contraction on.

Each runs the same seeded case as the exact oracle and is compared with it:
max|dh| and max|dhmax| after the run, the relative volume drift of each run
(mass-balance identity, src/common/diag.c:5-14), and the largest distance in
ulps between the two dt_base sequences, step by step.

    python -m oracle.tolerance_study [--size 512] [--hours 6] [--K 4]

This is test infrastructure (tests/test_tolerance_study.py runs a small case);
it builds into a temporary directory and changes nothing in the tree.
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

BASE = ["-std=gnu11", "-O2", "-fno-fast-math", "-fPIC", "-D_ISOC11_SOURCE", "-shared", "-fopenmp",
        "-w", f"-I{os.path.join(ROOT, 'include')}", f"-I{HERE}"]
EXACT = ["-ffp-contract=off"]
FMA = ["-ffp-contract=fast", "-march=x86-64-v3"]
VARIANTS = {
    "exact": EXACT,
    "fma": FMA,
    "cbrtf": EXACT + ["-DORACLE_CBRT_SQ(hf,sq)=cbrtf((hf)*(hf))"],
    "exp2log2": EXACT + ["-DORACLE_CBRT_SQ(hf,sq)=exp2f(log2f(hf)*(2.0f/3.0f))"],
    "exp2log2_fma": FMA + ["-DORACLE_CBRT_SQ(hf,sq)=exp2f(log2f(hf)*(2.0f/3.0f))",
                           "-DORACLE_SQRT_SLOPE(x)=sqrtf(x)"],
}


def build(variant: str, outdir: str) -> str:
    """liboracle_f32.so of one variant in outdir/variant/."""
    d = os.path.join(outdir, variant)
    os.makedirs(d, exist_ok=True)
    out = os.path.join(d, "liboracle_f32.so")
    subprocess.run(["gcc", *BASE, *VARIANTS[variant], os.path.join(HERE, "fca_oracle.c"), "-lm",
                    "-o", out], check=True)
    return out


def run(lib: str, size: int, T_sim: float, K: int = 4, seed: int = 20261203):
    """One run of the 50 mm/h uniform-rain case from a dry start (the
    generator of the benchmark grids) on a size^2 grid."""
    from floodca_amd import abi as A
    from floodca_amd import synth
    from oracle.binding import Oracle

    a = A.abi("f32")
    g = synth.make_grid(size, size, seed=seed, dtype=a.dtype)
    cfg = A.default_cfg(a)
    A.set_dx(a, cfg, g.dx)
    cfg.hyeto = A.FCA_HYETO_UNIFORM
    cfg.T_sim = T_sim
    cfg.phys.K = cfg.neigh = K
    acc, log = Oracle("f32", path=lib).run(g, cfg, log_cap=1 << 20)
    drift = (acc.V + acc.V_sea + acc.V_bnd - acc.V_rain - acc.V0 - acc.clamp_err) / \
        max(acc.V_rain + acc.V0, 1e-30)
    return g, acc, np.array([r.dt_base for r in log], dtype=np.float32), drift


def ulp_distance(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Distance in float32 ulps of two arrays of positive floats."""
    ia = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    ib = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def compare(ref, var) -> dict:
    g0, acc0, db0, drift0 = ref
    g1, acc1, db1, drift1 = var
    n = min(db0.size, db1.size)
    u = ulp_distance(db0[:n], db1[:n])
    over = np.nonzero(u > 1)[0]
    return {
        "steps": int(acc1.step), "steps_exact": int(acc0.step),
        "dt_max_ulp": int(u.max()) if n else 0,
        "dt_steps_over_1ulp": int(over.size),
        "dt_first_over_1ulp": int(over[0]) if over.size else -1,
        "max_dh_m": float(np.abs(g1.h.astype(np.float64) - g0.h).max()),
        "max_dhmax_m": float(np.abs(g1.hmax.astype(np.float64) - g0.hmax).max()),
        "drift": float(drift1), "drift_exact": float(drift0),
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--hours", type=float, default=6.0)
    ap.add_argument("--K", type=int, default=4)
    ap.add_argument("--variants", default="fma,cbrtf,exp2log2,exp2log2_fma")
    args = ap.parse_args(argv)
    with tempfile.TemporaryDirectory() as tmp:
        ref = run(build("exact", tmp), args.size, args.hours * 3600.0, args.K)
        print(f"exact: {ref[1].step} steps, drift {ref[3]:.2e}")
        print("| variant | dt_base max ulp | steps > 1 ulp (first) | max abs dh (m) | "
              "max abs dhmax (m) | drift |")
        print("|---|---|---|---|---|---|")
        for v in args.variants.split(","):
            r = compare(ref, run(build(v, tmp), args.size, args.hours * 3600.0, args.K))
            print(f"| {v} | {r['dt_max_ulp']} | {r['dt_steps_over_1ulp']} of {r['steps']} "
                  f"({r['dt_first_over_1ulp']}) | {r['max_dh_m']:.2e} | {r['max_dhmax_m']:.2e} | "
                  f"{r['drift']:.2e} |", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
