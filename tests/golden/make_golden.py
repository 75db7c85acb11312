"""Writes the golden fixtures under tests/golden/ from the reference build.

Each fixture is one small run of the reference's own fca_step loop (the
reference's src/common compiled here into oracle/_ref, contraction off): the
inputs (z, n, C, Rday, mask, h0), the configuration, and the reference's
outputs (h, hmax, tmax, the per-step t/dt_base log and the accumulators).
The GPU box has no reference tree, so these vectors carry the reference's
results there; tests/test_golden.py checks the oracle against them on the CPU
and tests/test_gpu_golden.py checks the HIP backend against them on the GPU.

Also writes abi_layout.json: sizeof/offsetof of every ABI struct field
compiled against the reference's own src/common/fca_types.h.

Run (in a tree with the reference present):  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from conftest import make_case  # noqa: E402
from floodca_amd import abi as A  # noqa: E402
from oracle import Reference  # noqa: E402

REF = os.environ.get("REF", "/root/reference")

# name: (nr, nc, kwargs)
CASES = {
    "g33x17_k4_open_uniform": (33, 17, dict(K=4, T_sim=5400.0)),
    "g33x17_k8_closed_uniform": (33, 17, dict(K=8, boundary=A.FCA_BND_CLOSED, T_sim=5400.0)),
    "g57x41_k4_open_triangle_walls": (57, 41, dict(K=4, hyeto=A.FCA_HYETO_TRIANGLE, D_h=1.5,
                                                   phi=0.3, T_sim=7200.0, rday=30.0,
                                                   walls=True)),
    "g57x41_k8_open_steps_u3": (57, 41, dict(K=8, U=3, steps=250, T_sim=1e9,
                                             out_every=1800.0)),
    "g7x5_k4_open_dry": (7, 5, dict(K=4, h0=0.0, T_sim=3600.0)),
    "g128x96_k4_open_uniform": (128, 96, dict(K=4, T_sim=3600.0, seed=3)),
}

CFG_KEYS = ("T_sim", "out_every", "hyeto", "D_h", "phi", "steps", "neigh")
PHYS_KEYS = ("dx", "A", "g", "alpha", "tau_h", "tau_l", "z_sea", "dt_max", "dt_min", "S_bnd",
             "U", "K", "boundary")


def cfg_dict(cfg):
    d = {k: getattr(cfg, k) for k in CFG_KEYS}
    d.update({f"phys.{k}": getattr(cfg.phys, k) for k in PHYS_KEYS})
    return d


def make_fixture(name, nr, nc, kw, prec):
    a, g, cfg = make_case(nr, nc, prec=prec, **kw)
    inp = g.copy()
    ref = Reference(prec)
    acc, log = ref.run(g, cfg, log_cap=100000)
    accd = {k: getattr(acc, k) for k in ("V", "V0", "V_rain", "V_sea", "V_bnd", "clamp_err",
                                          "cap_active", "cap_total", "dt_min_hits", "t",
                                          "step", "dt_base", "force_day")}
    path = os.path.join(HERE, f"{name}_{prec}.npz")
    np.savez_compressed(path, z=inp.z, n=inp.n, C=inp.C, Rday=inp.Rday, mask=inp.mask,
                        h0=inp.h, hmax0=inp.hmax, tmax0=inp.tmax, dx=np.float64(inp.dx),
                        h=g.h, hmax=g.hmax, tmax=g.tmax, log=log,
                        cfg=json.dumps(cfg_dict(cfg)), acc=json.dumps(accd))
    return path, acc.step


PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "fca_types.h"
#define F(T, f) printf("\"%s.%s\": [%zu, %zu],\n", #T, #f, offsetof(T, f), sizeof(((T *)0)->f));
#define S(T) printf("\"%s\": [0, %zu],\n", #T, sizeof(T));
int main(void) {
  printf("{\n");
  S(FcaPhys) S(FcaRunCfg) S(FcaComm) S(FcaGrid) S(FcaAcc) S(FcaTimers)
  %FIELDS%
  printf("\"end\": [0, 0]\n}\n");
  return 0;
}
"""


def abi_fields():
    a = A.abi("f32")
    out = []
    for sname in ("FcaPhys", "FcaRunCfg", "FcaComm", "FcaGrid", "FcaAcc", "FcaTimers"):
        for fname, _ in getattr(a, sname)._fields_:
            c = "lambda" if fname == "lambda_" else fname
            out.append((sname, c))
    return out


def write_abi_layout(include_dir, defs):
    fields = "\n  ".join(f"F({s}, {f})" for s, f in abi_fields())
    src = PROBE.replace("%FIELDS%", fields)
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "probe.c")
        exe = os.path.join(td, "probe")
        open(c, "w").write(src)
        subprocess.run(["gcc", *defs, f"-I{include_dir}", c, "-o", exe], check=True)
        return json.loads(subprocess.run([exe], capture_output=True, text=True,
                                         check=True).stdout)


def main():
    for prec in ("f32", "f64"):
        for name, (nr, nc, kw) in CASES.items():
            p, steps = make_fixture(name, nr, nc, kw, prec)
            print(f"{os.path.relpath(p, ROOT)}: {steps} steps")
    lay = {"f32": write_abi_layout(os.path.join(REF, "src", "common"), []),
           "f64": write_abi_layout(os.path.join(REF, "src", "common"), ["-DFCA_DOUBLE"])}
    with open(os.path.join(HERE, "abi_layout.json"), "w") as fp:
        json.dump(lay, fp, indent=1, sort_keys=True)
    print("tests/golden/abi_layout.json")


if __name__ == "__main__":
    main()
