"""The C-ABI boundary (CPU): struct layouts and exported symbols.

include/floodca_hip.h restates src/common/fca_types.h so the library links
beside the reference's backends. Three views must agree field by field, in
both precisions: a probe compiled from our header, the reference header's
layout recorded in tests/golden/abi_layout.json, and the ctypes mirrors.
"""
import ctypes as C
import json
import os
import re
import subprocess

import pytest

from floodca_amd import abi as A
from floodca_amd.build import lib_path

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys_path_golden = os.path.join(ROOT, "tests", "golden")

import sys  # noqa: E402

sys.path.insert(0, sys_path_golden)
from make_golden import PROBE, abi_fields  # noqa: E402


def probe_layout(include_dir, defs, tmp_path):
    fields = "\n  ".join(f"F({s}, {f})" for s, f in abi_fields())
    src = PROBE.replace('#include "fca_types.h"', '#include "floodca_hip.h"')
    src = src.replace("%FIELDS%", fields)
    c, exe = tmp_path / "p.c", tmp_path / "p"
    c.write_text(src)
    subprocess.run(["gcc", *defs, f"-I{include_dir}", str(c), "-o", str(exe)], check=True)
    return json.loads(subprocess.run([str(exe)], capture_output=True, text=True,
                                     check=True).stdout)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_header_matches_reference_layout(prec, tmp_path):
    ours = probe_layout(os.path.join(ROOT, "include"), ["-DFCA_DOUBLE"] if prec == "f64" else [],
                        tmp_path)
    with open(os.path.join(sys_path_golden, "abi_layout.json")) as fp:
        ref = json.load(fp)[prec]
    assert ours == ref


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_ctypes_mirrors_match_reference_layout(prec):
    a = A.abi(prec)
    with open(os.path.join(sys_path_golden, "abi_layout.json")) as fp:
        ref = json.load(fp)[prec]
    for sname in ("FcaPhys", "FcaRunCfg", "FcaComm", "FcaGrid", "FcaAcc", "FcaTimers"):
        cls = getattr(a, sname)
        assert C.sizeof(cls) == ref[sname][1], sname
        for fname, _ in cls._fields_:
            key = f"{sname}.{'lambda' if fname == 'lambda_' else fname}"
            fld = getattr(cls, fname)
            assert [fld.offset, fld.size] == ref[key], key


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "floodca_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fca_hip_\w+)\s*\(", txt)))


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_library_loads_and_exports_every_declared_symbol(prec):
    path = lib_path(prec)
    assert os.path.exists(path), "run python -m floodca_amd.build"
    syms = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                          check=True).stdout.split()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if n not in syms]
    assert not missing, missing
    lib = C.CDLL(path)  # loads without a GPU
    for n in names:
        getattr(lib, n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_crop_table_mirror_matches_the_header(prec, tmp_path):
    """FcaHipCrop (no counterpart in the reference): the ctypes mirror against
    a probe compiled from include/floodca_hip.h."""
    src = tmp_path / "crop.c"
    exe = tmp_path / "crop"
    src.write_text("""#include <stddef.h>
#include <stdio.h>
#include "floodca_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %d\\n", sizeof(FcaHipCrop), offsetof(FcaHipCrop, r0),
         offsetof(FcaHipCrop, c0), offsetof(FcaHipCrop, nrows), offsetof(FcaHipCrop, ncols),
         offsetof(FcaHipCrop, rday), offsetof(FcaHipCrop, n_class), FCA_HIP_MAX_CLASSES);
  return 0;
}
""")
    defs = ["-DFCA_DOUBLE"] if prec == "f64" else []
    subprocess.run(["gcc", *defs, f"-I{os.path.join(ROOT, 'include')}", str(src), "-o", str(exe)],
                   check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    a = A.abi(prec)
    T = a.FcaHipCrop
    assert got == [C.sizeof(T), T.r0.offset, T.c0.offset, T.nrows.offset, T.ncols.offset,
                   T.rday.offset, T.n_class.offset, A.FCA_HIP_MAX_CLASSES]
