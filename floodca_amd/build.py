"""Builds the in-tree native libraries.

* floodca_amd/libfloodca_hip.so, libfloodca_hip_f64.so -- the product: the HIP
  kernels (hipcc, gfx950) and the C host (gcc), one library per precision.
* floodca_amd/libfca_nn.so -- the surrogate network's kernels (hipcc, gfx950).
* tests/emu/libwalk_emu_{f32,f64}.so -- host emulation of the fused walk for the
  CPU test suite (g++), never loaded by the product.
* oracle/ -- the checker (oracle/build_oracle.sh), and, where the reference
  tree is present, the reference's driver around this library
  (oracle/build_hip_main.sh -> oracle/_ref/floodca_hip_{f32,f64}).

Run ``python -m floodca_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "floodca_amd")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
INC = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("FCA_OFFLOAD_ARCH", "gfx950")

# Contraction off everywhere: the step must round like the reference build.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", f"-I{INC}", f"-I{CSRC}"]
C_FLAGS = ["-O2", "-std=gnu11", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
           "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{INC}", f"-I{CSRC}"]
EMU_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
             "-Wall", "-Wno-unused-variable", f"-I{INC}", f"-I{CSRC}"]

PRECS = {"f32": ([], "libfloodca_hip.so"), "f64": (["-DFCA_DOUBLE"], "libfloodca_hip_f64.so")}


def load_hip_library(path: str):
    """ctypes handle of one of this package's HIP libraries, with one HIP
    runtime per process: PyTorch's ROCm build carries its own copy of
    libamdhip64 under a file name its libraries ask for by path, so a process
    that loads ours (linked against /opt/rocm's) first and imports torch
    later holds two runtimes, and torch's may then find no device. Importing
    torch first, where it is installed, makes our libraries bind to the copy
    torch loaded (the same soname)."""
    import ctypes

    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    return ctypes.CDLL(path)


def lib_path(prec: str = "f32") -> str:
    """The in-tree backend library; FCA_HIP_LIB (float) / FCA_HIP_LIB_F64 override it."""
    env = os.environ.get("FCA_HIP_LIB" if prec == "f32" else "FCA_HIP_LIB_F64")
    return env or os.path.join(PKG, PRECS[prec][1])


def _run(cmd, quiet=False):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(ln for ln in r.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in ln)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip() and not quiet:
        print(out)
    return out


def build_hip(prec: str, extra=(), out: str | None = None, tag: str = "") -> str:
    os.makedirs(BUILD, exist_ok=True)
    defs, name = PRECS[prec]
    kobj = os.path.join(BUILD, f"fca_kernels_{prec}{tag}.o")
    cobj = os.path.join(BUILD, f"hip_backend_{prec}{tag}.o")
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    _run([hipcc, *HIP_FLAGS, *defs, *extra, "-c", os.path.join(CSRC, "fca_kernels.hip"),
          "-o", kobj])
    _run(["gcc", *C_FLAGS, *defs, "-c", os.path.join(CSRC, "hip_backend.c"), "-o", cobj])
    out = out or os.path.join(PKG, name)
    tmp = out + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", kobj, cobj, "-o", tmp])
    os.replace(tmp, out)
    return out


def nn_lib_path() -> str:
    """The in-tree surrogate kernel library; FCA_NN_LIB overrides it."""
    return os.environ.get("FCA_NN_LIB") or os.path.join(PKG, "libfca_nn.so")


def build_nn(_=None) -> str:
    out = nn_lib_path()
    tmp = out + ".tmp"
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    _run([hipcc, *HIP_FLAGS, "-shared", os.path.join(CSRC, "fca_nn.hip"), "-o", tmp])
    os.replace(tmp, out)
    return out


# FCA_EMU_SANITIZE=1 builds the emulation with AddressSanitizer and UBSan
# (host code; the walk's indexing checked on the CPU). Run the CPU suite on it
# with both runtimes preloaded, e.g.
#   LD_PRELOAD="$(gcc -print-file-name=libasan.so) $(gcc -print-file-name=libubsan.so)"
#   ASAN_OPTIONS=detect_leaks=0 python -m pytest tests/test_emu.py
# and build again without the variable afterwards.
SAN_FLAGS = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
             "-fno-omit-frame-pointer"]


def build_emu(prec: str) -> str:
    defs, _ = PRECS[prec]
    out = os.path.join(ROOT, "tests", "emu", f"libwalk_emu_{prec}.so")
    san = SAN_FLAGS if os.environ.get("FCA_EMU_SANITIZE") == "1" else []
    _run(["g++", *EMU_FLAGS, *san, *defs, os.path.join(ROOT, "tests", "emu", "walk_emu.cpp"),
          "-o", out, "-lm"])
    return out


def build_mathcheck() -> str:
    """Test-only HIP library evaluating the device scalar math (tests/gpu)."""
    out = os.path.join(ROOT, "tests", "gpu", "libmathcheck.so")
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    tmp = out + ".tmp"
    _run([hipcc, *HIP_FLAGS, "-shared", os.path.join(ROOT, "tests", "gpu", "mathcheck.hip"),
          "-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle() -> None:
    _run(["bash", os.path.join(ROOT, "oracle", "build_oracle.sh")])


def build_hip_main() -> None:
    """oracle/_ref/floodca_hip_{f32,f64}: the reference's driver around this
    library (oracle/build_hip_main.sh; needs the reference tree and the built
    libraries, else nothing is built)."""
    _run(["bash", os.path.join(ROOT, "oracle", "build_hip_main.sh")])


def build_all(verbose: bool = True) -> list[str]:
    if shutil.which(os.path.join(ROCM, "bin", "hipcc")) is None:
        raise RuntimeError(f"hipcc not found under {ROCM}")
    jobs = [(build_hip, "f32"), (build_hip, "f64"), (build_emu, "f32"), (build_emu, "f64"),
            (build_nn, None)]
    with ThreadPoolExecutor(max_workers=5) as ex:
        futs = [ex.submit(f, p) for f, p in jobs] + [ex.submit(build_oracle),
                                                     ex.submit(build_mathcheck)]
        outs = [f.result() for f in futs]
    build_hip_main()  # links against the libraries just built
    if verbose:
        for o in outs:
            if o:
                print("built", os.path.relpath(o, ROOT))
    return [o for o in outs if o]


if __name__ == "__main__":
    build_all()
    sys.exit(0)
