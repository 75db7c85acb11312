"""Device scalar math vs libm over every input (GPU).

The fused step computes the reference's float-mode detours with hardware
approximations plus exactness certificates (fca_device.h). Each device function
is a deterministic function of one float, so evaluating it on the GPU for every
float bit pattern in its domain and comparing with the reference expression
(glibc via the oracle, or numpy where numpy is IEEE-exact) proves it equal.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOP = 0x7F800000  # +inf bit pattern
CHUNK = 1 << 26


def _lib():
    lib = C.CDLL(os.path.join(ROOT, "tests", "gpu", "libmathcheck.so"))
    lib.mc_eval.restype = C.c_int
    lib.mc_eval.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_float,
                            C.c_int, C.c_void_p]
    return lib


def dev(lib, kind, first, count, g=9.81, d=90.0, rd=None, dok=1):
    out = np.empty(count, dtype=np.float32)
    rd = np.float32(1) / np.float32(d) if rd is None else rd
    rc = lib.mc_eval(kind, first, count, g, d, float(rd), dok, out.ctypes.data)
    assert rc == 0, f"mc_eval failed: HIP error {rc}"
    return out


def bits(first, count):
    return np.arange(first, first + count, dtype=np.uint64).astype(np.uint32)


def same(a, b):
    """Bitwise equality, NaN == NaN."""
    return np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.timeout(600)
@pytest.mark.parametrize("kind,name", [(0, "cbrt_sq"), (1, "sqrt"), (2, "crit_speed")])
def test_device_libm_detours_match_glibc_on_every_positive_float(kind, name):
    lib, o = _lib(), Oracle("f32")
    print(f"exhaustive {name}", flush=True)
    bad = []
    for first in range(1, TOP, CHUNK):
        n = min(CHUNK, TOP - first)
        d = dev(lib, kind, first, n)
        r = o.libm_range(kind, first, n, 9.81)
        idx = same(d, r)
        if idx.size:
            bad.extend((hex(first + int(i)), float(d[i]), float(r[i])) for i in idx[:5])
    assert not bad, f"{name}: {len(bad)}+ mismatches, e.g. {bad[:5]}"


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_cbrt_newton_error_is_inside_the_certificate_guard():
    """The certificate accepts a Newton value 2^-38 away from a midpoint; the
    value itself must be closer than that to hf^(2/3) (checked against the
    device's double cbrt, itself within 2^-52)."""
    lib = _lib()
    lo = np.float32(2.0 ** -15).view(np.uint32)
    hi = np.float32(2.0 ** 15).view(np.uint32)
    worst = 0.0
    for first in range(int(lo), int(hi) + 1, CHUNK):
        n = min(CHUNK, int(hi) + 1 - first)
        worst = max(worst, float(dev(lib, 6, first, n).max()))
    print(f"cbrt Newton max relative error 2^{np.log2(worst):.2f}")
    assert worst < 2.0 ** -39


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_rcp_cr_is_correctly_rounded_over_its_domain():
    lib = _lib()
    lo = np.float32(2.0 ** -100).view(np.uint32)
    hi = np.float32(2.0 ** 100).view(np.uint32)
    for first in range(int(lo), int(hi) + 1, CHUNK):
        n = min(CHUNK, int(hi) + 1 - first)
        d = dev(lib, 3, first, n)
        x = bits(first, n).view(np.float32)
        idx = same(d, np.float32(1) / x)
        assert idx.size == 0, f"rcp mismatch at {hex(first + int(idx[0]))}"


@pytest.mark.gpu
@pytest.mark.timeout(600)
@pytest.mark.parametrize("dconst", [90.0, 8100.0, float(np.float32(np.sqrt(2.0)) * np.float32(90.0))])
def test_constant_division_over_its_domain(dconst):
    """div_c with the host-verified flag set, for the divisors of the 90 m grid."""
    lib = _lib()
    d = np.float32(dconst)
    lo = np.float32(2.0 ** -80).view(np.uint32)
    hi = np.float32(2.0 ** 80).view(np.uint32)
    for first in range(int(lo), int(hi) + 1, CHUNK):
        n = min(CHUNK, int(hi) + 1 - first)
        got = dev(lib, 4, first, n, d=float(d))
        x = bits(first, n).view(np.float32)
        idx = same(got, x / d)
        assert idx.size == 0, f"x/{d} mismatch at {hex(first + int(idx[0]))}"


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_rough_sqrt_is_within_the_filter_margin():
    """The cap filters assume the hardware sqrt is within 2^-21 relative."""
    lib = _lib()
    lo = np.float32(2.0 ** -100).view(np.uint32)
    hi = np.float32(2.0 ** 100).view(np.uint32)
    worst = 0.0
    for first in range(int(lo), int(hi) + 1, CHUNK):
        n = min(CHUNK, int(hi) + 1 - first)
        got = dev(lib, 5, first, n).astype(np.float64)
        x = bits(first, n).view(np.float32).astype(np.float64)
        rel = np.abs(got / np.sqrt(x) - 1.0)
        worst = max(worst, float(rel.max()))
    print(f"rough sqrt max relative error {worst:.3e}")
    assert worst < 2.0 ** -21
