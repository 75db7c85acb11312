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


class McArgs(C.Structure):
    _fields_ = [("g", C.c_float), ("d", C.c_float), ("rd", C.c_float), ("e", C.c_float),
                ("f", C.c_float), ("dok", C.c_int), ("k", C.c_int)]


def _lib():
    from floodca_amd.build import load_hip_library
    lib = load_hip_library(os.path.join(ROOT, "tests", "gpu", "libmathcheck.so"))
    lib.mc_eval.restype = C.c_int
    lib.mc_eval.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.POINTER(McArgs), C.c_void_p]
    return lib


def dev(lib, kind, first, count, g=9.81, d=90.0, rd=None, dok=1, e=0.0, f=0.0, k=0):
    out = np.empty(count, dtype=np.float32)
    rd = np.float32(1) / np.float32(d) if rd is None else rd
    a = McArgs(g, d, float(rd), e, f, dok, k)
    rc = lib.mc_eval(kind, first, count, C.byref(a), out.ctypes.data)
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


def _range(lo, hi):
    lo = int(np.float32(lo).view(np.uint32))
    hi = int(np.float32(hi).view(np.uint32))
    for first in range(lo, hi + 1, CHUNK):
        yield first, min(CHUNK, hi + 1 - first)


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_sqrt_in_range_is_correctly_rounded_on_every_float_of_its_range():
    lib = _lib()
    for first, n in _range(2.0 ** -96, 2.0 ** 126):
        got = dev(lib, 7, first, n)
        x = bits(first, n).view(np.float32)
        idx = same(got, np.sqrt(x))
        assert idx.size == 0, f"sqrt mismatch at {hex(first + int(idx[0]))}"


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_rcp_in_range_is_correctly_rounded_on_every_float_of_its_range():
    lib = _lib()
    for first, n in _range(2.0 ** -100, 2.0 ** 100):
        got = dev(lib, 8, first, n)
        x = bits(first, n).view(np.float32)
        idx = same(got, np.float32(1) / x)
        assert idx.size == 0, f"rcp mismatch at {hex(first + int(idx[0]))}"


@pytest.mark.gpu
@pytest.mark.timeout(600)
@pytest.mark.parametrize("dconst", [90.0, float(np.float32(np.sqrt(2.0)) * np.float32(90.0))])
def test_div_c_big_over_its_domain(dconst):
    lib = _lib()
    d = np.float32(dconst)
    for first, n in _range(2.0 ** -80, 2.0 ** 80):
        got = dev(lib, 9, first, n, d=float(d))
        x = bits(first, n).view(np.float32)
        idx = same(got, x / d)
        assert idx.size == 0, f"x/{d} mismatch at {hex(first + int(idx[0]))}"


@pytest.mark.gpu
def test_div_c_keeps_signed_zeros():
    lib = _lib()
    for b in (0x00000000, 0x80000000):
        got = dev(lib, 4, b, 1, d=8100.0)
        assert got.view(np.uint32)[0] == b


@pytest.mark.gpu
@pytest.mark.timeout(600)
@pytest.mark.parametrize("k", [-300, -40, -3, -1, 0, 1, 2, 40])
def test_crit_cap_equals_min_with_the_critical_speed(k):
    """v stepped k ulps from the critical speed, for every depth in
    [1e-3, 1e3]: the filtered cap returns exactly min(v, vcr)."""
    lib = _lib()
    for first, n in _range(1e-3, 1e3):
        got = dev(lib, 10, first, n, k=k)
        hf = bits(first, n).view(np.float32)
        vcr = np.sqrt(np.float64(np.float32(9.81)) * hf.astype(np.float64)).astype(np.float32)
        v = (vcr.view(np.int32) + np.int32(k)).view(np.float32)
        idx = same(got, np.minimum(v, vcr))
        assert idx.size == 0, f"k={k}: mismatch at {hex(first + int(idx[0]))}"


@pytest.mark.gpu
@pytest.mark.timeout(600)
@pytest.mark.parametrize("n,ax", [(0.02, 0.5), (0.15, 3e-4), (0.035, 0.0), (1e-13, 2.0),
                                  (0.06, 40.0)])
@pytest.mark.parametrize("k", [-1, -4, -4000, 1])
def test_stage_d_bound_never_skips_a_faster_cell(n, ax, k):
    """For every depth in [1e-3, 1e3] and thr one or more ulps below (k < 0) or
    above the cell's exact speed, the bound never declares 'cannot exceed' for
    a cell that does; for thr below the speed it must ask for the exact value."""
    lib = _lib()
    for first, cnt in _range(1e-3, 1e3):
        got = dev(lib, 11, first, cnt, e=n, f=ax, k=k)
        assert not np.any(got == 1.0), f"skipped a faster cell (n={n}, ax={ax}, k={k})"
        if k < 0 and ax > 0:
            assert np.all(got == 0.5)


@pytest.mark.gpu
@pytest.mark.timeout(900)
@pytest.mark.parametrize("den", [480.0, 960.0, float(np.float32(8.0) * np.float32(52.371)),
                                 float(np.float32(16.0) * np.float32(0.0137)),
                                 float(np.float32(8.0) * np.float32(1.9999999))])
def test_step_constant_division_is_correctly_rounded(den):
    """div_step_cr (the level cap's |dL|*A / ((2K)*dt)) against IEEE division
    for every dividend in [2^-80, 2^80], for step denominators including one
    whose significand is all ones."""
    lib = _lib()
    d = np.float32(den)
    rd = np.float32(1) / d
    for first, n in _range(2.0 ** -80, 2.0 ** 80):
        got = dev(lib, 12, first, n, d=float(d), rd=float(rd))
        x = bits(first, n).view(np.float32)
        idx = same(got, x / d)
        assert idx.size == 0, f"x/{d} mismatch at {hex(first + int(idx[0]))}"


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_hardware_log2_error_bound_on_every_normal_float():
    """The stage D filter in logarithms (speed_may_exceed_fast) assumes the
    hardware log2 is within 2^-14 of log2 on every positive normal float."""
    lib = _lib()
    print("exhaustive log2", flush=True)
    worst = 0.0
    for first in range(0x00800000, TOP, CHUNK):
        n = min(CHUNK, TOP - first)
        worst = max(worst, float(dev(lib, 13, first, n).max()))
    print(f"max |lg - log2| = {worst:.3e}", flush=True)
    assert worst <= 2.0 ** -14


@pytest.mark.gpu
@pytest.mark.timeout(600)
@pytest.mark.parametrize("n,ax", [(0.02, 0.5), (0.15, 3e-4), (0.035, 0.0), (1e-13, 2.0),
                                  (0.06, 40.0)])
@pytest.mark.parametrize("k", [-1, -4, -4000, 1])
def test_stage_d_log_bound_never_skips_a_faster_cell(n, ax, k):
    """As the sixth-power bound: for every depth in [1e-3, 1e3] and thr one or
    more ulps below (k < 0) or above the cell's exact speed, the log2 bound
    never skips a cell that is faster than thr."""
    lib = _lib()
    for first, cnt in _range(1e-3, 1e3):
        got = dev(lib, 14, first, cnt, e=n, f=ax, k=k)
        assert not np.any(got == 1.0), f"skipped a faster cell (n={n}, ax={ax}, k={k})"
        if k < 0 and ax > 0:
            assert np.all(got == 0.5)


def _div_pairs_lib():
    lib = _lib()
    lib.mc_div_pairs.restype = C.c_int
    lib.mc_div_pairs.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p]
    return lib


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_two_fma_division_is_exact_for_every_significand_pair():
    """Markstein's refinement without a certificate (div_step_fast): over all
    2^23 divisor significands x all 2^23 dividend significands, every
    quotient is correctly rounded (the all-ones divisor significand, the
    hardest case for a reciprocal, included)."""
    lib = _div_pairs_lib()
    print("exhaustive division pairs", flush=True)
    bad = []
    step = 1 << 20
    for first in range(0, 1 << 23, step):
        out = np.empty(step, dtype=np.uint32)
        rc = lib.mc_div_pairs(first, step, out.ctypes.data)
        assert rc == 0, f"mc_div_pairs failed: HIP error {rc}"
        bad += (np.flatnonzero(out) + first).tolist()
        print(f"divisors {first:#x}..{first + step - 1:#x}: {np.count_nonzero(out)} failing",
              flush=True)
    assert bad == []


@pytest.mark.gpu
@pytest.mark.timeout(600)
@pytest.mark.parametrize("den", [480.0, 960.0, float(np.float32(8.0) * np.float32(52.371)),
                                 float(np.float32(16.0) * np.float32(0.0137)),
                                 float(np.float32(8.0) * np.float32(1.9999999))])
def test_step_division_with_the_fast_flag_is_correctly_rounded(den):
    """div_step_fast as the LAND form calls it, against IEEE division for
    every dividend in [2^-80, 2^80]."""
    lib = _lib()
    d = np.float32(den)
    rd = np.float32(1) / d
    for first, n in _range(2.0 ** -80, 2.0 ** 80):
        got = dev(lib, 16, first, n, d=float(d), rd=float(rd))
        x = bits(first, n).view(np.float32)
        idx = same(got, x / d)
        assert idx.size == 0, f"x/{d} mismatch at {hex(first + int(idx[0]))}"


@pytest.mark.gpu
@pytest.mark.timeout(600)
@pytest.mark.parametrize("g", [9.81, 9.80665, 1.62, 24.79])
def test_certified_critical_speed_on_every_depth(g):
    """crit_speed_cert, the LAND form's branch-free critical speed: for every
    h in [2^-15, 2^15] it either defers (rarely) or equals the reference's
    (float)sqrt((double)g*(double)h) bit for bit."""
    lib = _lib()
    gf = np.float32(g)
    deferred = total = 0
    for first, n in _range(2.0 ** -15, 2.0 ** 15):
        got = dev(lib, 17, first, n, g=float(gf))
        h = bits(first, n).view(np.float32)
        ref = np.sqrt(np.float64(gf) * h.astype(np.float64)).astype(np.float32)
        ok = got >= 0
        idx = same(got[ok], ref[ok])
        assert idx.size == 0, f"g={g}: mismatch at h={h[ok][idx[0]]!r}"
        deferred += int(np.count_nonzero(~ok))
        total += n
    assert deferred < total * 2.0 ** -10


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_certified_cube_root_on_every_depth():
    """cbrt_sq_cert, the LAND form's branch-free hf^(2/3): for every hf in
    [2^-15, 2^15] it either defers (rarely) or equals glibc's
    (float)cbrt((double)hf*hf) bit for bit."""
    lib, o = _lib(), Oracle("f32")
    deferred = total = 0
    for first, n in _range(2.0 ** -15, 2.0 ** 15):
        got = dev(lib, 18, first, n)
        ok = got >= 0
        ref = o.libm_range(0, first, n)
        idx = same(got[ok], ref[ok])
        assert idx.size == 0, f"mismatch at bits {hex(first + int(np.flatnonzero(ok)[idx[0]]))}"
        deferred += int(np.count_nonzero(~ok))
        total += n
    assert deferred < total * 2.0 ** -10


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_wave_max_is_the_wavefront_maximum(dtype):
    """The stage-D threshold reduction: every lane of each 64-lane wavefront
    gets the largest of the wavefront's non-negative values, wherever that
    maximum sits (one wavefront per lane position), among zeros, subnormals,
    ties and random values."""
    lib = _lib()
    lib.mc_wave_max.restype = C.c_int
    lib.mc_wave_max.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_uint32]
    rng = np.random.default_rng(7)
    waves = []
    for pos in range(64):  # the maximum at each lane in turn
        w = rng.random(64).astype(dtype)
        w[pos] = dtype(2.0)
        waves.append(w)
    waves.append(np.zeros(64, dtype))
    w = np.zeros(64, dtype)
    w[63] = np.finfo(dtype).smallest_subnormal
    waves.append(w)
    w = np.zeros(64, dtype)
    w[0] = np.finfo(dtype).smallest_subnormal
    waves.append(w)
    waves.append(np.full(64, 3.5, dtype))
    while len(waves) % 4:
        waves.append((rng.random(64) * 10.0 ** rng.integers(-30, 30, 64)).astype(dtype))
    x = np.ascontiguousarray(np.stack(waves))
    out = np.empty_like(x)
    rc = lib.mc_wave_max(int(dtype == np.float64), x.ctypes.data, out.ctypes.data, x.size)
    assert rc == 0, f"mc_wave_max failed: HIP error {rc}"
    want = np.repeat(x.max(axis=1, keepdims=True), 64, axis=1)
    assert np.array_equal(out.view(np.uint8), want.view(np.uint8))
