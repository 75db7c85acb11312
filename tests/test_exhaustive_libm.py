"""The reference's float-mode libm detours are correctly rounded (CPU).

physics.h:19-24 computes hf^(2/3) as (float)cbrt((double)hf*(double)hf) with
glibc's double cbrt, which is not itself correctly rounded. This test checks
every positive finite float: the glibc expression always equals the correctly
rounded hf^(2/3) (decided by exact midpoint cube comparisons), so any correctly
rounded implementation -- the device's -- reproduces the reference bit for bit.
The sqrt detour (float)sqrt((double)x) is correctly rounded by the double
rounding theorem (53 >= 2*24 + 2); it is checked here on a sample and
exhaustively against the device in tests/test_gpu_math.py.
"""
import numpy as np

from oracle import Oracle

TOP = 0x7F800000  # +inf


def test_cbrt_sq_is_correctly_rounded_for_every_positive_float():
    o = Oracle("f32")
    bad = 0
    step = 1 << 28
    for first in range(1, TOP, step):
        bad += o.lib.or_cbrt_sq_cr_mismatches(first, min(step, TOP - first))
    assert bad == 0


def test_sqrt_detour_equals_float_sqrt_sample():
    o = Oracle("f32")
    first = 0x30000000
    ref = o.libm_range(1, first, 1 << 22)
    x = np.arange(first, first + (1 << 22), dtype=np.uint32).view(np.float32)
    assert np.array_equal(ref, np.sqrt(x))
