"""Device-to-device copy of a known size, to calibrate HBM traffic counters.

``python -m floodca_amd.copyprobe --mib 1024 --reps 3`` copies a buffer with
hipMemcpy (the runtime's blit kernel) so that rocprofv3 --pmc FETCH_SIZE /
WRITE_SIZE of that dispatch can be compared with the bytes it must move: each
copy reads and writes exactly ``mib`` MiB.
"""
from __future__ import annotations

import argparse
import ctypes as C
import sys


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    try:
        rt = C.CDLL("libamdhip64.so")
    except OSError:
        rt = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    n = args.mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    if rt.hipMalloc(C.byref(a), C.c_size_t(n)) or rt.hipMalloc(C.byref(b), C.c_size_t(n)):
        print("hipMalloc failed", file=sys.stderr)
        return 1
    rt.hipMemset(a, 1, C.c_size_t(n))
    rt.hipDeviceSynchronize()
    for _ in range(args.reps):
        rt.hipMemcpy(b, a, C.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    rt.hipDeviceSynchronize()
    rt.hipFree(a)
    rt.hipFree(b)
    print(f"copied {args.mib} MiB x {args.reps}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
