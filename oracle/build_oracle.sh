#!/usr/bin/env bash
# Builds the oracle (test infrastructure, never shipped):
#   oracle/liboracle_f32.so, oracle/liboracle_f64.so -- the CPU restatement
#       (fca_oracle.c), built from this repository only.
#   oracle/_ref/... -- the reference's own C sources compiled where they lie,
#       only when the reference tree is present (REF, default /root/reference):
#       libfca_ref_{f32,f64}.so  parity builds (contraction off) + ref_shim.c
#       floodca_serial_{f32,f64} floodca_omp_{f32,f64}  timing builds with the
#           reference's stock Release flags (-O3, contraction left at GCC's
#           default), targeting x86-64-v3 so they run on any AVX2 host.
# Usage: oracle/build_oracle.sh [--no-ref]
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$(cd "$HERE/.." && pwd)"
REF="${REF:-/root/reference}"
CC="${CC:-gcc}"
PARITY=(-std=gnu11 -O2 -ffp-contract=off -fno-fast-math -fPIC -D_ISOC11_SOURCE)

for prec in f32 f64; do
    def=()
    [[ $prec == f64 ]] && def=(-DFCA_DOUBLE)
    "$CC" "${PARITY[@]}" "${def[@]}" -Wall -fopenmp -shared -I"$ROOT/include" \
        "$HERE/fca_oracle.c" -lm -o "$HERE/liboracle_${prec}.so"
done

if [[ "${1:-}" == "--no-ref" || ! -d "$REF/src/common" ]]; then
    exit 0
fi
mkdir -p "$HERE/_ref"
SRC=("$REF"/src/common/config.c "$REF"/src/common/io.c "$REF"/src/common/grid.c
     "$REF"/src/common/rain.c "$REF"/src/common/diag.c "$REF"/src/common/timer.c
     "$REF"/src/common/log.c "$REF"/src/common/model.c "$REF"/src/common/cli.c)
INC=(-I"$REF/src/common" -I"$REF/src/kernels")
STOCK=(-std=gnu11 -O3 -march=x86-64-v3 -D_ISOC11_SOURCE)
for prec in f32 f64; do
    def=()
    [[ $prec == f64 ]] && def=(-DFCA_DOUBLE)
    "$CC" "${PARITY[@]}" "${def[@]}" -w -shared "${INC[@]}" "${SRC[@]}" \
        "$HERE/ref_shim.c" -lm -o "$HERE/_ref/libfca_ref_${prec}.so"
    "$CC" "${STOCK[@]}" "${def[@]}" -w "${INC[@]}" "${SRC[@]}" \
        "$REF/src/serial/main_serial.c" -lm -o "$HERE/_ref/floodca_serial_${prec}"
    "$CC" "${STOCK[@]}" "${def[@]}" -w -fopenmp -DFCA_OMP "${INC[@]}" "${SRC[@]}" \
        "$REF/src/omp/main_omp.c" -lm -o "$HERE/_ref/floodca_omp_${prec}"
done
