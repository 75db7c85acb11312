#!/usr/bin/env bash
# Builds floodca_hip_{f32,f64} and floodca_ens_hip_{f32,f64} into oracle/_ref/:
# the reference's own command-line driver code (src/common, compiled where it
# lies) around floodca_amd/csrc/main_hip.c and main_ens_hip.c, linked against
# this repository's libfloodca_hip*.so -- the executables a maintainer would
# add beside the reference's floodca_cuda and floodca_ens_mpi_cuda. Used by
# the tests to run the library behind the reference's CLI, grid loading, rain
# files and CSV report. Needs the reference tree
# (REF, default /root/reference) and the built libraries; otherwise does
# nothing.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$(cd "$HERE/.." && pwd)"
REF="${REF:-/root/reference}"
CC="${CC:-gcc}"
[[ -d "$REF/src/common" ]] || exit 0
mkdir -p "$HERE/_ref"
SRC=("$REF"/src/common/config.c "$REF"/src/common/io.c "$REF"/src/common/grid.c
     "$REF"/src/common/rain.c "$REF"/src/common/diag.c "$REF"/src/common/timer.c
     "$REF"/src/common/log.c "$REF"/src/common/model.c "$REF"/src/common/cli.c)
INC=(-I"$REF/src/common" -I"$REF/src/kernels")
FLAGS=(-std=gnu11 -O2 -ffp-contract=off -fno-fast-math -D_ISOC11_SOURCE)
for prec in f32 f64; do
    def=()
    lib=libfloodca_hip.so
    if [[ $prec == f64 ]]; then
        def=(-DFCA_DOUBLE)
        lib=libfloodca_hip_f64.so
    fi
    [[ -f "$ROOT/floodca_amd/$lib" ]] || continue
    "$CC" "${FLAGS[@]}" "${def[@]}" -w "${INC[@]}" "$ROOT/floodca_amd/csrc/main_hip.c" \
        "${SRC[@]}" -L"$ROOT/floodca_amd" -l:"$lib" -Wl,-rpath,'$ORIGIN/../../floodca_amd' -lm \
        -o "$HERE/_ref/floodca_hip_${prec}"
    "$CC" "${FLAGS[@]}" "${def[@]}" -w "${INC[@]}" -I"$ROOT/include" \
        "$ROOT/floodca_amd/csrc/main_ens_hip.c" "${SRC[@]}" -L"$ROOT/floodca_amd" -l:"$lib" \
        -Wl,-rpath,'$ORIGIN/../../floodca_amd' -lm -o "$HERE/_ref/floodca_ens_hip_${prec}"
done
