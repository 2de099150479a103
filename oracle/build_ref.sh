#!/bin/sh
# Builds oracle/_ref/ref_harness from the reference's unmodified sources
# (src/constants.f90, src/utils.f90, src/kdtree.f90 under /root/reference) and
# the harness oracle/ref_harness.f90, with the flang of the ROCm toolchain.
# Outputs go to oracle/_ref/ only.  Exits 0 without building when the
# reference tree or flang is absent (the GPU host has neither need nor tree).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
REF=${TPCF_REFERENCE:-/root/reference}
FC=${FC:-/opt/rocm/llvm/bin/flang}
OUT="$HERE/_ref"
if [ ! -f "$REF/src/utils.f90" ] || [ ! -x "$FC" ]; then
    echo "build_ref: reference sources or flang not found; skipped"
    exit 0
fi
mkdir -p "$OUT"
cd "$OUT"
"$FC" -O2 -c "$REF/src/constants.f90" -o constants.o
"$FC" -O2 -c "$REF/src/utils.f90" -o utils.o
"$FC" -O2 -c "$REF/src/kdtree.f90" -o kdtree.o
"$FC" -O2 -c "$HERE/ref_harness.f90" -o ref_harness.o
"$FC" -O2 -o ref_harness ref_harness.o kdtree.o utils.o constants.o
echo "build_ref: $OUT/ref_harness"
