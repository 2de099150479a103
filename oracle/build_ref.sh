#!/bin/sh
# Builds the reference's own code into oracle/_ref/ with the flang of the ROCm
# toolchain, from the sources where they lie under the reference tree:
#
#   ref_harness      src/constants.f90, src/utils.f90, src/kdtree.f90 (unmodified)
#                    + oracle/ref_harness.f90 (binning and kd-tree vectors)
#   libref2pcf.so    the same modules + src/kmeans.f90 (unmodified) + the
#                    correlation subroutines tangential_shear_tpcf and
#                    scalar_kappa_tpcf of src/tpcf.f90 + oracle/ref_binding.f90
#                    (bind(C) entry points), OpenMP on
#
# src/tpcf.f90 as a whole does not parse with flang (line 1160: the I/O list
# "((g_t(j), g_x(j), block_weight(i), 0), j=1,bins%n_bin)" in the bootstrap's
# per-block file output is not standard Fortran).  The two correlation
# subroutines are therefore taken verbatim, by their subroutine/end subroutine
# lines, into a temporary file inside a module that uses the same modules as
# the reference's module tpcf (src/tpcf.f90:1-8).  Nothing in them is changed.
#
# Outputs go to oracle/_ref/ only (git-ignored): the binary and the shared
# library, no source and no intermediate.  Exits 0 without building when
# the reference tree or flang is absent.
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
# intermediates (objects, module files and the extract of src/tpcf.f90) live in
# a temporary directory outside the repository; only the two products are kept
WORK=$(mktemp -d)
trap 'rm -rf "$WORK"' EXIT
cd "$WORK"
FFLAGS="-O2 -fPIC -fopenmp"
"$FC" $FFLAGS -c "$REF/src/constants.f90" -o constants.o
"$FC" $FFLAGS -c "$REF/src/utils.f90" -o utils.o
"$FC" $FFLAGS -c "$REF/src/kdtree.f90" -o kdtree.o
"$FC" $FFLAGS -c "$REF/src/kmeans.f90" -o kmeans.o
"$FC" $FFLAGS -c "$HERE/ref_harness.f90" -o ref_harness.o
"$FC" $FFLAGS -o "$OUT/ref_harness" ref_harness.o kdtree.o utils.o constants.o

{
    echo "module ref_tpcf_pairs"
    echo "use, intrinsic :: ISO_FORTRAN_ENV"
    echo "use kdtree_module"
    echo "use utils"
    echo "implicit none"
    echo "contains"
    sed -n '/^ *subroutine tangential_shear_tpcf(/,/^ *end subroutine scalar_kappa_tpcf/p' "$REF/src/tpcf.f90"
    echo "end module ref_tpcf_pairs"
} > ref_tpcf_pairs.f90
"$FC" $FFLAGS -c ref_tpcf_pairs.f90 -o ref_tpcf_pairs.o
"$FC" $FFLAGS -c "$HERE/ref_binding.f90" -o ref_binding.o
"$FC" $FFLAGS -shared -o "$OUT/libref2pcf.so" ref_binding.o ref_tpcf_pairs.o kmeans.o kdtree.o utils.o constants.o
echo "build_ref: $OUT/ref_harness $OUT/libref2pcf.so"
