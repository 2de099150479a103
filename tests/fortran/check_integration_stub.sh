#!/bin/sh
# Compile check of the reference-side binding ("Option 2" of INTEGRATION.md):
# builds the stub text between the option2-stub markers of INTEGRATION.md
# into the reference's own module tpcf and links a program that calls
# shear_2pcf and kappa_2pcf against 2pcf_amd/lib2pcf_amd.so.  The module is
# assembled from src/tpcf.f90 unmodified: its use lines, its derived types
# (src/tpcf.f90:17-43) and allocate_shear_2pcf_results / shear_2pcf /
# kappa_2pcf (src/tpcf.f90:99-188), with the stub's procedures in place of
# the two workers; it is compiled with the reference's constants, utils,
# kdtree and kmeans modules and this project's tpcf_gpu_c.  Nothing is run.
#
# Also builds tests/fortran/bins_forms.F90 twice, against the reference's
# utils and against this project's tpcf_gpu, into $1 and $2 when given (the
# CPU test runs both and compares their output).
#
# Reads the reference tree where it lies; every intermediate stays in a
# temporary directory.  Exits 0 with "skipped" when the reference tree,
# flang or the library is absent.
set -e
ROOT=$(cd "$(dirname "$0")/../.." && pwd)
REF=${TPCF_REFERENCE:-/root/reference}
FC=${FC:-/opt/rocm/llvm/bin/flang}
LIB="$ROOT/2pcf_amd/lib2pcf_amd.so"
if [ ! -f "$REF/src/tpcf.f90" ] || [ ! -x "$FC" ] || [ ! -f "$LIB" ]; then
    echo "check_integration_stub: reference tree, flang or library not found; skipped"
    exit 0
fi
WORK=$(mktemp -d)
trap 'rm -rf "$WORK"' EXIT
cd "$WORK"
FFLAGS="-O1 -fopenmp"
for m in constants utils kdtree kmeans; do
    "$FC" $FFLAGS -c "$REF/src/$m.f90" -o $m.o
done
"$FC" $FFLAGS -c "$ROOT/2pcf_amd/fortran/tpcf_gpu.f90" -o tpcf_gpu.o

{
    echo "module tpcf"
    sed -n '2,/^implicit none/p' "$REF/src/tpcf.f90" | grep '^ *use '
    echo "use tpcf_gpu_c"
    echo "implicit none"
    echo "private"
    echo "public shear_2pcf, kappa_2pcf, allocate_shear_2pcf_results"
    echo "public shear_2pcf_results, shape_catalog, kappa_map, lens_catalog, lensing_survey_field"
    sed -n '/^type shear_2pcf_results/,/^end type lensing_survey_field/p' "$REF/src/tpcf.f90"
    echo "contains"
    sed -n '/^ *subroutine allocate_shear_2pcf_results(/,/^ *end subroutine allocate_shear_2pcf_results/p' \
        "$REF/src/tpcf.f90"
    sed -n '/^ *subroutine shear_2pcf(/,/^ *end subroutine kappa_2pcf/p' "$REF/src/tpcf.f90"
    sed -n '/<!-- option2-stub:begin -->/,/<!-- option2-stub:end -->/p' "$ROOT/INTEGRATION.md" | grep -v '^<!--' |
        grep -v '^```'
    echo "end module tpcf"
} > tpcf_stub.f90
grep -q 'subroutine tangential_shear_tpcf' tpcf_stub.f90
grep -q 'subroutine scalar_kappa_tpcf' tpcf_stub.f90
grep -q 'call tangential_shear_tpcf' tpcf_stub.f90
"$FC" $FFLAGS -c tpcf_stub.f90 -o tpcf_stub.o

cat > driver.f90 <<'EOF'
program stub_driver
    use utils
    use tpcf
    implicit none
    type(lensing_survey_field) :: field
    type(binning) :: bins
    type(shear_2pcf_results) :: res
    integer, parameter :: n = 3
    call create_bins(x_min=1.0, x_max=10.0, n_bin=4, spacing=log_binning, bins=bins)
    field%lenses%n = n
    allocate(field%lenses%x(n), field%lenses%y(n), field%lenses%z(n), field%lenses%val(n), field%lenses%w(n))
    field%shapes%n = n
    allocate(field%shapes%x(n), field%shapes%y(n), field%shapes%z(n), field%shapes%e1(n), field%shapes%e2(n), &
             field%shapes%m(n), field%shapes%w(n))
    field%convergence%n = n
    allocate(field%convergence%x(n), field%convergence%y(n), field%convergence%kappa_E(n), &
             field%convergence%kappa_B(n), field%convergence%w(n))
    call shear_2pcf(field, bins, res, .false., 1, .true., .false.)
    call kappa_2pcf(field, bins, res, .false., 1, .true., .false.)
    print *, res%n_pairs
end program stub_driver
EOF
"$FC" $FFLAGS -c driver.f90 -o driver.o
"$FC" $FFLAGS -o stub_driver driver.o tpcf_stub.o tpcf_gpu.o kmeans.o kdtree.o utils.o constants.o \
    -L"$ROOT/2pcf_amd" -l2pcf_amd -Wl,-rpath,"$ROOT/2pcf_amd"
echo "check_integration_stub: module tpcf with the stub compiled and linked against lib2pcf_amd.so"

if [ -n "$2" ]; then
    "$FC" $FFLAGS -cpp -DUSE_REFERENCE -c "$ROOT/tests/fortran/bins_forms.F90" -o bins_ref.o
    "$FC" $FFLAGS -o "$1" bins_ref.o utils.o constants.o
    "$FC" $FFLAGS -cpp -c "$ROOT/tests/fortran/bins_forms.F90" -o bins_gpu.o
    "$FC" $FFLAGS -o "$2" bins_gpu.o tpcf_gpu.o -L"$ROOT/2pcf_amd" -l2pcf_amd -Wl,-rpath,"$ROOT/2pcf_amd"
    echo "check_integration_stub: bins_forms built against utils and tpcf_gpu"
fi
