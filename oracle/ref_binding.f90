! ref_binding.f90 -- C-callable entry points around the reference's own
! correlation subroutines tangential_shear_tpcf (src/tpcf.f90:190-378) and
! scalar_kappa_tpcf (src/tpcf.f90:380-545), plus create_bins /
! find_bin_index (src/utils.f90:101-230).  oracle/build_ref.sh compiles them,
! unmodified, together with this file into oracle/_ref/libref2pcf.so.
!
! TEST INFRASTRUCTURE ONLY: used by tests/golden/make_ref_vectors.py to write
! golden vectors and, when it has been built, by bench.py as the timed CPU
! baseline ("reference" kind).  Nothing in the product links or loads it.
!
! Arrays are FP32 (the reference's default real), marks are (n_bin, n_A) in
! Fortran order, n_pairs is int64.  use_tree /= 0 selects the kd-tree search,
! n_threads is passed to the OpenMP region as num_threads.
module ref_binding
    use, intrinsic :: iso_c_binding
    use utils
    use ref_tpcf_pairs
    implicit none
contains

    subroutine ref_shear_c(n_a, xa, ya, za, va, wa, n_b, xb, yb, zb, e1, e2, wb, mb, &
                           x_min, x_max, n_bin, spacing, p, n_threads, use_tree, &
                           theta, gt, gx, w, npairs, xie, xib, km, wm, mw) bind(C, name="ref_shear_c")
        integer(c_int), value :: n_a, n_b, n_bin, spacing, n_threads, use_tree
        real(c_float), value :: x_min, x_max, p
        real(c_float), intent(in) :: xa(n_a), ya(n_a), za(n_a), va(n_a), wa(n_a)
        real(c_float), intent(in) :: xb(n_b), yb(n_b), zb(n_b), e1(n_b), e2(n_b), wb(n_b), mb(n_b)
        real(c_float), intent(out) :: theta(n_bin), gt(n_bin), gx(n_bin), w(n_bin)
        integer(c_int64_t), intent(out) :: npairs(n_bin)
        real(c_float), intent(out) :: xie(n_bin, n_a), xib(n_bin, n_a), km(n_bin, n_a), wm(n_bin, n_a), mw(n_a)
        type(binning) :: bins
        call create_bins(x_min=x_min, x_max=x_max, n_bin=n_bin, p=p, spacing=spacing, bins=bins)
        call tangential_shear_tpcf(xa, ya, za, va, wa, xb, yb, zb, e1, e2, wb, mb, bins, &
                                   theta, gt, gx, w, npairs, .false., int(n_threads), use_tree /= 0, .false., &
                                   .true., xie, xib, km, wm, mw)
        call destroy_bins(bins)
    end subroutine ref_shear_c

    subroutine ref_kappa_c(n_a, xa, ya, va, wa, n_b, xb, yb, ke, kb, wb, &
                           x_min, x_max, n_bin, spacing, p, n_threads, use_tree, &
                           theta, gt, gx, w, npairs, xie, xib, km, wm, mw) bind(C, name="ref_kappa_c")
        integer(c_int), value :: n_a, n_b, n_bin, spacing, n_threads, use_tree
        real(c_float), value :: x_min, x_max, p
        real(c_float), intent(in) :: xa(n_a), ya(n_a), va(n_a), wa(n_a)
        real(c_float), intent(in) :: xb(n_b), yb(n_b), ke(n_b), kb(n_b), wb(n_b)
        real(c_float), intent(out) :: theta(n_bin), gt(n_bin), gx(n_bin), w(n_bin)
        integer(c_int64_t), intent(out) :: npairs(n_bin)
        real(c_float), intent(out) :: xie(n_bin, n_a), xib(n_bin, n_a), km(n_bin, n_a), wm(n_bin, n_a), mw(n_a)
        type(binning) :: bins
        call create_bins(x_min=x_min, x_max=x_max, n_bin=n_bin, p=p, spacing=spacing, bins=bins)
        call scalar_kappa_tpcf(xa, ya, va, wa, xb, yb, ke, kb, wb, bins, &
                               theta, gt, gx, w, npairs, .false., int(n_threads), use_tree /= 0, .false., &
                               .true., xie, xib, km, wm, mw)
        call destroy_bins(bins)
    end subroutine ref_kappa_c

    subroutine ref_bins_c(x_min, x_max, n_bin, spacing, p, n, x, edges, centers, idx) bind(C, name="ref_bins_c")
        real(c_float), value :: x_min, x_max, p
        integer(c_int), value :: n_bin, spacing, n
        real(c_float), intent(in) :: x(n)
        real(c_float), intent(out) :: edges(n_bin + 1), centers(n_bin)
        integer(c_int), intent(out) :: idx(n)
        type(binning) :: bins
        integer :: i
        call create_bins(x_min=x_min, x_max=x_max, n_bin=n_bin, p=p, spacing=spacing, bins=bins)
        edges = bins%bin_edges
        centers = bins%bin_centers
        do i = 1, n
            idx(i) = find_bin_index(x(i), bins)
        end do
        call destroy_bins(bins)
    end subroutine ref_bins_c

end module ref_binding
