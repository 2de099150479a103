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
! n_threads is passed to the OpenMP region as num_threads, sph /= 0 sets
! spherical_coords.  ref_sph_pair_c evaluates distance_sphere and
! angles_sphere (src/utils.f90:512-528, :618-639) for arrays of pairs.
module ref_binding
    use, intrinsic :: iso_c_binding
    use utils
    use kmeans_module
    use ref_tpcf_pairs
    implicit none
contains

    ! distance_euclidean is elemental, which may not be passed as an actual
    ! argument: this wrapper has the interface partition_data expects.
    function ref_euclid(x1, y1, x2, y2) result(d)
        real, intent(in) :: x1, y1, x2, y2
        real :: d
        d = distance_euclidean(x1, y1, x2, y2)
    end function ref_euclid

    ! qsort_arg (src/utils.f90:239-298): the 1-based sorting permutation of a.
    subroutine ref_qsort_arg_c(n, a, order, reverse) bind(C, name="ref_qsort_arg_c")
        integer(c_int), value :: n, reverse
        real(c_float), intent(in) :: a(n)
        integer(c_int), intent(out) :: order(n)
        call qsort_arg(a, order, reverse=(reverse /= 0))
    end subroutine ref_qsort_arg_c

    ! partition_data (src/kmeans.f90:12-263) with fixed initial centroids and
    ! one thread; labels are 1-based as the reference returns them.
    subroutine ref_kmeans_c(n, x, y, w, k, equal_size, scaling, iterations, reassign_tol, weight_tol, &
                            labels, centroids) bind(C, name="ref_kmeans_c")
        integer(c_int), value :: n, k, equal_size, iterations
        real(c_float), value :: scaling, reassign_tol, weight_tol
        real(c_float), intent(in) :: x(n), y(n), w(n)
        integer(c_int), intent(out) :: labels(n)
        real(c_float), intent(inout) :: centroids(k, 2)
        real, allocatable :: pts(:, :)
        integer, allocatable :: lab(:)
        allocate(pts(n, 2), lab(n))
        pts(:, 1) = x
        pts(:, 2) = y
        call partition_data(pts, w, k, lab, centroids, ref_euclid, equal_size /= 0, &
                            kmeans_compute_fixed_initial_centroids, 1, scaling, iterations, reassign_tol, weight_tol)
        labels = lab
    end subroutine ref_kmeans_c

    subroutine ref_shear_c(n_a, xa, ya, za, va, wa, n_b, xb, yb, zb, e1, e2, wb, mb, &
                           x_min, x_max, n_bin, spacing, p, n_threads, use_tree, &
                           theta, gt, gx, w, npairs, xie, xib, km, wm, mw, sph) bind(C, name="ref_shear_c")
        integer(c_int), value :: n_a, n_b, n_bin, spacing, n_threads, use_tree, sph
        real(c_float), value :: x_min, x_max, p
        real(c_float), intent(in) :: xa(n_a), ya(n_a), za(n_a), va(n_a), wa(n_a)
        real(c_float), intent(in) :: xb(n_b), yb(n_b), zb(n_b), e1(n_b), e2(n_b), wb(n_b), mb(n_b)
        real(c_float), intent(out) :: theta(n_bin), gt(n_bin), gx(n_bin), w(n_bin)
        integer(c_int64_t), intent(out) :: npairs(n_bin)
        real(c_float), intent(out) :: xie(n_bin, n_a), xib(n_bin, n_a), km(n_bin, n_a), wm(n_bin, n_a), mw(n_a)
        type(binning) :: bins
        call create_bins(x_min=x_min, x_max=x_max, n_bin=n_bin, p=p, spacing=spacing, bins=bins)
        call tangential_shear_tpcf(xa, ya, za, va, wa, xb, yb, zb, e1, e2, wb, mb, bins, &
                                   theta, gt, gx, w, npairs, sph /= 0, int(n_threads), use_tree /= 0, .false., &
                                   .true., xie, xib, km, wm, mw)
        call destroy_bins(bins)
    end subroutine ref_shear_c

    subroutine ref_kappa_c(n_a, xa, ya, va, wa, n_b, xb, yb, ke, kb, wb, &
                           x_min, x_max, n_bin, spacing, p, n_threads, use_tree, &
                           theta, gt, gx, w, npairs, xie, xib, km, wm, mw, sph) bind(C, name="ref_kappa_c")
        integer(c_int), value :: n_a, n_b, n_bin, spacing, n_threads, use_tree, sph
        real(c_float), value :: x_min, x_max, p
        real(c_float), intent(in) :: xa(n_a), ya(n_a), va(n_a), wa(n_a)
        real(c_float), intent(in) :: xb(n_b), yb(n_b), ke(n_b), kb(n_b), wb(n_b)
        real(c_float), intent(out) :: theta(n_bin), gt(n_bin), gx(n_bin), w(n_bin)
        integer(c_int64_t), intent(out) :: npairs(n_bin)
        real(c_float), intent(out) :: xie(n_bin, n_a), xib(n_bin, n_a), km(n_bin, n_a), wm(n_bin, n_a), mw(n_a)
        type(binning) :: bins
        call create_bins(x_min=x_min, x_max=x_max, n_bin=n_bin, p=p, spacing=spacing, bins=bins)
        call scalar_kappa_tpcf(xa, ya, va, wa, xb, yb, ke, kb, wb, bins, &
                               theta, gt, gx, w, npairs, sph /= 0, int(n_threads), use_tree /= 0, .false., &
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

    subroutine ref_covariance_c(n_samples, n_data, x, cov) bind(C, name="ref_covariance_c")
        integer(c_int), value :: n_samples, n_data
        real(c_float), intent(in) :: x(n_samples, n_data)
        real(c_float), intent(out) :: cov(n_data, n_data)
        call covariance_matrix(x, cov)
    end subroutine ref_covariance_c

    subroutine ref_sph_pair_c(n, lon1, lat1, lon2, lat2, d, c2, s2) bind(C, name="ref_sph_pair_c")
        integer(c_int), value :: n
        real(c_float), intent(in) :: lon1(n), lat1(n), lon2(n), lat2(n)
        real(c_float), intent(out) :: d(n), c2(n), s2(n)
        real :: cs(2)
        integer :: i
        do i = 1, n
            d(i) = distance_sphere(lon1(i), lat1(i), lon2(i), lat2(i))
            cs = angles_sphere(lon1(i), lat1(i), lon2(i), lat2(i), d(i))
            c2(i) = 2.0*cs(1)**2 - 1.0
            s2(i) = 2.0*cs(1)*cs(2)
        end do
    end subroutine ref_sph_pair_c

    ! Writes out what the Fortran runtime holds for standard output (the
    ! reference's find_bin_index prints when FP32 rounding puts a distance
    ! past the last bin, src/utils.f90:132-136), so a caller that points
    ! standard output elsewhere around a call can restore it afterwards.
    subroutine ref_flush_c() bind(C, name="ref_flush_c")
        flush(6)
    end subroutine ref_flush_c

end module ref_binding
