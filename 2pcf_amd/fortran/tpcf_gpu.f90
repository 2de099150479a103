! tpcf_gpu.f90 -- Fortran host side of the gfx950 pair-counting library.
!
! module tpcf_gpu_c   ISO_C_BINDING interfaces of include/tpcf_gpu.h, one to one.
! module tpcf_gpu     the reference's correlation call with its own signature:
!                       call shear_2pcf(field, bins, results, spherical_coords, num_threads, use_kd_tree, verbose)
!                       call kappa_2pcf(field, bins, results, spherical_coords, num_threads, use_kd_tree, verbose)
!                     (src/tpcf.f90:134-160, :162-188) running on the GPU, with the
!                     derived types those calls take (src/tpcf.f90:17-43,
!                     src/utils.f90:18-24) and create_bins (src/utils.f90:139-230),
!                     so that a driver swaps "use tpcf" / "use utils, only: binning,
!                     create_bins" for "use tpcf_gpu".  Extension: shear_2pcf_patches /
!                     kappa_2pcf_patches also return the per-patch-pair raw sums;
!                     partition_data_gpu (the Voronoi form of partition_data,
!                     src/kmeans.f90:12-263) and block_bootstrap_gpu (the resampling
!                     of src/tpcf.f90:1165-1191 and covariance_matrix,
!                     src/utils.f90:745-764) run the bootstrap chain on the GPU.
!
! Errors follow the reference's style (src/utils.f90:36-45): a message and
! "error stop".  The default real kind may be 4 (FP32 path) or 8 (a
! -fdefault-real-8 build, FP64 path).
module tpcf_gpu_c
    use, intrinsic :: iso_c_binding
    implicit none

    integer(c_int32_t), parameter :: TPCF_MAX_BIN = 1024
    integer(c_int32_t), parameter :: TPCF_LIN = 0, TPCF_LOG = 1, TPCF_IHS = 2, TPCF_SQRT = 3
    integer(c_int32_t), parameter :: TPCF_OK = 0, TPCF_ERR_ARG = 1, TPCF_ERR_BINS = 2, TPCF_ERR_ALLOC = 3, &
                                     TPCF_ERR_HIP = 4, TPCF_ERR_UNSUPPORTED = 5, TPCF_ERR_NONFINITE = 6, &
                                     TPCF_ERR_NODEVICE = 7, TPCF_ERR_PATCH = 8, TPCF_ERR_GRID = 9, &
                                     TPCF_ERR_MASK = 10
    integer(c_int32_t), parameter :: TPCF_DEVICE_INPUTS = 1, TPCF_SPHERICAL = 2, TPCF_VERBOSE = 4, &
                                     TPCF_NO_KD_TREE = 8, TPCF_DEVICE_MARKS = 64
    integer(c_int32_t), parameter :: TPCF_OPT_CELL_SIZE = 1, TPCF_OPT_COUNTERS = 2, TPCF_OPT_CELL_TARGET = 3, &
                                     TPCF_OPT_SPLIT = 4

    type, bind(C) :: tpcf_bins
        real(c_float) :: x_min, x_max, bin_size
        integer(c_int32_t) :: n_bin, spacing
        real(c_float) :: ihs_theta
        real(c_double) :: edges(TPCF_MAX_BIN + 1)
        real(c_float) :: bin_centers(TPCF_MAX_BIN)
    end type tpcf_bins

    type, bind(C) :: tpcf_out
        integer(c_int32_t) :: real_bytes = 4
        type(c_ptr) :: theta = c_null_ptr, mean_theta = c_null_ptr, g_t = c_null_ptr, g_x = c_null_ptr, &
                       weights = c_null_ptr, n_pairs = c_null_ptr
        type(c_ptr) :: sum_theta = c_null_ptr, sum_gt = c_null_ptr, sum_gx = c_null_ptr, sum_w = c_null_ptr, &
                       sum_k = c_null_ptr
        type(c_ptr) :: xi_E_marks = c_null_ptr, xi_B_marks = c_null_ptr, K_marks = c_null_ptr, &
                       w_marks = c_null_ptr, marks_weight = c_null_ptr
        type(c_ptr) :: pp_n = c_null_ptr, pp_theta = c_null_ptr, pp_gt = c_null_ptr, pp_gx = c_null_ptr, &
                       pp_w = c_null_ptr, pp_k = c_null_ptr
    end type tpcf_out

    type, bind(C) :: tpcf_kmeans_opts
        integer(c_int32_t) :: k = 0, max_iterations = 1000, init = 0, equal_size = 0
        real(c_double) :: reassignment_tol = 0.0d0, weight_equalizer_scaling = 0.0d0, weight_tol = 0.0d0
    end type tpcf_kmeans_opts

    type, bind(C) :: tpcf_blocks
        integer(c_int32_t) :: n_block = 0, n_bin = 0
        type(c_ptr) :: g_t = c_null_ptr, g_x = c_null_ptr, k = c_null_ptr, weight = c_null_ptr
    end type tpcf_blocks

    interface
        integer(c_int) function tpcf_create_bins(x_min, x_max, n_bin, spacing, ihs_theta, bins) bind(C)
            import :: c_int, c_float, c_int32_t, tpcf_bins
            real(c_float), value :: x_min, x_max
            integer(c_int32_t), value :: n_bin, spacing
            real(c_float), value :: ihs_theta
            type(tpcf_bins), intent(out) :: bins
        end function tpcf_create_bins

        integer(c_int32_t) function tpcf_find_bin_index(x, bins) bind(C)
            import :: c_int32_t, c_double, tpcf_bins
            real(c_double), value :: x
            type(tpcf_bins), intent(in) :: bins
        end function tpcf_find_bin_index

        integer(c_int) function tpcf_ctx_create(device, ctx) bind(C)
            import :: c_int, c_int32_t, c_ptr
            integer(c_int32_t), value :: device
            type(c_ptr), intent(out) :: ctx
        end function tpcf_ctx_create

        integer(c_int) function tpcf_ctx_destroy(ctx) bind(C)
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function tpcf_ctx_destroy

        integer(c_int) function tpcf_set_option(ctx, option, val) bind(C)
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: option
            real(c_double), value :: val
        end function tpcf_set_option

        type(c_ptr) function tpcf_last_error(ctx) bind(C)
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function tpcf_last_error

        ! The four correlation calls take every array as a C pointer so that
        ! one interface serves both kinds of default real.
        integer(c_int) function tpcf_shear_f32(ctx, n_A, x_A, y_A, z_A, val_A, w_A, n_B, x_B, y_B, z_B, e1_B, e2_B, &
                                               w_B, m_B, bins, flags, patch_A, patch_B, n_patch_A, n_patch_B, out) &
                                               bind(C)
            import :: c_int, c_int32_t, c_int64_t, c_ptr, tpcf_bins, tpcf_out
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: n_A, n_B
            type(c_ptr), value :: x_A, y_A, z_A, val_A, w_A, x_B, y_B, z_B, e1_B, e2_B, w_B, m_B
            type(tpcf_bins), intent(in) :: bins
            integer(c_int32_t), value :: flags
            type(c_ptr), value :: patch_A, patch_B
            integer(c_int32_t), value :: n_patch_A, n_patch_B
            type(tpcf_out), intent(inout) :: out
        end function tpcf_shear_f32

        integer(c_int) function tpcf_shear_f64(ctx, n_A, x_A, y_A, z_A, val_A, w_A, n_B, x_B, y_B, z_B, e1_B, e2_B, &
                                               w_B, m_B, bins, flags, patch_A, patch_B, n_patch_A, n_patch_B, out) &
                                               bind(C)
            import :: c_int, c_int32_t, c_int64_t, c_ptr, tpcf_bins, tpcf_out
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: n_A, n_B
            type(c_ptr), value :: x_A, y_A, z_A, val_A, w_A, x_B, y_B, z_B, e1_B, e2_B, w_B, m_B
            type(tpcf_bins), intent(in) :: bins
            integer(c_int32_t), value :: flags
            type(c_ptr), value :: patch_A, patch_B
            integer(c_int32_t), value :: n_patch_A, n_patch_B
            type(tpcf_out), intent(inout) :: out
        end function tpcf_shear_f64

        integer(c_int) function tpcf_kappa_f32(ctx, n_A, x_A, y_A, val_A, w_A, n_B, x_B, y_B, kappa_E, kappa_B, w_B, &
                                               bins, flags, patch_A, patch_B, n_patch_A, n_patch_B, out) bind(C)
            import :: c_int, c_int32_t, c_int64_t, c_ptr, tpcf_bins, tpcf_out
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: n_A, n_B
            type(c_ptr), value :: x_A, y_A, val_A, w_A, x_B, y_B, kappa_E, kappa_B, w_B
            type(tpcf_bins), intent(in) :: bins
            integer(c_int32_t), value :: flags
            type(c_ptr), value :: patch_A, patch_B
            integer(c_int32_t), value :: n_patch_A, n_patch_B
            type(tpcf_out), intent(inout) :: out
        end function tpcf_kappa_f32

        integer(c_int) function tpcf_kappa_f64(ctx, n_A, x_A, y_A, val_A, w_A, n_B, x_B, y_B, kappa_E, kappa_B, w_B, &
                                               bins, flags, patch_A, patch_B, n_patch_A, n_patch_B, out) bind(C)
            import :: c_int, c_int32_t, c_int64_t, c_ptr, tpcf_bins, tpcf_out
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: n_A, n_B
            type(c_ptr), value :: x_A, y_A, val_A, w_A, x_B, y_B, kappa_E, kappa_B, w_B
            type(tpcf_bins), intent(in) :: bins
            integer(c_int32_t), value :: flags
            type(c_ptr), value :: patch_A, patch_B
            integer(c_int32_t), value :: n_patch_A, n_patch_B
            type(tpcf_out), intent(inout) :: out
        end function tpcf_kappa_f64

        integer(c_int) function tpcf_kmeans_f32(ctx, n, x, y, w, opts, flags, labels, centroids, cluster_weights, &
                                                rounds, last_moves) bind(C)
            import :: c_ptr, c_int, c_int32_t, c_int64_t, tpcf_kmeans_opts
            type(c_ptr), value :: ctx, x, y, w, labels, centroids, cluster_weights
            integer(c_int64_t), value :: n
            type(tpcf_kmeans_opts), intent(in) :: opts
            integer(c_int32_t), value :: flags
            integer(c_int32_t), intent(out) :: rounds
            integer(c_int64_t), intent(out) :: last_moves
        end function tpcf_kmeans_f32

        integer(c_int) function tpcf_kmeans_f64(ctx, n, x, y, w, opts, flags, labels, centroids, cluster_weights, &
                                                rounds, last_moves) bind(C)
            import :: c_ptr, c_int, c_int32_t, c_int64_t, tpcf_kmeans_opts
            type(c_ptr), value :: ctx, x, y, w, labels, centroids, cluster_weights
            integer(c_int64_t), value :: n
            type(tpcf_kmeans_opts), intent(in) :: opts
            integer(c_int32_t), value :: flags
            integer(c_int32_t), intent(out) :: rounds
            integer(c_int64_t), intent(out) :: last_moves
        end function tpcf_kmeans_f64

        integer(c_int) function tpcf_bootstrap(ctx, blocks, n_boot, multiplicity, seed, supersampling, flags, &
                                               g_t, g_x, weights, cov_t, cov_x, multiplicity_out) bind(C)
            import :: c_ptr, c_int, c_int32_t, c_int64_t, c_double, tpcf_blocks
            type(c_ptr), value :: ctx, multiplicity, g_t, g_x, weights, cov_t, cov_x, multiplicity_out
            type(tpcf_blocks), intent(in) :: blocks
            integer(c_int32_t), value :: n_boot, flags
            integer(c_int64_t), value :: seed
            real(c_double), value :: supersampling
        end function tpcf_bootstrap
    end interface
end module tpcf_gpu_c


module tpcf_gpu
    use, intrinsic :: iso_c_binding
    use tpcf_gpu_c
    implicit none

    private
    public lin_binning, log_binning, ihs_binning, sqrt_binning
    public binning, create_bins, destroy_bins, find_bin_index
    public shape_catalog, kappa_map, lens_catalog, lensing_survey_field, shear_2pcf_results, patch_pair_sums
    public allocate_shear_2pcf_results, deallocate_shear_2pcf_results
    public shear_2pcf, kappa_2pcf, shear_2pcf_patches, kappa_2pcf_patches
    public tpcf_gpu_finalize, tpcf_gpu_set_option
    public partition_data_gpu, block_bootstrap_gpu

    integer, parameter :: lin_binning = 0, log_binning = 1, ihs_binning = 2, sqrt_binning = 3   ! src/utils.f90:7

    ! The component names and kinds of type(binning), src/utils.f90:18-24.
    type binning
        real :: x_min, x_max, bin_size = 0.0
        integer :: n_bin = 0
        integer :: spacing
        real :: ihs_theta
        real, allocatable :: bin_edges(:), bin_centers(:), bin_centers_eff(:)
    end type binning

    ! The derived types of the correlation call, src/tpcf.f90:17-43.
    type shear_2pcf_results
        integer :: n_bin
        real, allocatable, dimension(:) :: theta, mean_theta, g_t, g_x, weights, marks_weight
        real, allocatable, dimension(:,:) :: xi_E_marks, xi_B_marks, K_marks, w_marks
        integer(kind=8), allocatable, dimension(:) :: n_pairs
    end type shear_2pcf_results

    type shape_catalog
        integer :: n
        real, allocatable, dimension(:) :: x, y, z, e1, e2, m, w
    end type shape_catalog

    type kappa_map
        integer :: n
        real, allocatable, dimension(:) :: x, y, kappa_E, kappa_B, w
    end type kappa_map

    type lens_catalog
        integer :: n, n_bin
        real, allocatable, dimension(:) :: x, y, z, val, w, marks_weight
    end type lens_catalog

    type lensing_survey_field
        type(shape_catalog) :: shapes
        type(lens_catalog) :: lenses
        type(kappa_map) :: convergence
    end type lensing_survey_field

    ! Extension: raw FP64 sums per (bin, lens patch, source patch).  Summed
    ! over the source patch they are the per-block sums of src/tpcf.f90:1130-1137.
    type patch_pair_sums
        integer :: n_bin = 0, n_patch_A = 0, n_patch_B = 0
        integer(kind=8), allocatable :: n(:,:,:)
        real(kind=8), allocatable :: theta(:,:,:), g_t(:,:,:), g_x(:,:,:), w(:,:,:), k(:,:,:)
    end type patch_pair_sums

    type(c_ptr), save :: ctx = c_null_ptr

contains

    subroutine check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        character(len=512) :: text
        integer :: i
        if (rc == TPCF_OK) return
        text = ""
        if (c_associated(ctx)) then
            call c_f_pointer(tpcf_last_error(ctx), msg, [512])
            do i = 1, 512
                if (msg(i) == c_null_char) exit
                text(i:i) = msg(i)
            end do
        end if
        print *, "tpcf_gpu: ", what, " failed with status ", rc, ": ", trim(text)
        error stop
    end subroutine check

    subroutine ensure_ctx()
        if (.not. c_associated(ctx)) call check(tpcf_ctx_create(-1_c_int32_t, ctx), "tpcf_ctx_create")
    end subroutine ensure_ctx

    !> Releases the device context (device buffers, stream).
    subroutine tpcf_gpu_finalize()
        integer(c_int) :: rc
        if (c_associated(ctx)) rc = tpcf_ctx_destroy(ctx)
        ctx = c_null_ptr
    end subroutine tpcf_gpu_finalize

    !> tpcf_set_option on the module's context (cell size, work split, ...).
    subroutine tpcf_gpu_set_option(option, val)
        integer, intent(in) :: option
        real(kind=8), intent(in) :: val
        call ensure_ctx()
        call check(tpcf_set_option(ctx, int(option, c_int32_t), real(val, c_double)), "tpcf_set_option")
    end subroutine tpcf_gpu_set_option

    !> create_bins(x_min, x_max, bin_size, n_bin, p, spacing, bins),
    !> src/utils.f90:139-230: the same optional-argument forms; the edges are
    !> those the library bins with (FP64 closed form, rounded to real).
    subroutine create_bins(x_min, x_max, bin_size, n_bin, p, spacing, bins)
        real, optional, intent(in) :: x_min, x_max, bin_size
        integer, optional, intent(in) :: n_bin
        real, optional, intent(in) :: p
        integer, intent(in) :: spacing
        type(binning), intent(out) :: bins
        type(tpcf_bins) :: cb
        real :: theta, hi
        integer :: nb
        theta = 1.0
        if (present(p)) theta = p
        if (spacing < lin_binning .or. spacing > sqrt_binning) then
            print *, "Bin spacing unknown:", spacing
            error stop
        end if
        if (present(x_min) .and. present(x_max) .and. present(n_bin) .and. .not. present(bin_size)) then
            hi = x_max
            nb = n_bin
        else if (present(x_min) .and. present(bin_size) .and. present(n_bin) .and. .not. present(x_max)) then
            nb = n_bin
            select case (spacing)
            case (lin_binning)
                hi = x_min + bin_size * n_bin
            case (log_binning)
                hi = 10**(log10(x_min) + bin_size * n_bin)
            case (ihs_binning)
                hi = sinh((asinh(theta * x_min) / theta + bin_size * n_bin) * theta) / theta
            case default
                hi = (sqrt(x_min) + bin_size * n_bin)**2
            end select
        else
            print *, "Binning is ill-defined."
            error stop
        end if
        if (tpcf_create_bins(real(x_min, c_float), real(hi, c_float), int(nb, c_int32_t), int(spacing, c_int32_t), &
                             real(theta, c_float), cb) /= TPCF_OK) then
            print *, "Binning is ill-defined."
            error stop
        end if
        call from_c_bins(cb, bins)
    end subroutine create_bins

    subroutine from_c_bins(cb, bins)
        type(tpcf_bins), intent(in) :: cb
        type(binning), intent(out) :: bins
        bins%x_min = cb%x_min
        bins%x_max = cb%x_max
        bins%bin_size = cb%bin_size
        bins%n_bin = cb%n_bin
        bins%spacing = cb%spacing
        bins%ihs_theta = cb%ihs_theta
        allocate(bins%bin_edges(cb%n_bin + 1), bins%bin_centers(cb%n_bin))
        bins%bin_edges = real(cb%edges(1:cb%n_bin + 1))
        bins%bin_centers = real(cb%bin_centers(1:cb%n_bin))
    end subroutine from_c_bins

    subroutine to_c_bins(bins, cb)
        type(binning), intent(in) :: bins
        type(tpcf_bins), intent(out) :: cb
        real :: theta
        theta = 1.0
        if (bins%spacing == ihs_binning) theta = bins%ihs_theta
        if (tpcf_create_bins(real(bins%x_min, c_float), real(bins%x_max, c_float), int(bins%n_bin, c_int32_t), &
                             int(bins%spacing, c_int32_t), real(theta, c_float), cb) /= TPCF_OK) then
            print *, "Binning is ill-defined."
            error stop
        end if
    end subroutine to_c_bins

    subroutine destroy_bins(bins)
        type(binning) :: bins
        if (allocated(bins%bin_edges)) deallocate(bins%bin_edges)
        if (allocated(bins%bin_centers)) deallocate(bins%bin_centers)
    end subroutine destroy_bins

    !> find_bin_index (src/utils.f90:101-137) under the library's bin
    !> definition: 1-based, -1 outside [x_min, x_max].
    integer function find_bin_index(x, bins) result(i)
        real, intent(in) :: x
        type(binning), intent(in) :: bins
        type(tpcf_bins) :: cb
        call to_c_bins(bins, cb)
        i = tpcf_find_bin_index(real(x, c_double), cb)
        if (i >= 0) i = i + 1
    end function find_bin_index

    !> allocate_shear_2pcf_results, src/tpcf.f90:99-115.
    subroutine allocate_shear_2pcf_results(results, n_bin, n_A)
        type(shear_2pcf_results), intent(out) :: results
        integer, intent(in) :: n_bin, n_A
        if (.not. allocated(results%theta)) allocate(results%theta(n_bin), source=0.0)
        if (.not. allocated(results%mean_theta)) allocate(results%mean_theta(n_bin), source=0.0)
        if (.not. allocated(results%g_t)) allocate(results%g_t(n_bin), source=0.0)
        if (.not. allocated(results%g_x)) allocate(results%g_x(n_bin), source=0.0)
        if (.not. allocated(results%weights)) allocate(results%weights(n_bin), source=0.0)
        if (.not. allocated(results%n_pairs)) allocate(results%n_pairs(n_bin), source=int(0, kind=8))
        if (.not. allocated(results%marks_weight)) allocate(results%marks_weight(n_A), source=0.0)
        if (.not. allocated(results%xi_E_marks) .and. n_A > 0) allocate(results%xi_E_marks(n_bin, n_A), source=0.0)
        if (.not. allocated(results%xi_B_marks) .and. n_A > 0) allocate(results%xi_B_marks(n_bin, n_A), source=0.0)
        if (.not. allocated(results%K_marks) .and. n_A > 0) allocate(results%K_marks(n_bin, n_A), source=0.0)
        if (.not. allocated(results%w_marks) .and. n_A > 0) allocate(results%w_marks(n_bin, n_A), source=0.0)
    end subroutine allocate_shear_2pcf_results

    subroutine deallocate_shear_2pcf_results(results)
        type(shear_2pcf_results), intent(out) :: results
        ! intent(out) deallocates every allocatable component
        results%n_bin = 0
    end subroutine deallocate_shear_2pcf_results

    ! c_loc of an array, or NULL for a zero-sized one (c_loc needs size > 0).
    function ptr(a) result(p)
        real, intent(in), target :: a(:)
        type(c_ptr) :: p
        p = c_null_ptr
        if (size(a) > 0) p = c_loc(a(1))
    end function ptr

    function ptr2(a) result(p)
        real, intent(in), target :: a(:,:)
        type(c_ptr) :: p
        p = c_null_ptr
        if (size(a) > 0) p = c_loc(a(1, 1))
    end function ptr2

    subroutine results_out(results, out)
        type(shear_2pcf_results), intent(inout), target :: results
        type(tpcf_out), intent(out) :: out
        out%real_bytes = int(c_sizeof(results%weights(1:1)), c_int32_t)
        out%mean_theta = ptr(results%mean_theta)
        out%g_t = ptr(results%g_t)
        out%g_x = ptr(results%g_x)
        out%weights = ptr(results%weights)
        out%n_pairs = c_loc(results%n_pairs(1))
        out%marks_weight = ptr(results%marks_weight)
        if (allocated(results%xi_E_marks)) then
            out%xi_E_marks = ptr2(results%xi_E_marks)
            out%xi_B_marks = ptr2(results%xi_B_marks)
            out%K_marks = ptr2(results%K_marks)
            out%w_marks = ptr2(results%w_marks)
        end if
    end subroutine results_out

    integer(c_int32_t) function call_flags(spherical_coords, use_kd_tree, verbose) result(f)
        logical, intent(in) :: spherical_coords, use_kd_tree, verbose
        f = 0
        if (spherical_coords) f = ior(f, TPCF_SPHERICAL)
        if (verbose) f = ior(f, TPCF_VERBOSE)
        if (.not. use_kd_tree) f = ior(f, TPCF_NO_KD_TREE)
    end function call_flags

    subroutine patch_out(pp, n_bin, n_patch_A, n_patch_B, out)
        type(patch_pair_sums), intent(inout), target :: pp
        integer, intent(in) :: n_bin, n_patch_A, n_patch_B
        type(tpcf_out), intent(inout) :: out
        if (allocated(pp%n)) deallocate(pp%n, pp%theta, pp%g_t, pp%g_x, pp%w, pp%k)
        pp%n_bin = n_bin
        pp%n_patch_A = n_patch_A
        pp%n_patch_B = n_patch_B
        allocate(pp%n(n_bin, n_patch_A, n_patch_B), pp%theta(n_bin, n_patch_A, n_patch_B), &
                 pp%g_t(n_bin, n_patch_A, n_patch_B), pp%g_x(n_bin, n_patch_A, n_patch_B), &
                 pp%w(n_bin, n_patch_A, n_patch_B), pp%k(n_bin, n_patch_A, n_patch_B))
        out%pp_n = c_loc(pp%n(1, 1, 1))
        out%pp_theta = c_loc(pp%theta(1, 1, 1))
        out%pp_gt = c_loc(pp%g_t(1, 1, 1))
        out%pp_gx = c_loc(pp%g_x(1, 1, 1))
        out%pp_w = c_loc(pp%w(1, 1, 1))
        out%pp_k = c_loc(pp%k(1, 1, 1))
    end subroutine patch_out

    subroutine run_shear(field, bins, results, spherical_coords, use_kd_tree, verbose, &
                         patch_A, patch_B, n_patch_A, n_patch_B, pp)
        type(lensing_survey_field), intent(in), target :: field
        type(binning), intent(in) :: bins
        type(shear_2pcf_results), target :: results
        logical, intent(in) :: spherical_coords, use_kd_tree, verbose
        integer(c_int32_t), intent(in), optional, target :: patch_A(:), patch_B(:)
        integer, intent(in), optional :: n_patch_A, n_patch_B
        type(patch_pair_sums), intent(inout), optional, target :: pp
        type(tpcf_bins) :: cb
        type(tpcf_out) :: out
        type(c_ptr) :: pa, pb
        integer(c_int32_t) :: npa, npb
        integer(c_int) :: rc
        call ensure_ctx()
        call allocate_shear_2pcf_results(results, bins%n_bin, field%lenses%n)
        results%n_bin = bins%n_bin
        results%theta = bins%bin_centers
        call to_c_bins(bins, cb)
        call results_out(results, out)
        call patch_args(patch_A, patch_B, n_patch_A, n_patch_B, pa, pb, npa, npb)
        if (present(pp)) call patch_out(pp, bins%n_bin, int(npa), int(npb), out)
        associate (L => field%lenses, S => field%shapes)
            if (out%real_bytes == 4) then
                rc = tpcf_shear_f32(ctx, int(L%n, c_int64_t), ptr(L%x), ptr(L%y), ptr(L%z), ptr(L%val), ptr(L%w), &
                                    int(S%n, c_int64_t), ptr(S%x), ptr(S%y), ptr(S%z), ptr(S%e1), ptr(S%e2), &
                                    ptr(S%w), ptr(S%m), cb, call_flags(spherical_coords, use_kd_tree, verbose), &
                                    pa, pb, npa, npb, out)
            else
                rc = tpcf_shear_f64(ctx, int(L%n, c_int64_t), ptr(L%x), ptr(L%y), ptr(L%z), ptr(L%val), ptr(L%w), &
                                    int(S%n, c_int64_t), ptr(S%x), ptr(S%y), ptr(S%z), ptr(S%e1), ptr(S%e2), &
                                    ptr(S%w), ptr(S%m), cb, call_flags(spherical_coords, use_kd_tree, verbose), &
                                    pa, pb, npa, npb, out)
            end if
        end associate
        call check(rc, "shear_2pcf")
    end subroutine run_shear

    subroutine run_kappa(field, bins, results, spherical_coords, use_kd_tree, verbose, &
                         patch_A, patch_B, n_patch_A, n_patch_B, pp)
        type(lensing_survey_field), intent(in), target :: field
        type(binning), intent(in) :: bins
        type(shear_2pcf_results), target :: results
        logical, intent(in) :: spherical_coords, use_kd_tree, verbose
        integer(c_int32_t), intent(in), optional, target :: patch_A(:), patch_B(:)
        integer, intent(in), optional :: n_patch_A, n_patch_B
        type(patch_pair_sums), intent(inout), optional, target :: pp
        type(tpcf_bins) :: cb
        type(tpcf_out) :: out
        type(c_ptr) :: pa, pb
        integer(c_int32_t) :: npa, npb
        integer(c_int) :: rc
        call ensure_ctx()
        call allocate_shear_2pcf_results(results, bins%n_bin, field%lenses%n)
        results%n_bin = bins%n_bin
        results%theta = bins%bin_centers
        call to_c_bins(bins, cb)
        call results_out(results, out)
        call patch_args(patch_A, patch_B, n_patch_A, n_patch_B, pa, pb, npa, npb)
        if (present(pp)) call patch_out(pp, bins%n_bin, int(npa), int(npb), out)
        associate (L => field%lenses, K => field%convergence)
            if (out%real_bytes == 4) then
                rc = tpcf_kappa_f32(ctx, int(L%n, c_int64_t), ptr(L%x), ptr(L%y), ptr(L%val), ptr(L%w), &
                                    int(K%n, c_int64_t), ptr(K%x), ptr(K%y), ptr(K%kappa_E), ptr(K%kappa_B), &
                                    ptr(K%w), cb, call_flags(spherical_coords, use_kd_tree, verbose), &
                                    pa, pb, npa, npb, out)
            else
                rc = tpcf_kappa_f64(ctx, int(L%n, c_int64_t), ptr(L%x), ptr(L%y), ptr(L%val), ptr(L%w), &
                                    int(K%n, c_int64_t), ptr(K%x), ptr(K%y), ptr(K%kappa_E), ptr(K%kappa_B), &
                                    ptr(K%w), cb, call_flags(spherical_coords, use_kd_tree, verbose), &
                                    pa, pb, npa, npb, out)
            end if
        end associate
        call check(rc, "kappa_2pcf")
    end subroutine run_kappa

    subroutine patch_args(patch_A, patch_B, n_patch_A, n_patch_B, pa, pb, npa, npb)
        integer(c_int32_t), intent(in), optional, target :: patch_A(:), patch_B(:)
        integer, intent(in), optional :: n_patch_A, n_patch_B
        type(c_ptr), intent(out) :: pa, pb
        integer(c_int32_t), intent(out) :: npa, npb
        pa = c_null_ptr
        pb = c_null_ptr
        npa = 1
        npb = 1
        if (present(patch_A)) then
            if (size(patch_A) > 0) pa = c_loc(patch_A(1))
            npa = int(n_patch_A, c_int32_t)
        end if
        if (present(patch_B)) then
            if (size(patch_B) > 0) pb = c_loc(patch_B(1))
            npb = int(n_patch_B, c_int32_t)
        end if
    end subroutine patch_args

    !> shear_2pcf, src/tpcf.f90:134-160 (num_threads is accepted and unused).
    subroutine shear_2pcf(field, bins, results, spherical_coords, num_threads, use_kd_tree, verbose)
        type(lensing_survey_field), intent(in) :: field
        type(binning), intent(in) :: bins
        type(shear_2pcf_results) :: results
        logical, intent(in) :: spherical_coords, use_kd_tree, verbose
        integer, intent(in) :: num_threads
        call run_shear(field, bins, results, spherical_coords, use_kd_tree, verbose)
    end subroutine shear_2pcf

    !> kappa_2pcf, src/tpcf.f90:162-188 (num_threads is accepted and unused).
    subroutine kappa_2pcf(field, bins, results, spherical_coords, num_threads, use_kd_tree, verbose)
        type(lensing_survey_field), intent(in) :: field
        type(binning), intent(in) :: bins
        type(shear_2pcf_results) :: results
        logical, intent(in) :: spherical_coords, use_kd_tree, verbose
        integer, intent(in) :: num_threads
        call run_kappa(field, bins, results, spherical_coords, use_kd_tree, verbose)
    end subroutine kappa_2pcf

    !> shear_2pcf plus per-patch-pair sums; patch labels are 0-based
    !> (0 .. n_patch-1), pp arrays are (n_bin, n_patch_A, n_patch_B).
    subroutine shear_2pcf_patches(field, bins, results, patch_A, patch_B, n_patch_A, n_patch_B, pp, verbose)
        type(lensing_survey_field), intent(in) :: field
        type(binning), intent(in) :: bins
        type(shear_2pcf_results) :: results
        integer(c_int32_t), intent(in) :: patch_A(:), patch_B(:)
        integer, intent(in) :: n_patch_A, n_patch_B
        type(patch_pair_sums), intent(inout) :: pp
        logical, intent(in), optional :: verbose
        logical :: v
        v = .false.
        if (present(verbose)) v = verbose
        call run_shear(field, bins, results, .false., .true., v, patch_A, patch_B, n_patch_A, n_patch_B, pp)
    end subroutine shear_2pcf_patches

    subroutine kappa_2pcf_patches(field, bins, results, patch_A, patch_B, n_patch_A, n_patch_B, pp, verbose)
        type(lensing_survey_field), intent(in) :: field
        type(binning), intent(in) :: bins
        type(shear_2pcf_results) :: results
        integer(c_int32_t), intent(in) :: patch_A(:), patch_B(:)
        integer, intent(in) :: n_patch_A, n_patch_B
        type(patch_pair_sums), intent(inout) :: pp
        logical, intent(in), optional :: verbose
        logical :: v
        v = .false.
        if (present(verbose)) v = verbose
        call run_kappa(field, bins, results, .false., .true., v, patch_A, patch_B, n_patch_A, n_patch_B, pp)
    end subroutine kappa_2pcf_patches

    !> partition_data (src/kmeans.f90:12-263) on the GPU: data_points(n, 2)
    !> and weights as the reference takes them, 1-based cluster_indicies,
    !> cluster_centroids(k, 2).  given_centroids = .true. starts from
    !> cluster_centroids instead of the fixed grid; spherical coordinates are
    !> longitude / latitude in radians.  equal_size, weight_equalizer_scaling
    !> and weight_tol as the reference's arguments (the equaliser moves the
    !> cluster weights after every step of 256 points, not after every point).
    subroutine partition_data_gpu(data_points, weights, k, cluster_indicies, cluster_centroids, spherical_coords, &
                                  given_centroids, iterations, reassignment_tol, equal_size, &
                                  weight_equalizer_scaling, weight_tol)
        real, intent(in) :: data_points(:, :), weights(:)
        integer, intent(in) :: k
        integer, intent(out) :: cluster_indicies(:)
        real, intent(inout) :: cluster_centroids(k, 2)
        logical, optional, intent(in) :: spherical_coords, given_centroids, equal_size
        integer, optional, intent(in) :: iterations
        real, optional, intent(in) :: reassignment_tol, weight_equalizer_scaling, weight_tol
        real, allocatable, target :: x(:), y(:), w(:)
        integer(c_int32_t), allocatable, target :: lab(:)
        real(c_double), target :: cen(2, k), cw(k)
        type(tpcf_kmeans_opts) :: o
        integer(c_int32_t) :: flags, rounds
        integer(c_int64_t) :: moves, n
        integer(c_int) :: rc
        call ensure_ctx()
        n = size(data_points, 1)
        allocate(x(n), y(n), w(n), lab(n))
        x = data_points(:, 1)
        y = data_points(:, 2)
        w = weights
        o%k = k
        if (present(iterations)) o%max_iterations = iterations
        if (present(reassignment_tol)) o%reassignment_tol = real(reassignment_tol, c_double)
        if (present(weight_equalizer_scaling)) o%weight_equalizer_scaling = real(weight_equalizer_scaling, c_double)
        if (present(weight_tol)) o%weight_tol = real(weight_tol, c_double)
        if (present(equal_size)) then
            if (equal_size) o%equal_size = 1
        end if
        o%init = 0
        if (present(given_centroids)) then
            if (given_centroids) o%init = 1
        end if
        cen(1, :) = real(cluster_centroids(:, 1), c_double)
        cen(2, :) = real(cluster_centroids(:, 2), c_double)
        flags = 0
        if (present(spherical_coords)) then
            if (spherical_coords) flags = TPCF_SPHERICAL
        end if
        if (kind(1.0) == 4) then
            rc = tpcf_kmeans_f32(ctx, n, c_loc(x), c_loc(y), c_loc(w), o, flags, c_loc(lab), c_loc(cen), &
                                 c_loc(cw), rounds, moves)
        else
            rc = tpcf_kmeans_f64(ctx, n, c_loc(x), c_loc(y), c_loc(w), o, flags, c_loc(lab), c_loc(cen), &
                                 c_loc(cw), rounds, moves)
        end if
        call check(rc, "tpcf_kmeans")
        cluster_indicies = lab + 1
        cluster_centroids(:, 1) = real(cen(1, :))
        cluster_centroids(:, 2) = real(cen(2, :))
    end subroutine partition_data_gpu

    !> The block bootstrap of src/tpcf.f90:1165-1191 and covariance_matrix
    !> (src/utils.f90:745-764) on the GPU.  block_g_t, block_g_x, block_K
    !> (n_bin, n_blocks) and block_weight(n_blocks) as the reference forms
    !> them (src/tpcf.f90:1130-1137); n_draw = int(n_blocks * supersampling_factor)
    !> draws per sample from a generator keyed by seed; g_t_bootstrap and
    !> g_x_bootstrap (n_bootstrap_samples, n_bin) as src/main.f90:524-532.
    subroutine block_bootstrap_gpu(block_g_t, block_g_x, block_K, block_weight, n_bootstrap_samples, &
                                   supersampling_factor, seed, g_t_bootstrap, g_x_bootstrap, weights, &
                                   g_t_covariance, g_x_covariance)
        real, intent(in) :: block_g_t(:, :), block_g_x(:, :), block_K(:, :), block_weight(:)
        integer, intent(in) :: n_bootstrap_samples
        real, intent(in) :: supersampling_factor
        integer(kind=8), intent(in) :: seed
        real, intent(out) :: g_t_bootstrap(:, :), g_x_bootstrap(:, :), weights(:)
        real, intent(out) :: g_t_covariance(:, :), g_x_covariance(:, :)
        real(c_double), allocatable, target :: bt(:, :), bx(:, :), bk(:, :), bw(:)
        real(c_double), allocatable, target :: ot(:, :), ox(:, :), ow(:), ct(:, :), cx(:, :)
        type(tpcf_blocks) :: b
        integer :: n_bin, n_blk
        call ensure_ctx()
        n_bin = size(block_g_t, 1)
        n_blk = size(block_g_t, 2)
        allocate(bt(n_bin, n_blk), bx(n_bin, n_blk), bk(n_bin, n_blk), bw(n_blk))
        allocate(ot(n_bin, n_bootstrap_samples), ox(n_bin, n_bootstrap_samples), ow(n_bootstrap_samples))
        allocate(ct(n_bin, n_bin), cx(n_bin, n_bin))
        bt = block_g_t
        bx = block_g_x
        bk = block_K
        bw = block_weight
        b%n_block = n_blk
        b%n_bin = n_bin
        b%g_t = c_loc(bt)
        b%g_x = c_loc(bx)
        b%k = c_loc(bk)
        b%weight = c_loc(bw)
        call check(tpcf_bootstrap(ctx, b, int(n_bootstrap_samples, c_int32_t), c_null_ptr, int(seed, c_int64_t), &
                                  real(supersampling_factor, c_double), 0_c_int32_t, c_loc(ot), c_loc(ox), &
                                  c_loc(ow), c_loc(ct), c_loc(cx), c_null_ptr), "tpcf_bootstrap")
        g_t_bootstrap = real(transpose(ot))
        g_x_bootstrap = real(transpose(ox))
        weights = real(ow)
        g_t_covariance = real(ct)
        g_x_covariance = real(cx)
    end subroutine block_bootstrap_gpu

end module tpcf_gpu
