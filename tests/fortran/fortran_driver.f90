! fortran_driver.f90 -- a Fortran host program that calls the GPU library the
! way the reference's driver calls its correlation (src/main.f90:596-660):
! create_bins, then shear_2pcf or kappa_2pcf on a lensing_survey_field, through
! module tpcf_gpu.  Used by tests/test_gpu_fortran.py.
!
! Usage: fortran_driver IN OUT
!   IN : int32 kind (0 shear, 1 kappa), spacing, n_bin, n_A, n_B, n_patch_A, n_patch_B
!        real  x_min, x_max, p
!        real  lens x, y, z, val, w (n_A each)
!        real  shear: x, y, z, e1, e2, w, m (n_B each); kappa: x, y, kappa_E, kappa_B, w
!        int32 patch_A (n_A), patch_B (n_B)   -- only when n_patch_A * n_patch_B > 1
!   OUT: real theta, mean_theta, g_t, g_x, weights (n_bin each); int64 n_pairs (n_bin);
!        real xi_E_marks, xi_B_marks, K_marks, w_marks (n_bin, n_A each); real marks_weight (n_A);
!        then, with patches: int64 pp_n, real(8) pp_theta, pp_g_t, pp_g_x, pp_w, pp_k (n_bin, P_A, P_B)
program fortran_driver
    use, intrinsic :: iso_c_binding, only: c_int32_t
    use tpcf_gpu
    implicit none
    character(len=1024) :: fin, fout
    integer :: u, v, kind, spacing, n_bin, n_A, n_B, n_pa, n_pb
    real :: x_min, x_max, p
    type(binning) :: bins
    type(lensing_survey_field) :: field
    type(shear_2pcf_results) :: res
    type(patch_pair_sums) :: pp
    integer(c_int32_t), allocatable :: patch_A(:), patch_B(:)

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) kind, spacing, n_bin, n_A, n_B, n_pa, n_pb
    read(u) x_min, x_max, p

    field%lenses%n = n_A
    allocate(field%lenses%x(n_A), field%lenses%y(n_A), field%lenses%z(n_A), field%lenses%val(n_A), &
             field%lenses%w(n_A))
    read(u) field%lenses%x, field%lenses%y, field%lenses%z, field%lenses%val, field%lenses%w
    if (kind == 0) then
        field%shapes%n = n_B
        allocate(field%shapes%x(n_B), field%shapes%y(n_B), field%shapes%z(n_B), field%shapes%e1(n_B), &
                 field%shapes%e2(n_B), field%shapes%w(n_B), field%shapes%m(n_B))
        read(u) field%shapes%x, field%shapes%y, field%shapes%z, field%shapes%e1, field%shapes%e2, &
                field%shapes%w, field%shapes%m
    else
        field%convergence%n = n_B
        allocate(field%convergence%x(n_B), field%convergence%y(n_B), field%convergence%kappa_E(n_B), &
                 field%convergence%kappa_B(n_B), field%convergence%w(n_B))
        read(u) field%convergence%x, field%convergence%y, field%convergence%kappa_E, field%convergence%kappa_B, &
                field%convergence%w
    end if
    if (n_pa * n_pb > 1) then
        allocate(patch_A(n_A), patch_B(n_B))
        read(u) patch_A, patch_B
    end if
    close(u)

    call create_bins(x_min=x_min, x_max=x_max, n_bin=n_bin, p=p, spacing=spacing, bins=bins)
    if (n_pa * n_pb > 1) then
        if (kind == 0) then
            call shear_2pcf_patches(field, bins, res, patch_A, patch_B, n_pa, n_pb, pp)
        else
            call kappa_2pcf_patches(field, bins, res, patch_A, patch_B, n_pa, n_pb, pp)
        end if
    else
        if (kind == 0) then
            call shear_2pcf(field, bins, res, .false., 1, .true., .false.)
        else
            call kappa_2pcf(field, bins, res, .false., 1, .true., .false.)
        end if
    end if

    open(newunit=v, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(v) res%theta, res%mean_theta, res%g_t, res%g_x, res%weights, res%n_pairs
    if (n_A > 0) write(v) res%xi_E_marks, res%xi_B_marks, res%K_marks, res%w_marks
    write(v) res%marks_weight
    if (n_pa * n_pb > 1) write(v) pp%n, pp%theta, pp%g_t, pp%g_x, pp%w, pp%k
    close(v)
    call deallocate_shear_2pcf_results(res)
    call destroy_bins(bins)
    call tpcf_gpu_finalize()
end program fortran_driver
