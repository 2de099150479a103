! bins_forms.F90 -- calls create_bins (src/utils.f90:139-230) in one of its
! argument forms and prints n_bin, x_max, bin_size, the edges and the centres.
! Built twice by check_integration_stub.sh: against the reference's module
! utils (-DUSE_REFERENCE) and against this project's module tpcf_gpu, whose
! create_bins must accept and refuse the same forms with the same results.
!
! Form (first argument): 1..4 = (x_min, x_max, n_bin) with spacing lin, log,
! ihs (p = 2), sqrt; 5..8 = (x_min, bin_size, n_bin) with the same spacings;
! 9 = (x_min, x_max, bin_size) with log spacing, which the reference refuses
! ("Binning is ill-defined.", src/utils.f90:163-206).
program bins_forms
#ifdef USE_REFERENCE
    use utils
#else
    use tpcf_gpu
#endif
    implicit none
    type(binning) :: b
    integer :: form, i, sp
    character(len=16) :: arg
    integer, parameter :: spacings(4) = [lin_binning, log_binning, ihs_binning, sqrt_binning]
    real, parameter :: sizes(4) = [2.5, 0.1, 0.05, 0.15]

    call get_command_argument(1, arg)
    read (arg, *) form
    if (form >= 1 .and. form <= 4) then
        sp = spacings(form)
        call create_bins(x_min=1.0, x_max=40.0, n_bin=12, p=2.0, spacing=sp, bins=b)
    else if (form >= 5 .and. form <= 8) then
        sp = spacings(form - 4)
        call create_bins(x_min=1.0, bin_size=sizes(form - 4), n_bin=12, p=2.0, spacing=sp, bins=b)
    else
        call create_bins(x_min=1.0, x_max=40.0, bin_size=0.1, spacing=log_binning, bins=b)
    end if
    print '(i0)', b%n_bin
    print '(es15.7)', b%x_max
    print '(es15.7)', b%bin_size
    do i = 1, b%n_bin + 1
        print '(es15.7)', b%bin_edges(i)
    end do
    do i = 1, b%n_bin
        print '(es15.7)', b%bin_centers(i)
    end do
end program bins_forms
