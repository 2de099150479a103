! ref_harness.f90 -- drives the reference's own binning and kd-tree routines
! (module utils: create_bins / find_bin_index, src/utils.f90:101-230; module
! kdtree_module: grow_tree / search_within_r, src/kdtree.f90:73-375), compiled
! from the unmodified reference sources by oracle/build_ref.sh.
!
! TEST INFRASTRUCTURE ONLY: tests/golden/make_ref_vectors.py runs it in the
! build container to write golden vectors; nothing ships or runs it on a GPU.
!
! Usage: ref_harness bins IN OUT | ref_harness tree IN OUT
!   bins IN : int32 spacing, n_bin, n ; float32 x_min, x_max, p ; float32 x(n)
!        OUT: float32 edges(n_bin+1), centers(n_bin) ; int32 idx(n)
!   tree IN : int32 n, n_q ; float32 r ; float32 px(n), py(n), qx(n_q), qy(n_q)
!        OUT: per query: int32 count, int32 idx(count) (1-based, sorted), then
!             the same from the brute-force search
program ref_harness
    use utils
    use kdtree_module
    implicit none
    character(len=16) :: mode
    character(len=1024) :: fin, fout
    integer :: u, v

    call get_command_argument(1, mode)
    call get_command_argument(2, fin)
    call get_command_argument(3, fout)
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    open(newunit=v, file=trim(fout), access='stream', form='unformatted', status='replace')
    if (trim(mode) == 'bins') then
        call run_bins(u, v)
    else
        call run_tree(u, v)
    end if
    close(u)
    close(v)

contains

    subroutine run_bins(u, v)
        integer, intent(in) :: u, v
        integer :: spacing, n_bin, n, i
        real :: x_min, x_max, p
        real, allocatable :: x(:)
        integer, allocatable :: idx(:)
        type(binning) :: bins
        read(u) spacing, n_bin, n
        read(u) x_min, x_max, p
        allocate(x(n), idx(n))
        read(u) x
        call create_bins(x_min=x_min, x_max=x_max, n_bin=n_bin, p=p, spacing=spacing, bins=bins)
        do i = 1, n
            idx(i) = find_bin_index(x(i), bins)
        end do
        write(v) bins%bin_edges, bins%bin_centers, idx
    end subroutine run_bins

    subroutine run_tree(u, v)
        integer, intent(in) :: u, v
        integer :: n, n_q, i, n_found
        real :: r
        real, allocatable :: px(:), py(:), qx(:), qy(:)
        real, allocatable, target :: pts(:, :)
        integer, allocatable :: found(:)
        type(kdtree) :: tree
        read(u) n, n_q
        read(u) r
        allocate(px(n), py(n), qx(n_q), qy(n_q), pts(n, 2), found(n))
        read(u) px, py, qx, qy
        pts(:, 1) = px
        pts(:, 2) = py
        call grow_tree(tree, pts, leafsize=10)
        do i = 1, n_q
            call search_within_r(tree, [qx(i), qy(i)], r, found, n_found, .false.)
            call put_sorted(v, found, n_found)
            call search_within_r(tree, [qx(i), qy(i)], r, found, n_found, .true.)
            call put_sorted(v, found, n_found)
        end do
        call burn_tree(tree)
    end subroutine run_tree

    subroutine put_sorted(v, found, n_found)
        integer, intent(in) :: v, n_found
        integer, intent(inout) :: found(:)
        integer :: a, b, t
        do a = 2, n_found
            t = found(a)
            b = a - 1
            do while (b >= 1)
                if (found(b) <= t) exit
                found(b + 1) = found(b)
                b = b - 1
            end do
            found(b + 1) = t
        end do
        write(v) n_found
        if (n_found > 0) write(v) found(1:n_found)
    end subroutine put_sorted

end program ref_harness
