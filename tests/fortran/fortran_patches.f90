! fortran_patches.f90 -- a Fortran host program for the bootstrap chain on the
! GPU (module tpcf_gpu): partition_data_gpu (patch labels, the Voronoi form of
! partition_data, src/kmeans.f90:12-263) and block_bootstrap_gpu (the
! resampling of src/tpcf.f90:1165-1191 and covariance_matrix,
! src/utils.f90:745-764).  Reads a binary input file (argument 1) and writes
! the results (argument 2); tests/test_gpu_fortran.py compares them with the
! same calls made through ctypes.
program fortran_patches
    use tpcf_gpu
    implicit none
    integer :: u, v, n, k, n_blk, n_bin, n_boot
    integer(kind=8) :: seed
    real :: supersampling
    real, allocatable :: pts(:, :), w(:), cen(:, :), bt(:, :), bx(:, :), bk(:, :), bw(:)
    real, allocatable :: gt(:, :), gx(:, :), ww(:), ct(:, :), cx(:, :)
    integer, allocatable :: lab(:)
    character(len=512) :: fin, fout

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(newunit=u, file=trim(fin), access="stream", form="unformatted", status="old")
    read(u) n, k, n_blk, n_bin, n_boot
    read(u) seed
    read(u) supersampling
    allocate(pts(n, 2), w(n), cen(k, 2), lab(n))
    read(u) pts(:, 1), pts(:, 2), w
    allocate(bt(n_bin, n_blk), bx(n_bin, n_blk), bk(n_bin, n_blk), bw(n_blk))
    read(u) bt, bx, bk, bw
    close(u)

    cen = 0.0
    call partition_data_gpu(pts, w, k, lab, cen)
    allocate(gt(n_boot, n_bin), gx(n_boot, n_bin), ww(n_boot), ct(n_bin, n_bin), cx(n_bin, n_bin))
    call block_bootstrap_gpu(bt, bx, bk, bw, n_boot, supersampling, seed, gt, gx, ww, ct, cx)

    open(newunit=v, file=trim(fout), access="stream", form="unformatted", status="replace")
    write(v) lab, cen, gt, gx, ww, ct, cx
    close(v)
    call tpcf_gpu_finalize()
end program fortran_patches
