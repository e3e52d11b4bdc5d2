! Reads a restart record file with the statements of the reference's restart read
! (record layout of Input/Init.F90:37-46) and prints KI and sums, for the record-format test.
program read_restart
    implicit none
    integer :: ni1, nj1, nk1, kind_dual, ki, i, j, k, ll
    integer, parameter :: ml = 7
    real(8), allocatable :: v(:, :, :, :), rmiu(:, :, :), wt0(:, :, :, :), wt1(:, :, :, :)
    character(len=512) :: path, arg
    call get_command_argument(1, path)
    call get_command_argument(2, arg); read (arg, *) ni1
    call get_command_argument(3, arg); read (arg, *) nj1
    call get_command_argument(4, arg); read (arg, *) nk1
    call get_command_argument(5, arg); read (arg, *) kind_dual
    allocate (v(ml, ni1, nj1, nk1), rmiu(ni1, nj1, nk1), wt0(ml, ni1, nj1, nk1), wt1(ml, ni1, nj1, nk1))
    wt0 = 0; wt1 = 0
    open (unit=15, file=trim(path), form="unformatted", access="sequential", status="old")
    read (15) ki
    read (15) ((((v(ll, i, j, k), k=1, nk1), j=1, nj1), i=1, ni1), ll=1, ml)
    read (15) (((rmiu(i, j, k), k=1, nk1), j=1, nj1), i=1, ni1)
    if (kind_dual == 1) then
        read (15) ((((wt0(ll, i, j, k), k=1, nk1), j=1, nj1), i=1, ni1), ll=1, ml)
    else if (kind_dual == 2) then
        read (15) ((((wt0(ll, i, j, k), k=1, nk1), j=1, nj1), i=1, ni1), ll=1, ml)
        read (15) ((((wt1(ll, i, j, k), k=1, nk1), j=1, nj1), i=1, ni1), ll=1, ml)
    end if
    close (15)
    print '(i0)', ki
    ! one value per component and array at a fixed cell, and weighted sums
    print '(es25.17)', v(3, 2, 1, 1), v(7, ni1, nj1, nk1), rmiu(1, nj1, 2), wt1(5, 1, 2, nk1)
    print '(es25.17)', sum(v), sum(rmiu), sum(wt0), sum(wt1)
end program read_restart
