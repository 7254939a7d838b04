!   Probe of the d2bi host's input readers for the CPU tests: reads an atom
!   file with read_atoms and prints ok, the atom count, the supercell
!   (columns) and every position wrapped into the voxel supercell the host
!   builds (a_super / (n, n, n)), or parses one string with parse_reals.
program host_io_probe
    use iso_fortran_env, only: real64
    use d2bi_host_io
    implicit none
    character(len=4096) :: mode, arg
    real(real64), allocatable :: x(:, :)
    real(real64) :: a(3, 3), v(9), av(3, 3)
    logical :: ok
    integer :: i, m, n3(3)
    call get_command_argument(1, mode)
    call get_command_argument(2, arg)
    if (trim(mode) == "parse") then
        call parse_reals(trim(arg), v, m)
        write(*, '(i0)') m
        do i = 1, m
            write(*, '(es26.17e3)') v(i)
        end do
    else
        call read_atoms(trim(arg), x, a, ok)
        write(*, '(l1)') ok
        if (.not. ok) stop
        write(*, '(i0)') size(x, 2)
        write(*, '(3es26.17e3)') a
        n3 = 3
        do i = 1, 3
            av(:, i) = a(:, i) / n3(i)
        end do
        do i = 1, size(x, 2)
            write(*, '(6es26.17e3)') x(:, i), wrap_pbc(av, n3, x(:, i))
        end do
    end if
end program host_io_probe
