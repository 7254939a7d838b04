!   d2bi_host_io.f90 -- command line, atom files and extinction tables for
!   the d2bi host.  Written for this library from the behaviour of the
!   reference driver (d2bi.f90), Lib_CommandLineArguments (-key v1,v2 and
!   -nokey), Lib_XYZFiles (LAMMPS write_data and extended .xyz input) and
!   Lib_ReadExtinctionDistances.
module d2bi_host_io
    use iso_fortran_env, only: real64, int64, error_unit
    implicit none
    private

    public :: cla_init, cla_has, cla_string, cla_reals, cla_ints, cla_logical, cla_unused
    public :: read_atoms, read_extinction_distance, cubic_symmetry

    integer, parameter :: MAXARG = 256
    character(len=1024) :: keys(MAXARG), vals(MAXARG)
    logical             :: used(MAXARG)
    integer             :: nargs = 0

contains

    !   split "-key value" / "-key" / "-nokey" arguments (Lib_CommandLineArguments)
    subroutine cla_init()
        integer :: i, n
        character(len=1024) :: a, b
        n = command_argument_count()
        nargs = 0
        i = 1
        do while (i <= n)
            call get_command_argument(i, a)
            if (a(1:1) /= "-" .or. len_trim(a) < 2) then
                write(error_unit, '(a)') "d2bi error - unexpected argument """//trim(a)//""""
                stop 2
            end if
            nargs = nargs + 1
            keys(nargs) = a(2:)
            vals(nargs) = ""
            used(nargs) = .false.
            if (i < n) then
                call get_command_argument(i + 1, b)
                if (.not. is_key(b)) then
                    vals(nargs) = b
                    i = i + 1
                end if
            end if
            i = i + 1
        end do
    end subroutine cla_init

    logical function is_key(s)
        character(len=*), intent(in) :: s
        real(real64) :: v
        integer :: ios
        is_key = .false.
        if (len_trim(s) < 2) return
        if (s(1:1) /= "-") return
        read(s, *, iostat=ios) v           ! "-0.01" is a value, "-sg" a key
        is_key = (ios /= 0)
    end function is_key

    integer function find(key)
        character(len=*), intent(in) :: key
        integer :: i
        find = 0
        do i = nargs, 1, -1
            if (trim(keys(i)) == key .or. trim(keys(i)) == "no"//key) then
                find = i
                return
            end if
        end do
    end function find

    logical function cla_has(key)
        character(len=*), intent(in) :: key
        cla_has = find(key) > 0
    end function cla_has

    subroutine cla_string(key, s)
        character(len=*), intent(in)    :: key
        character(len=*), intent(inout) :: s
        integer :: i
        i = find(key)
        if (i == 0) return
        used(i) = .true.
        s = vals(i)
        if (len_trim(s) >= 2) then          ! -o "" arrives as an empty value
            if (s(1:1) == '"' .and. s(len_trim(s):len_trim(s)) == '"') s = s(2:len_trim(s) - 1)
        end if
    end subroutine cla_string

    !   comma separated reals; n returns the count read (0 if the key is absent)
    subroutine cla_reals(key, x, n)
        character(len=*), intent(in)      :: key
        real(real64), intent(inout)       :: x(:)
        integer, intent(out)              :: n
        integer :: i, p0, p1, ios
        character(len=1024) :: s
        n = 0
        i = find(key)
        if (i == 0) return
        used(i) = .true.
        s = vals(i)
        p0 = 1
        do while (p0 <= len_trim(s) .and. n < size(x))
            p1 = index(s(p0:), ",")
            if (p1 == 0) then
                p1 = len_trim(s) + 1
            else
                p1 = p0 + p1 - 1
            end if
            read(s(p0:p1 - 1), *, iostat=ios) x(n + 1)
            if (ios /= 0) then
                write(error_unit, '(a)') "d2bi error - can't parse -"//trim(key)//" "//trim(s)
                stop 2
            end if
            n = n + 1
            p0 = p1 + 1
        end do
    end subroutine cla_reals

    subroutine cla_ints(key, k, n)
        character(len=*), intent(in) :: key
        integer, intent(inout)       :: k(:)
        integer, intent(out)         :: n
        real(real64) :: x(size(k))
        call cla_reals(key, x, n)
        if (n > 0) k(1:n) = nint(x(1:n))
    end subroutine cla_ints

    !   -key (true) / -nokey (false) / -key T,F,...
    subroutine cla_logical(key, b, n)
        character(len=*), intent(in) :: key
        logical, intent(inout)       :: b(:)
        integer, intent(out)         :: n
        integer :: i, p0, p1
        character(len=1024) :: s
        n = 0
        i = find(key)
        if (i == 0) return
        used(i) = .true.
        if (trim(keys(i)) == "no"//key) then
            b(1) = .false.
            n = 1
            return
        end if
        s = vals(i)
        if (len_trim(s) == 0) then
            b(1) = .true.
            n = 1
            return
        end if
        p0 = 1
        do while (p0 <= len_trim(s) .and. n < size(b))
            p1 = index(s(p0:), ",")
            if (p1 == 0) then
                p1 = len_trim(s) + 1
            else
                p1 = p0 + p1 - 1
            end if
            n = n + 1
            b(n) = scan(s(p0:p1 - 1), "TtYy1") > 0
            p0 = p1 + 1
        end do
    end subroutine cla_logical

    subroutine cla_unused()
        integer :: i
        do i = 1, nargs
            if (.not. used(i)) write(*, '(a)') "d2bi warning - unused option -"//trim(keys(i))
        end do
    end subroutine cla_unused

    !   Atom positions and supercell.  LAMMPS write_data (Lib_XYZFiles.f90:628,
    !   1061): header "N atoms", "xlo xhi" lines, an "Atoms" section of
    !   "id type x y z ..." in file order; positions are not shifted by xlo.
    !   Extended .xyz: line 1 N, line 2 with Lattice="ax ay az bx by bz cx cy cz",
    !   then "element x y z ..." lines.
    subroutine read_atoms(filename, x, a_super, ok)
        character(len=*), intent(in)              :: filename
        real(real64), allocatable, intent(out)    :: x(:, :)
        real(real64), intent(out)                 :: a_super(3, 3)
        logical, intent(out)                      :: ok
        integer :: u, ios, n, i, j, k, id, typ
        character(len=4096) :: line
        real(real64) :: lo(3), hi(3), v(9)
        character(len=64) :: el
        ok = .false.
        a_super = 0
        open(newunit=u, file=trim(filename), action="read", status="old", iostat=ios)
        if (ios /= 0) return
        read(u, '(a)', iostat=ios) line
        if (ios /= 0) then
            close(u)
            return
        end if
        read(line, *, iostat=ios) n
        if (ios == 0 .and. len_trim(adjustl(line)) > 0 .and. verify(trim(adjustl(line)), "0123456789 ") == 0) then
            ! extended xyz
            read(u, '(a)', iostat=ios) line
            i = index(line, 'Lattice="')
            if (i == 0) i = index(line, "Lattice='")
            if (i == 0) then
                close(u)
                return
            end if
            j = i + 9 + index(line(i + 9:), line(i + 8:i + 8)) - 2
            read(line(i + 9:j), *, iostat=ios) v
            if (ios /= 0) then
                close(u)
                return
            end if
            a_super = reshape(v, [3, 3])
            allocate(x(3, n))
            do k = 1, n
                read(u, *, iostat=ios) el, x(1:3, k)
                if (ios /= 0) then
                    close(u)
                    return
                end if
            end do
            close(u)
            ok = .true.
            return
        end if
        ! LAMMPS data file
        n = -1
        lo = 0
        hi = 0
        do
            read(u, '(a)', iostat=ios) line
            if (ios /= 0) exit
            if (index(line, " atoms") > 0 .and. n < 0) read(line, *) n
            if (index(line, "xlo") > 0) read(line, *) lo(1), hi(1)
            if (index(line, "ylo") > 0) read(line, *) lo(2), hi(2)
            if (index(line, "zlo") > 0) read(line, *) lo(3), hi(3)
            if (line(1:5) == "Atoms") exit
        end do
        if (n < 0 .or. ios /= 0) then
            close(u)
            return
        end if
        read(u, '(a)', iostat=ios) line             ! blank line after "Atoms"
        allocate(x(3, n))
        do k = 1, n
            read(u, *, iostat=ios) id, typ, x(1:3, k)
            if (ios /= 0) then
                close(u)
                return
            end if
        end do
        close(u)
        do i = 1, 3
            a_super(i, i) = hi(i) - lo(i)
        end do
        ok = .true.
    end subroutine read_atoms

    !   cubic symmetry-standard reflection: |h| >= |k| >= |l|
    !   (Lib_ReadExtinctionDistances.f90:111-150)
    function cubic_symmetry(hkl) result(s)
        integer, intent(in) :: hkl(3)
        integer :: s(3), t(3), i, j, m
        t = abs(hkl)
        do i = 1, 2
            do j = i + 1, 3
                if (t(j) > t(i)) then
                    m = t(i); t(i) = t(j); t(j) = m
                end if
            end do
        end do
        s = t
    end function cubic_symmetry

    !   xi_0 and xi_g for g = [hkl] from an extinction distance table: two
    !   comment lines then "element h k l T V Re(U_g) Im(U_g) xi_g"
    !   (Lib_ReadExtinctionDistances.f90:29-109)
    subroutine read_extinction_distance(filename, hkl, xi0, xig, ok)
        character(len=*), intent(in) :: filename
        integer, intent(in)          :: hkl(3)
        real(real64), intent(out)    :: xi0, xig
        logical, intent(out)         :: ok
        integer :: u, ios, i, idat(3), s(3)
        real(real64) :: rdat(5)
        character(len=256) :: el, line
        logical :: f0, fg
        xi0 = huge(1.0)
        xig = huge(1.0)
        f0 = .false.
        fg = .false.
        ok = .false.
        open(newunit=u, file=trim(filename), action="read", status="old", iostat=ios)
        if (ios /= 0) return
        read(u, '(a)', iostat=ios) line
        read(u, '(a)', iostat=ios) line
        s = cubic_symmetry(hkl)
        do i = 1, 100000
            read(u, *, iostat=ios) el, idat, rdat
            if (ios /= 0) exit
            idat = cubic_symmetry(idat)
            if (all(idat == 0)) then
                xi0 = rdat(5)
                f0 = .true.
            end if
            if (all(idat == s)) then
                xig = rdat(5)
                fg = .true.
            end if
            if (f0 .and. fg) exit
        end do
        close(u)
        ok = f0 .and. fg
    end subroutine read_extinction_distance

end module d2bi_host_io
