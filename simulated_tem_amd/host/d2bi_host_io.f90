!   d2bi_host_io.f90 -- command line, atom files and extinction tables for
!   the d2bi host.  Written for this library from the behaviour of the
!   reference driver (d2bi.f90), Lib_CommandLineArguments (-key v1,v2 and
!   -nokey), Lib_XYZFiles (LAMMPS write_data, extended / PARCAS .xyz and
!   commented-supercell .xyz input) and Lib_ReadExtinctionDistances.
module d2bi_host_io
    use iso_fortran_env, only: real64, int64, error_unit
    implicit none
    private

    public :: cla_init, cla_has, cla_string, cla_reals, cla_ints, cla_logical, cla_unused
    public :: read_atoms, read_extinction_distance, cubic_symmetry, parse_reals, wrap_pbc

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

    !   Atom positions and supercell a_super (columns = cell vectors), as
    !   XYZFile's readHeader / input / getSupercell read them
    !   (Lib_XYZFiles.f90:520-1151).  A first line beginning "LAMMPS" marks a
    !   LAMMPS write_data file: "N atoms", "xlo xhi" .. "zlo zhi" and the
    !   triclinic "xy xz yz" line in the header, then an "Atoms" section of
    !   "id type x y z ..." lines in file order, positions not shifted by the
    !   box origin (:628-740, 1061-1151).  Otherwise an .xyz file: comment
    !   lines (# % !), the atom count, a description line, then "species x y z
    !   ..." lines (:548-620, 756-880); its supercell (:928-1054) is
    !       Lattice="..."   3 values (cuboid) or 9 (columns), extended xyz
    !       boxsize ...     the first 3 values (cuboid), PARCAS
    !   or, in the comment lines, "supercell nx ny nz" followed by three
    !   lines of cell vectors (rows) or one of 9 values (columns), each
    !   scaled by its repeat count.
    subroutine read_atoms(filename, x, a_super, ok)
        character(len=*), intent(in)              :: filename
        real(real64), allocatable, intent(out)    :: x(:, :)
        real(real64), intent(out)                 :: a_super(3, 3)
        logical, intent(out)                      :: ok
        integer :: u, ios
        character(len=4096) :: line
        ok = .false.
        a_super = reshape([1.0_real64, 0.0_real64, 0.0_real64, 0.0_real64, 1.0_real64, 0.0_real64, &
                           0.0_real64, 0.0_real64, 1.0_real64], [3, 3])
        open(newunit=u, file=trim(filename), action="read", status="old", iostat=ios)
        if (ios /= 0) return
        read(u, '(a)', iostat=ios) line
        if (ios /= 0) then
            close(u)
            return
        end if
        rewind(u)
        line = adjustl(line)
        if (line(1:6) == "LAMMPS") then
            call read_lammps(u, x, a_super, ok)
        else
            call read_xyz(u, x, a_super, ok)
        end if
        close(u)
    end subroutine read_atoms

    subroutine read_lammps(u, x, a_super, ok)
        integer, intent(in)                       :: u
        real(real64), allocatable, intent(out)    :: x(:, :)
        real(real64), intent(inout)               :: a_super(3, 3)
        logical, intent(out)                      :: ok
        integer :: ios, n, k, id, typ, i
        character(len=4096) :: line
        real(real64) :: lo(3), hi(3), off(3)
        ok = .false.
        n = -1
        lo = 0
        hi = 0
        off = 0
        do i = 1, 100000
            read(u, '(a)', iostat=ios) line
            if (ios /= 0) return
            line = adjustl(line)
            if (index(line, "atoms") > 0) then
                read(line, *, iostat=ios) n
                if (ios /= 0) return
            end if
            if (index(line, "xlo") > 0) read(line, *, iostat=ios) lo(1), hi(1)
            if (index(line, "ylo") > 0) read(line, *, iostat=ios) lo(2), hi(2)
            if (index(line, "zlo") > 0) read(line, *, iostat=ios) lo(3), hi(3)
            if (index(line, "xy xz yz") > 0) read(line, *, iostat=ios) off
            if (ios /= 0) return
            if (line(1:5) == "Atoms") exit
        end do
        if (n < 0) return
        read(u, '(a)', iostat=ios) line             ! blank line after "Atoms"
        allocate(x(3, n))
        do k = 1, n
            read(u, *, iostat=ios) id, typ, x(1:3, k)
            if (ios /= 0) return
        end do
        if (.not. (norm2(hi) > norm2(lo))) return
        hi = hi - lo
        a_super = 0
        do i = 1, 3
            a_super(i, i) = hi(i)
        end do
        a_super(1, 2) = off(1)
        a_super(1, 3) = off(2)
        a_super(2, 3) = off(3)
        ok = .true.
    end subroutine read_lammps

    subroutine read_xyz(u, x, a_super, ok)
        integer, intent(in)                       :: u
        real(real64), allocatable, intent(out)    :: x(:, :)
        real(real64), intent(inout)               :: a_super(3, 3)
        logical, intent(out)                      :: ok
        integer :: ios, n, k, nh, i, j, m
        character(len=4096) :: line, desc
        character(len=64) :: el
        real(real64) :: v(9), rep(3)
        ok = .false.
        nh = 0
        do
            read(u, '(a)', iostat=ios) line
            if (ios /= 0) return
            line = adjustl(line)
            if (scan(line(1:1), "#%!") == 0) exit
            nh = nh + 1
        end do
        read(line, *, iostat=ios) n
        if (ios /= 0 .or. n < 0) return
        read(u, '(a)', iostat=ios) desc
        if (ios /= 0) return
        allocate(x(3, n))
        do k = 1, n
            read(u, *, iostat=ios) el, x(1:3, k)
            if (ios /= 0) return
        end do
        i = index(trim(desc), 'Lattice="')
        if (i /= 0) then
            call parse_reals(desc(i + 9:), v, m)
            if (m == 3) then
                a_super = 0
                a_super(1, 1) = v(1); a_super(2, 2) = v(2); a_super(3, 3) = v(3)
                ok = .true.
            else if (m == 9) then
                a_super = reshape(v, [3, 3])
                ok = .true.
            end if
            return
        end if
        i = index(trim(desc), "boxsize")
        if (i /= 0) then
            call parse_reals(desc(i + 8:), v, m)
            if (m >= 3) then
                a_super = 0
                a_super(1, 1) = v(1); a_super(2, 2) = v(2); a_super(3, 3) = v(3)
                ok = .true.
            end if
            return
        end if
        ! "supercell nx ny nz" in a comment line, then the cell vectors
        rewind(u)
        do k = 1, nh
            read(u, '(a)', iostat=ios) line
            if (ios /= 0) return
            j = max(index(line, "supercell"), index(line, "super cell"))
            if (j == 0) cycle
            call parse_reals(line(j + 9:), v, m)
            rep = 1
            if (m == 3) rep = v(1:3)
            do i = 1, 3
                read(u, '(a)', iostat=ios) line
                if (ios /= 0) return
                call parse_reals(line(3:), v, m)
                if (m == 3) then
                    a_super(i, 1:3) = v(1:3)
                else if (m == 9) then
                    a_super = reshape(v, [3, 3])
                    exit
                else
                    return
                end if
            end do
            do i = 1, 3
                a_super(:, i) = a_super(:, i) * nint(rep(i))
            end do
            ok = .true.
            return
        end do
    end subroutine read_xyz

    !   Up to size(v) reals from the tokens of s (separators: space , ; tab CR
    !   NUL), counting the tokens that begin with a number character
    !   0-9 . + - e E d D (parse into reals, NBAX_StringTokenizers.f90:273-425):
    !   the token's leading run of such characters is split as
    !   aaaa[.bbbb][e cccc] and read as (aaaa + 0.b + 0.0b + ...) * 10**cccc,
    !   sign from the first character of aaaa and of cccc.
    subroutine parse_reals(s, v, m)
        character(len=*), intent(in)  :: s
        real(real64), intent(inout)   :: v(:)
        integer, intent(out)          :: m
        integer :: p, q, L
        character(len=32) :: tok
        m = 0
        L = len_trim(s)
        p = 1
        do while (p <= L .and. m < size(v))
            do while (p <= L)
                if (scan(s(p:p), " ,;"//achar(9)//achar(13)//achar(0)) == 0) exit
                p = p + 1
            end do
            if (p > L) exit
            q = p
            do while (q <= L)
                if (scan(s(q:q), " ,;"//achar(9)//achar(13)//achar(0)) /= 0) exit
                q = q + 1
            end do
            tok = s(p:q - 1)
            p = q
            if (token_real(tok, v(m + 1))) m = m + 1
        end do
    end subroutine parse_reals

    logical function token_real(tok, r)
        character(len=*), intent(in)  :: tok
        real(real64), intent(inout)   :: r
        character(len=64) :: str
        character(len=32) :: aa, bb, cc
        integer :: i, j, n, d, ic
        real(real64) :: xa, xb, xx
        token_real = .false.
        str = ""
        n = 0
        do i = 1, len(tok)
            if (n == 0 .and. tok(i:i) == " ") cycle
            if (index("0123456789.+-eEdD", tok(i:i)) == 0) exit
            n = n + 1
            str(n:n) = tok(i:i)
        end do
        if (n == 0) return
        do i = 1, n
            if (index("dDE", str(i:i)) > 0) str(i:i) = "e"
        end do
        aa = "0"
        bb = "0"
        cc = "0"
        i = index(str, ".")
        j = index(str, "e")
        if (i == 1) then
            if (j > 1) then
                bb = str(2:j - 1)
                cc = str(j + 1:len_trim(str))
            else
                bb = str(2:len_trim(str))
            end if
        else if (i > 1) then
            aa = str(1:i - 1)
            if (j > 1) then
                bb = str(i + 1:j - 1)
                cc = str(j + 1:len_trim(str))
            else
                bb = str(i + 1:len_trim(str))
            end if
        else
            if (j > 1) then
                aa = str(1:j - 1)
                cc = str(j + 1:len_trim(str))
            else
                aa = str(1:len_trim(str))
            end if
        end if
        xa = 0
        do i = 1, len_trim(aa)
            d = iachar(aa(i:i)) - iachar("0")
            if (d >= 0 .and. d <= 9) xa = xa * 10 + d
        end do
        ic = 0
        do i = 1, len_trim(cc)
            d = iachar(cc(i:i)) - iachar("0")
            if (d >= 0 .and. d <= 9) ic = ic * 10 + d
        end do
        if (cc(1:1) == "-") ic = -ic
        xx = 1
        xb = 0
        do i = 1, len_trim(bb)
            xx = xx * 0.1_real64
            xb = xb + xx * (iachar(bb(i:i)) - iachar("0"))
        end do
        r = (xa + xb) * (10.0_real64**ic)
        if (aa(1:1) == "-") r = -r
        if (abs(r) == 0) r = 0
        token_real = .true.
    end function token_real

    !   wrapPBC0 (Lib_SimpleSupercells.f90:992-1021): cell coordinates by the
    !   inverse of a (inverse3Mat, :1205-1236: cofactors times 1/det, det the
    !   mean of the three row expansions, :1238-1255), each wrapped into
    !   [0, n), and back to real space
    function wrap_pbc(a, n3, y) result(yp)
        real(real64), intent(in) :: a(3, 3), y(3)
        integer, intent(in)      :: n3(3)
        real(real64) :: yp(3), xx(3), ia(3, 3), dd(9), idd
        integer :: i, k
        dd(1) = a(1, 1) * (a(2, 2) * a(3, 3) - a(2, 3) * a(3, 2))
        dd(2) = a(1, 2) * (a(2, 3) * a(3, 1) - a(2, 1) * a(3, 3))
        dd(3) = a(1, 3) * (a(2, 1) * a(3, 2) - a(2, 2) * a(3, 1))
        dd(4) = a(2, 1) * (a(3, 2) * a(1, 3) - a(3, 3) * a(1, 2))
        dd(5) = a(2, 2) * (a(3, 3) * a(1, 1) - a(3, 1) * a(1, 3))
        dd(6) = a(2, 3) * (a(3, 1) * a(1, 2) - a(3, 2) * a(1, 1))
        dd(7) = a(3, 1) * (a(1, 2) * a(2, 3) - a(1, 3) * a(2, 2))
        dd(8) = a(3, 2) * (a(1, 3) * a(2, 1) - a(1, 1) * a(2, 3))
        dd(9) = a(3, 3) * (a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1))
        idd = (1.0d0 / 3.0d0) * sum(dd)
        if (abs(idd) < tiny(1.0d0)) then
            ia = 0
            do i = 1, 3
                ia(i, i) = 1
            end do
        else
            idd = 1.0d0 / idd
            ia(1, 1) = (a(2, 2) * a(3, 3) - a(2, 3) * a(3, 2)) * idd
            ia(2, 1) = (a(2, 3) * a(3, 1) - a(2, 1) * a(3, 3)) * idd
            ia(3, 1) = (a(2, 1) * a(3, 2) - a(2, 2) * a(3, 1)) * idd
            ia(1, 2) = (a(1, 3) * a(3, 2) - a(1, 2) * a(3, 3)) * idd
            ia(2, 2) = (a(1, 1) * a(3, 3) - a(1, 3) * a(3, 1)) * idd
            ia(3, 2) = (a(1, 2) * a(3, 1) - a(1, 1) * a(3, 2)) * idd
            ia(1, 3) = (a(1, 2) * a(2, 3) - a(1, 3) * a(2, 2)) * idd
            ia(2, 3) = (a(1, 3) * a(2, 1) - a(1, 1) * a(2, 3)) * idd
            ia(3, 3) = (a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1)) * idd
        end if
        xx = ia(:, 1) * y(1) + ia(:, 2) * y(2) + ia(:, 3) * y(3)
        do i = 1, 3
            k = floor(xx(i))
            xx(i) = xx(i) - k
            k = mod(k + ishft(n3(i), 8), n3(i))
            xx(i) = k + xx(i)
        end do
        yp = a(:, 1) * xx(1) + a(:, 2) * xx(2) + a(:, 3) * xx(3)
    end function wrap_pbc

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
