!   d2bi_main.f90 -- the d2bi command-line host on the MI355X imaging library.
!
!   Written for this library from the behaviour of the reference driver
!   (src/d2bi.f90): same options, input files and output images; the imaging
!   of each condition (computeD2BI_img, :1394) and the PNG blur
!   (gaussianBlurImage, :1543) run on the GPU through the C-ABI.
!
!   Supported: -f -a0 -lattice (cubic) -o -g -k -sg -dark -V -nPrecAngle
!   -precAngle -nTomoAngle -tomoAngle -density -M -xpad -ypad -zpad -d
!   -addOffset -U -R -xifile -xi0 -xig -ng -png -png_min -png_max -png_blur,
!   plus -mode fp64|fast and -device n.  -alterg/-dp, -surf and hcp are
!   not part of this build and stop with a message.  A tomography series
!   (-nTomoAngle > 1) is imaged as one batch (d2bi_gpu_image_batch: one
!   captured graph for every tilt); -nobatch images the tilts one by one.
program d2bi_host
    use iso_fortran_env, only: real64, int64, output_unit
    use iso_c_binding
    use d2bi_gpu_mod
    use d2bi_host_io
    implicit none

    character(len=12), parameter :: VERSION = "1.1.6-gpu"
    real(real64), parameter     :: PI = 3.141592653589790d0
    real(real64), parameter     :: THETA_MAX = 5.0d0          ! Lib_DiffractionConditions.f90:26
    real(real64), parameter     :: UNSET = transfer(int(z'0BADF00D0BADF00D', int64), 1.0d0)
    integer, parameter          :: NDIFFMAX = 100

    character(len=1024) :: filename = "", outfile0 = "", xifile = "", latname = "bcc", modename = "fast"
    real(real64) :: a0in(3), a0, xi0 = 103.889565d0, xig = 207.533203d0, V = 200.0d0
    real(real64) :: precAngle = 0.005d0, tomoAngle = 20.0d0, png_min = 0.0d0, png_max = 0.0d0
    real(real64) :: png_blur = 2.50d0
    real(real64) :: darr(4 * NDIFFMAX), g3a(3, NDIFFMAX), k3a(3, NDIFFMAX), sga(NDIFFMAX), nga(NDIFFMAX), ng
    logical      :: darka(NDIFFMAX), bl(NDIFFMAX)
    real(real64) :: d0(3), addOffset(3), U(3, 3), R(3, 3), dat(12)
    integer      :: M(2) = [4, 4], nPrec = 1, nTomo = 1, device = 0, mode
    logical      :: useDensity = .false., xpad = .false., ypad = .false., zpad = .false.
    logical      :: opPng = .true., haveU = .false., haveR = .false., useBatch = .true.
    integer      :: nDiff = 1, n, ii, nn, kk, rc, Mx, My, Mz, latcode
    real(real64), allocatable :: x(:, :), img(:, :), blurred(:, :)
    real(c_float), allocatable :: pngbuf(:)
    real(real64) :: a_super(3, 3), a_vox(3, 3), yy(3), xmin(3), xmax(3), d(3), rot(3, 3)
    real(real64) :: theta, pad, len1, len2, modg, sg, t_d2bi, t0, t1, U0(3, 3), R0(3, 3), Rt(3, 3)
    integer      :: sN(3)
    type(d2bi_record) :: rec, rec0
    type(d2bi_plan)   :: plan
    type(d2bi_stats)  :: st
    type(c_ptr)       :: atoms
    type(d2bi_record), allocatable :: recs(:)
    type(c_ptr), allocatable       :: hs(:)
    real(real64), allocatable      :: pas(:), imgs(:)
    integer(c_int32_t), allocatable :: nps(:)
    integer(int64), allocatable    :: offs(:)
    real(real64) :: thetas(1000)
    logical      :: ok, dark
    character(len=1024) :: outfile, fname
    character(len=64)   :: s1, s2, s3

    call cla_init()
    if (.not. cla_has("f") .or. .not. cla_has("a0") .or. .not. cla_has("g") .or. .not. cla_has("k")) then
        write(*, '(a)') "d2bi v"//trim(VERSION)//" - simulated TEM in the dynamical two-beam approximation (MI355X)"
        write(*, '(a)') "usage: d2bi -f file -a0 a [-lattice bcc] -g h,k,l[,...] -k u,v,w [-sg s | -ng n] [-o prefix]"
        write(*, '(a)') "       [-dark|-nodark] [-V keV] [-M xy[,z]] [-d x,y,z] [-U 9 reals] [-R 9 reals]"
        write(*, '(a)') "       [-xifile f | -xi0 a -xig b] [-density] [-nPrecAngle n -precAngle rad]"
        write(*, '(a)') "       [-nTomoAngle n -tomoAngle deg [-nobatch]] [-xpad|-ypad|-zpad] [-png_blur A] [-mode fp64|fast]"
        stop 1
    end if
    call cla_string("f", filename)
    a0in = UNSET
    call cla_reals("a0", a0in, n)
    if (n == 1) a0in(2:3) = a0in(1)
    call cla_string("lattice", latname)
    if (trim(latname) == "hcp") error stop "d2bi error - hcp lattices are not part of this build"
    select case (trim(latname))
        case ("bcc"); latcode = 2
        case ("fcc"); latcode = 3
        case default; latcode = 1
    end select
    outfile0 = filename
    call cla_string("o", outfile0)
    bl(1) = opPng
    call cla_logical("png", bl, n)
    if (n > 0) opPng = bl(1)
    call cla_reals("png_min", dat, n); if (n > 0) png_min = dat(1)
    call cla_reals("png_max", dat, n); if (n > 0) png_max = dat(1)
    call cla_reals("png_blur", dat, n); if (n > 0) png_blur = dat(1)
    if (cla_has("png_min") .or. cla_has("png_max") .or. cla_has("png_blur")) opPng = .true.
    call cla_string("xifile", xifile)
    call cla_reals("xi0", dat, n); if (n > 0) xi0 = dat(1)
    call cla_reals("xig", dat, n); if (n > 0) xig = dat(1)
    if (len_trim(xifile) > 0 .and. (cla_has("xi0") .or. cla_has("xig"))) &
        error stop "d2bi error - attempt to define xi_0 or xi_g and read from file"
    if (cla_has("alterg") .or. cla_has("dp")) error stop "d2bi error - -alterg/-dp are not part of this build"
    ! g-vectors and zone axes (d2bi.f90:244-325)
    darr = 0
    call cla_reals("g", darr, n)
    if (mod(n, 3) /= 0) error stop "d2bi error - can't parse array of miller index g-vectors."
    nDiff = n / 3
    do ii = 1, nDiff
        g3a(:, ii) = darr(3 * ii - 2:3 * ii)
    end do
    call cla_reals("k", darr, n)
    if (mod(n, 3) /= 0) error stop "d2bi error - can't parse array of miller index k-vectors."
    if (n / 3 == nDiff) then
        do ii = 1, nDiff
            k3a(:, ii) = darr(3 * ii - 2:3 * ii)
        end do
    else if (n / 3 == 1) then
        do ii = 1, nDiff
            k3a(:, ii) = darr(1:3)
        end do
    else if (nDiff == 1) then
        nDiff = n / 3
        do ii = 1, nDiff
            k3a(:, ii) = darr(3 * ii - 2:3 * ii)
            g3a(:, ii) = g3a(:, 1)
        end do
    else
        error stop "d2bi error - array of miller index k-vectors must match g-vectors."
    end if
    sga = UNSET
    call cla_reals("sg", sga, n)
    if (n == 1) then
        sga(1:nDiff) = sga(1)
    else if (nDiff == 1 .and. n > 1) then
        nDiff = n
        do ii = 2, nDiff
            k3a(:, ii) = k3a(:, 1)
            g3a(:, ii) = g3a(:, 1)
        end do
    else if (n * (n - nDiff) /= 0) then
        error stop "d2bi error - array of -sg deviation parameters must match k,g-vectors "
    end if
    nga = UNSET
    call cla_reals("ng", nga, n)
    if (n == 1) then
        nga(1:nDiff) = nga(1)
    else if (nDiff == 1 .and. n > 1) then
        nDiff = n
        do ii = 2, nDiff
            k3a(:, ii) = k3a(:, 1)
            g3a(:, ii) = g3a(:, 1)
        end do
        sga(1:nDiff) = sga(1)
    else if (n * (n - nDiff) /= 0) then
        error stop "d2bi error - array of -ng parameters must match k,g-vectors "
    end if
    if (cla_has("ng") .and. cla_has("sg")) error stop "d2bi error - attempt to define both n_g and s_g"
    if (cla_has("ng") .and. len_trim(xifile) == 0) error stop "d2bi error - the (g,ng) tilt search needs -xifile"
    darka = .true.
    call cla_logical("dark", darka, n)
    if (n == 1) darka(1:nDiff) = darka(1)
    call cla_reals("V", dat, n); if (n > 0) V = dat(1)
    call cla_ints("nPrecAngle", sN(1:1), n); if (n > 0) nPrec = sN(1)
    call cla_reals("precAngle", dat, n); if (n > 0) precAngle = dat(1)
    call cla_ints("nTomoAngle", sN(1:1), n); if (n > 0) nTomo = sN(1)
    if (cla_has("nobatch")) useBatch = .false.
    if (nTomo > size(thetas)) error stop "d2bi error - at most 1000 tomography steps"
    call cla_reals("tomoAngle", dat, n); if (n > 0) tomoAngle = dat(1)
    bl(1) = .false.; call cla_logical("density", bl, n); if (n > 0) useDensity = bl(1)
    call cla_ints("M", M, n)
    if (n == 1) M(2) = M(1)
    bl(1) = .false.; call cla_logical("xpad", bl, n); if (n > 0) xpad = bl(1)
    bl(1) = .false.; call cla_logical("ypad", bl, n); if (n > 0) ypad = bl(1)
    bl(1) = .false.; call cla_logical("zpad", bl, n); if (n > 0) zpad = bl(1)
    d0 = UNSET
    call cla_reals("d", dat, n)
    if (n == 1) then
        d0(1:2) = 0; d0(3) = dat(1)
    else if (n == 2) then
        d0(1:2) = dat(1:2); d0(3) = 0
    else if (n == 3) then
        d0 = dat(1:3)
    end if
    addOffset = UNSET
    call cla_reals("addOffset", addOffset, n)
    call cla_reals("U", dat, n)
    if (n == 9) then
        U = reshape(dat(1:9), [3, 3])
        haveU = .true.
        if (abs(det3(U) - 1) > 0.001d0) error stop "d2bi error - input U matrix has a determinant /= 1"
    end if
    call cla_reals("R", dat, n)
    if (n == 9) then
        R = reshape(dat(1:9), [3, 3])
        haveR = .true.
        if (abs(det3(R) - 1) > 0.001d0) error stop "d2bi error - input R matrix has a determinant  /= 1"
    end if
    call cla_string("mode", modename)
    mode = D2BI_MODE_FAST
    if (trim(modename) == "fp64") mode = D2BI_MODE_FP64
    call cla_ints("device", sN(1:1), n); if (n > 0) device = sN(1)
    bl(1) = .false.; call cla_logical("surf", bl, n)
    call cla_logical("opxyz", bl, n)
    call cla_unused()

    inquire(file=trim(filename), exist=ok)
    if (.not. ok) error stop "d2bi error - could not find input file"
    if (nPrec < 1) nPrec = 1
    if (len_trim(outfile0) == 0) opPng = .false.
    if (xpad .and. d0(1) == UNSET) d0(1) = 0
    if (ypad .and. d0(2) == UNSET) d0(2) = 0
    if (zpad .and. d0(3) == UNSET) d0(3) = 0

    write(*, '(a)') ""
    write(*, '(a)') "d2bi v"//trim(VERSION)
    write(*, '(a)') "input filename     : "//trim(filename)
    write(*, '(a,3f10.5,a)') " a0                 : ", a0in, " (A)"
    write(*, '(a,i3,a)') " diffraction conditions ", nDiff, "  mode "//trim(modename)

    !   cubic lattice: a0 and the conventional cell (d2bi.f90:976-983)
    a0 = ((a0in(1) + a0in(2)) + a0in(3)) / 3
    rec0%a_cell = 0
    rec0%a_cell(1) = a0; rec0%a_cell(5) = a0; rec0%a_cell(9) = a0

    !   read atoms (d2bi.f90:796-966)
    call cpu_time(t0)
    call read_atoms(filename, x, a_super, ok)
    if (.not. ok) error stop "d2bi error - could not read atoms / supercell from input file"
    write(*, '(a,i10)') " atoms read         : ", size(x, 2)
    if (addOffset(1) /= UNSET) then
        do ii = 1, size(x, 2)
            x(1:3, ii) = x(1:3, ii) + addOffset
        end do
    end if
    call voxel_super(a_super, M(1), M(2), a0, a_vox, sN)
    do ii = 1, size(x, 2)
        x(1:3, ii) = wrap_pbc(a_vox, sN, x(1:3, ii))
    end do
    !   padding for free surfaces (d2bi.f90:1015-1095)
    if (xpad .or. ypad .or. zpad) then
        if (nTomo > 1) then
            theta = (tomoAngle + THETA_MAX) * PI / 180.0 + precAngle
        else
            theta = THETA_MAX * PI / 180.0 + precAngle
        end if
        if (xpad) call pad_axis(1)
        if (ypad) call pad_axis(2)
        if (zpad) call pad_axis(3)
        if (zpad .and. (d0(3) == 0 .or. d0(3) == UNSET)) d0(3) = pad
    end if
    call voxel_super(a_super, M(1), M(2), a0, a_vox, sN)
    rec0%super_a = reshape(a_vox, [9])
    rec0%super_n = sN
    rec0%a0 = a0
    rec0%E = V

    rc = d2bi_gpu_init(int(device, c_int))
    if (rc /= 0) call fail("d2bi_gpu_init")
    rc = d2bi_gpu_atoms_upload(x, 3_c_int64_t, int(size(x, 2), c_int64_t), atoms)
    if (rc /= 0) call fail("atom upload")
    t_d2bi = 0

    do nn = 1, nDiff
        rec = rec0
        rec%g = g3a(:, nn)
        sg = sga(nn)
        ng = nga(nn)
        dark = darka(nn)
        !   bright field with ng > 1 images the [hkl] point at ng = 1 (d2bi.f90:502-510)
        if (.not. dark .and. ng /= UNSET .and. ng > 1) then
            write(*, '(a)') "d2bi warning - conflicting options '-nodark/-bright' and '-ng > 1'"
            write(*, '(a)') "   setting ng = 1."
            ng = 1
        end if
        write(*, '(a)') ""
        write(*, '(a,i4,a,i4)') " diffraction condition ", nn, "/", nDiff
        write(*, '(a,3f8.2)') "Miller index   g = ", rec%g
        write(*, '(a,3f8.2)') "Miller index   k = ", k3a(:, nn)
        if (sg /= UNSET) then
            modg = (2 * 3.141592654d0 / a0) * norm2(g3a(:, nn))
            if (abs(2 * 3.141592654d0 * sg) > modg) &
                error stop "d2bi error - input parameter s_g/|g| must be < 1/(2pi) ( 1 - |g|/(2|k|) )"
        end if
        if (len_trim(xifile) > 0) then
            call read_extinction_distance(xifile, nint(g3a(:, nn)), xi0, xig, ok)
            if (.not. ok) error stop "d2bi error - failed to find xi_g for [000] and g=[hkl] in extinction distances file"
        end if
        !   image extents from the atoms in the standard frame where -d is unset (d2bi.f90:1206-1230)
        d = d0
        rc = d2bi_rotate_to_standard_frame(k3a(:, nn), rec%a_cell, g3a(:, nn), rot)
        xmin = huge(1.0); xmax = -huge(1.0)
        do ii = 1, size(x, 2)
            yy = rot(:, 1) * x(1, ii) + rot(:, 2) * x(2, ii) + rot(:, 3) * x(3, ii)
            xmin = min(xmin, yy)
            xmax = max(xmax, yy)
        end do
        do ii = 1, 3
            if (d0(ii) == UNSET) d(ii) = xmax(ii) - xmin(ii)
        end do
        if (any(d <= 0)) error stop "d2bi error - '-d 0' (image space from the supercell) is not part of this build"
        if (sg /= UNSET) then
            rc = d2bi_setup_tilt_sg(rec, k3a(:, nn), sg, int(M(1), c_int32_t), int(M(2), c_int32_t), d)
        else if (ng /= UNSET) then
            !   setSupercellTilt1a: the reference sets the table's extinction
            !   distances only after the tilt (d2bi.f90:1265), so the search
            !   sees the constructor's 103.9, 207.5 A
            rec%xi0 = 103.9d0
            rec%xig = 207.5d0
            write(*, '(a,f8.3)') "d2bi info - setting (g,ng) condition and finding crystal rotation and tilt, ng = ", ng
            rc = d2bi_setup_tilt_ng(rec, k3a(:, nn), ng, int(latcode, c_int32_t), trim(xifile)//c_null_char, &
                                    int(M(1), c_int32_t), int(M(2), c_int32_t), d)
        else if (haveR) then
            if (.not. haveU) U = reshape([1, 0, 0, 0, 1, 0, 0, 0, 1], [3, 3])
            rc = d2bi_setup_tilt_ur(rec, reshape(U, [9]), reshape(R, [9]), int(M(1), c_int32_t), &
                                    int(M(2), c_int32_t), d)
        else
            rc = d2bi_setup_tilt_sg(rec, k3a(:, nn), 0.0d0, int(M(1), c_int32_t), int(M(2), c_int32_t), d)
        end if
        if (rc /= 0) call fail("diffraction condition setup")
        rec%xi0 = xi0
        rec%xig = xig
        rec%use_density = merge(1, 0, useDensity)
        U0 = reshape(rec%U, [3, 3])
        R0 = reshape(rec%R, [3, 3])
        write(*, '(a,f12.6,a)') "d2bi info - deviation parameter ", rec%sg, " (1/A)"

        !   output file name (d2bi.f90:1289-1330)
        if (len_trim(outfile0) > 0) then
            write(s1, '(i2)') nint(g3a(1, nn)); write(s2, '(i2)') nint(g3a(2, nn)); write(s3, '(i2)') nint(g3a(3, nn))
            outfile = trim(outfile0)//".g"//trim(adjustl(s1))//"_"//trim(adjustl(s2))//"_"//trim(adjustl(s3))
            write(s1, '(f5.2)') k3a(1, nn); write(s2, '(f5.2)') k3a(2, nn); write(s3, '(f5.2)') k3a(3, nn)
            outfile = trim(outfile)//".k"//trim(adjustl(s1))//"_"//trim(adjustl(s2))//"_"//trim(adjustl(s3))
            write(s1, '(f8.5)') rec%sg
            if (dark) then
                outfile = trim(outfile)//"_df_sg"//trim(adjustl(s1))
            else
                outfile = trim(outfile)//"_bf_sg"//trim(adjustl(s1))
            end if
            if (nPrec > 1) outfile = trim(outfile)//"_p"
        end if

        if (nTomo > 1 .and. useBatch) then
            call tomography_batch()
            cycle
        end if
        do kk = 0, nTomo - 1
            if (nTomo > 1) then
                theta = tomoAngle - 2 * tomoAngle * kk / (nTomo - 1)
                write(*, '(a,i4,a,i4,a,f10.4,a)') " d2bi info - tomography step ", kk + 1, "/", nTomo, &
                    " theta = ", theta, " deg"
                theta = theta * 3.141592654d0 / 180.0d0
                Rt = reshape([1.0d0, 0.0d0, 0.0d0, 0.0d0, cos(theta), sin(theta), 0.0d0, -sin(theta), &
                              cos(theta)], [3, 3])
                Rt = matmul(Rt, R0)
                rc = d2bi_setup_tilt_ur(rec, reshape(U0, [9]), reshape(Rt, [9]), int(M(1), c_int32_t), &
                                        int(M(2), c_int32_t), d)
                if (rc /= 0) call fail("tomography setup")
            end if
            rc = d2bi_plan_image(rec, int(size(x, 2), c_int64_t), precAngle, int(nPrec, c_int32_t), plan)
            if (rc /= 0) call fail("plan")
            Mx = plan%Mx; My = plan%My; Mz = plan%Mz
            if (allocated(img)) deallocate(img)
            allocate(img(0:Mx - 1, 0:My - 1))
            call cpu_time(t0)
            rc = d2bi_gpu_image_dev(rec, atoms, precAngle, int(nPrec, c_int32_t), int(mode, c_int32_t), img, st)
            if (rc /= 0) call fail("computeD2BI_img")
            call cpu_time(t1)
            t_d2bi = t_d2bi + st%ms_total / 1000
            if (.not. dark) img = 1 - img
            write(*, '(a,3i6,a,f10.3,a)') " d2bi info - voxels ", Mx, My, Mz, "  GPU time ", st%ms_total, " ms"
            if (.not. dark) then
                write(*, '(a,3f12.6,a)') "d2bi info - minmaxavg d2bimg     ", minval(img), maxval(img), &
                    sum(img) / (size(img, 1) * size(img, 2)), " (transmitted beam) "
            else
                write(*, '(a,3f12.6,a)') "d2bi info - minmaxavg d2bimg     ", minval(img), maxval(img), &
                    sum(img) / (size(img, 1) * size(img, 2)), " (diffracted beam) "
            end if
            if (len_trim(outfile0) > 0) then
                fname = outfile
                if (nTomo > 1) then
                    write(s1, '(i4.4)') kk
                    fname = trim(outfile)//"."//trim(s1)
                end if
                call write_dat(trim(fname), img, plan)
                if (opPng) call write_png(trim(fname), img, plan%aperpx)
            end if
        end do
    end do

    rc = d2bi_gpu_atoms_free(atoms)
    rc = d2bi_gpu_finalize()
    write(*, '(a)') ""
    write(*, '(a,f16.6)') "   D2BI calc (GPU) ", t_d2bi
    write(*, '(a)') ""

contains

    !   The tomography series of one diffraction condition as one batch: the
    !   records of every tilt R = Rx(theta) R0 (d2bi.f90:1352-1372), imaged by
    !   d2bi_gpu_image_batch, then written step by step as the sequential loop
    !   writes them.
    subroutine tomography_batch()
        integer(int64) :: tot
        real(real64)   :: ms
        allocate(recs(nTomo), hs(nTomo), pas(nTomo), nps(nTomo), offs(nTomo + 1))
        tot = 0
        do kk = 0, nTomo - 1
            theta = tomoAngle - 2 * tomoAngle * kk / (nTomo - 1)
            thetas(kk + 1) = theta
            write(*, '(a,i4,a,i4,a,f10.4,a)') " d2bi info - tomography step ", kk + 1, "/", nTomo, &
                " theta = ", theta, " deg"
            theta = theta * 3.141592654d0 / 180.0d0
            Rt = reshape([1.0d0, 0.0d0, 0.0d0, 0.0d0, cos(theta), sin(theta), 0.0d0, -sin(theta), &
                          cos(theta)], [3, 3])
            Rt = matmul(Rt, R0)
            recs(kk + 1) = rec
            rc = d2bi_setup_tilt_ur(recs(kk + 1), reshape(U0, [9]), reshape(Rt, [9]), int(M(1), c_int32_t), &
                                    int(M(2), c_int32_t), d)
            if (rc /= 0) call fail("tomography setup")
            rc = d2bi_plan_image(recs(kk + 1), int(size(x, 2), c_int64_t), precAngle, int(nPrec, c_int32_t), plan)
            if (rc /= 0) call fail("plan")
            offs(kk + 1) = tot
            tot = tot + int(plan%Mx, int64) * plan%My
            hs(kk + 1) = atoms
            pas(kk + 1) = precAngle
            nps(kk + 1) = int(nPrec, c_int32_t)
        end do
        offs(nTomo + 1) = tot
        allocate(imgs(tot))
        rc = d2bi_gpu_image_batch(recs, hs, int(nTomo, c_int32_t), pas, nps, int(mode, c_int32_t), imgs, st)
        if (rc /= 0) call fail("computeD2BI_img (batch)")
        t_d2bi = t_d2bi + st%ms_total / 1000
        ms = st%ms_total
        write(*, '(a,i6,a,f10.3,a)') " d2bi info - tomography batch of ", nTomo, " tilts, GPU time ", ms, " ms"
        do kk = 0, nTomo - 1
            rec = recs(kk + 1)                 ! the step's record, as the sequential loop leaves it
            rc = d2bi_plan_image(recs(kk + 1), int(size(x, 2), c_int64_t), precAngle, int(nPrec, c_int32_t), plan)
            Mx = plan%Mx; My = plan%My; Mz = plan%Mz
            if (allocated(img)) deallocate(img)
            allocate(img(0:Mx - 1, 0:My - 1))
            img = reshape(imgs(offs(kk + 1) + 1:offs(kk + 2)), [Mx, My])
            if (.not. dark) img = 1 - img
            write(*, '(a,i4,a,f10.4,a,3i6)') " d2bi info - tomography step ", kk + 1, " theta = ", thetas(kk + 1), &
                " deg  voxels ", Mx, My, Mz
            if (.not. dark) then
                write(*, '(a,3f12.6,a)') "d2bi info - minmaxavg d2bimg     ", minval(img), maxval(img), &
                    sum(img) / (size(img, 1) * size(img, 2)), " (transmitted beam) "
            else
                write(*, '(a,3f12.6,a)') "d2bi info - minmaxavg d2bimg     ", minval(img), maxval(img), &
                    sum(img) / (size(img, 1) * size(img, 2)), " (diffracted beam) "
            end if
            if (len_trim(outfile0) > 0) then
                write(s1, '(i4.4)') kk
                fname = trim(outfile)//"."//trim(s1)
                call write_dat(trim(fname), img, plan)
                if (opPng) call write_png(trim(fname), img, plan%aperpx)
            end if
        end do
        deallocate(recs, hs, pas, nps, offs, imgs)
    end subroutine tomography_batch

    subroutine fail(what)
        character(len=*), intent(in) :: what
        write(*, '(a)') "d2bi error - "//what//": "//d2bi_error_message()
        stop 1
    end subroutine fail

    real(real64) function det3(A)
        real(real64), intent(in) :: A(3, 3)
        det3 = A(1, 1) * (A(2, 2) * A(3, 3) - A(2, 3) * A(3, 2)) - A(1, 2) * (A(2, 1) * A(3, 3) - A(2, 3) * A(3, 1)) &
             + A(1, 3) * (A(2, 1) * A(3, 2) - A(2, 2) * A(3, 1))
    end function det3

    !   voxelised input supercell (d2bi.f90:986-992)
    subroutine voxel_super(a_super, m1, m2, a0, a_vox, n3)
        real(real64), intent(in)  :: a_super(3, 3), a0
        integer, intent(in)       :: m1, m2
        real(real64), intent(out) :: a_vox(3, 3)
        integer, intent(out)      :: n3(3)
        n3(1) = max(3, floor(m1 * norm2(a_super(:, 1)) / a0))
        n3(2) = max(3, floor(m1 * norm2(a_super(:, 2)) / a0))
        n3(3) = max(3, floor(m2 * norm2(a_super(:, 3)) / a0))
        a_vox(:, 1) = a_super(:, 1) / n3(1)
        a_vox(:, 2) = a_super(:, 2) / n3(2)
        a_vox(:, 3) = a_super(:, 3) / n3(3)
    end subroutine voxel_super

    !   -xpad/-ypad/-zpad (d2bi.f90:1036-1082)
    subroutine pad_axis(ax)
        integer, intent(in) :: ax
        integer :: o1, o2
        select case (ax)                  ! the two other axes
            case (1); o1 = 2; o2 = 3
            case (2); o1 = 1; o2 = 3
            case default; o1 = 1; o2 = 2
        end select
        len1 = norm2(a_super(ax, :))
        len2 = max(norm2(a_super(o1, :)), norm2(a_super(o2, :)))
        if (d0(1) /= UNSET) len2 = max(len2, d0(1) / cos(theta))
        if (d0(2) /= UNSET) len2 = max(len2, d0(2) / cos(theta))
        pad = len1 * cos(theta) + len2 * sin(theta)
        x(ax, :) = x(ax, :) + (pad - len1) * 1.05 / 2
        a_super(ax, :) = a_super(ax, :) * (pad * 1.05d0 / len1)
        write(*, '(a,i2,a,f12.3)') " d2bi info - padding axis ", ax, " supercell side ", norm2(a_super(ax, :))
    end subroutine pad_axis

    !   .dat image file (d2bi.f90:1429-1457): header then f12.7 values, x fastest
    subroutine write_dat(base, img, plan)
        character(len=*), intent(in)  :: base
        real(real64), intent(in)      :: img(0:, 0:)
        type(d2bi_plan), intent(in)   :: plan
        integer :: u
        real(real64) :: RUm(3, 3)
        open(newunit=u, file=base//".tmp", action="write", status="replace")
        write(u, '(a)') "# d2bi v."//trim(VERSION)
        write(u, '(a,3f10.3)') "# k       ", k3a(:, nn)
        write(u, '(a,3f10.3,f12.5)') "# g       ", rec%g, plan%modg
        write(u, '(a,3i8)') "# M       ", M
        write(u, '(a,3f10.0)') "# V (keV) ", V
        write(u, '(a,3f12.5)') "# xi (A)  ", rec%xi0, rec%xig
        write(u, '(a,3f12.5)') "# s_g (1/A)    ", rec%sg
        write(u, '(a,3f12.3)') "# img space (A)", plan%xmax
        write(u, '(a,3f12.5)') "# scale (A/px) ", plan%aperpx
        RUm = reshape(plan%RU, [3, 3])
        write(u, '(a,9f12.6)') "# rotation ", RUm
        write(u, '(2i8)') size(img, 1), size(img, 2)
        write(u, '(1000f12.7)') img
        close(u)
        call execute_command_line("mv "//base//".tmp "//base//".dat")
        write(*, '(a)') "writing output to """//base//".dat"""
    end subroutine write_dat

    !   outputImgAsPng (d2bi.f90:1509-1538): blur on the GPU, then normalise
    !   or contrast-scale and write 16-bit greyscale (Lib_Png.f90:117-181)
    subroutine write_png(base, img, aperpx)
        character(len=*), intent(in) :: base
        real(real64), intent(in)     :: img(0:, 0:)
        real(real64), intent(in)     :: aperpx
        real(real64) :: aa, bb, lo, hi, one
        integer :: nx, ny, i, j
        nx = size(img, 1); ny = size(img, 2)
        if (allocated(blurred)) deallocate(blurred)
        allocate(blurred(0:nx - 1, 0:ny - 1))
        if (png_blur / aperpx > 0) then
            rc = d2bi_gpu_blur(img, int(nx, c_int32_t), int(ny, c_int32_t), png_blur / aperpx, blurred)
            if (rc /= 0) call fail("blur")
        else
            blurred = img
        end if
        one = 1.0d0 - 1.0d0 / 65536
        if (abs(png_max - png_min) == 0) then
            lo = minval(blurred); hi = maxval(blurred)
            if (hi - lo < 1.0d-16) then
                aa = 0; bb = 0.5d0
            else
                aa = one / (hi - lo); bb = -aa * lo
            end if
        else
            blurred = (blurred - png_min) / (png_max - png_min)
            blurred = max(0.0d0, min(1.0d0, blurred))
            aa = 1; bb = 0
        end if
        if (allocated(pngbuf)) deallocate(pngbuf)
        allocate(pngbuf(nx * ny))
        do j = 0, ny - 1
            do i = 0, nx - 1
                pngbuf(1 + i + nx * j) = real(aa * blurred(i, j) + bb, c_float)
            end do
        end do
        rc = d2bi_write_png16(base//".png"//c_null_char, int(nx, c_int), int(ny, c_int), pngbuf)
        if (rc /= 0) write(*, '(a)') "d2bi warning - png write failed"
        write(*, '(a)') "writing output to """//base//".png"""
    end subroutine write_png

end program d2bi_host
