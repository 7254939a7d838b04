!   d2bi_gpu_mod.f90 -- iso_c_binding view of the imaging library's C-ABI
!   (include/d2bi_gpu.h).  The derived types are bind(c) mirrors of the C
!   structs; every interface is one entry point of libd2bi_gpu.so.
module d2bi_gpu_mod
    use iso_c_binding
    implicit none
    private

    integer(c_int), parameter, public :: D2BI_MODE_FP64 = 0
    integer(c_int), parameter, public :: D2BI_MODE_FAST = 1
    integer, parameter, public        :: D2BI_MAX_UVW = 125
    integer, parameter, public        :: D2BI_MAX_PREC = 256
    integer, parameter, public        :: D2BI_MAX_APERTURE = 1024
    integer(c_int64_t), parameter, public :: D2BI_SCRATCH_BYTES = 67108864_c_int64_t
    integer(c_int), parameter, public :: D2BI_OK = 0, D2BI_ERR_ARG = -1, D2BI_ERR_HIP = -2, &
                                         D2BI_ERR_CAPACITY = -3, D2BI_ERR_NOINIT = -4, D2BI_ERR_NODEVICE = -5

    !   type DynamicalTwoBeamImaging fields read by computeD2BI_img
    !   (Lib_DynamicalTwoBeamImaging.f90:64-90)
    type, bind(c), public :: d2bi_record
        real(c_double)     :: a0
        real(c_double)     :: a_cell(9)
        real(c_double)     :: g(3)
        real(c_double)     :: E
        real(c_double)     :: xi0, xig
        real(c_double)     :: U(9), R(9)
        real(c_double)     :: delta(3)
        real(c_double)     :: super_a(9)
        integer(c_int32_t) :: super_n(3)
        integer(c_int32_t) :: nuvw
        integer(c_int32_t) :: uvw(3*D2BI_MAX_UVW)
        real(c_double)     :: gk_a(9)
        integer(c_int32_t) :: gk_n(3)
        integer(c_int32_t) :: use_density
        real(c_double)     :: sg
    end type d2bi_record

    !   computeD2BI_img prologue (Lib_DynamicalTwoBeamImaging.f90:1635-1836)
    type, bind(c), public :: d2bi_plan
        integer(c_int32_t) :: Mx, My, Mz, deltaM, nMax, Ex, Ey, Ez, nPrec, pad0
        real(c_double)     :: sigma, i2s2, rc2
        real(c_double)     :: a_col(9), ia_col(9), lc_delta(3), RU(9), gg(3), xmax(3)
        real(c_double)     :: aperpx, modg, dz, b0, bg
        real(c_double)     :: aa(D2BI_MAX_PREC)
    end type d2bi_plan

    type, bind(c), public :: d2bi_stats
        integer(c_int64_t) :: n_points, n_guard
        integer(c_int32_t) :: max_per_cell, n_cells
        real(c_double)     :: ms_bin, ms_voxel, ms_integrate, ms_total
        integer(c_int64_t) :: branch_bytes
        integer(c_int32_t) :: n_branches, pad0
    end type d2bi_stats

    !   d2bi_synth_bcc parameters (no counterpart in the reference)
    type, bind(c), public :: d2bi_synth_params
        real(c_double)     :: a0
        integer(c_int32_t) :: nx, ny, nz, n_centres
        real(c_double)     :: A_min, A_max, u_cap
        integer(c_int32_t) :: n_voids, pad0
        real(c_double)     :: r_min, r_max, noise, offset(3)
        integer(c_int64_t) :: seed
    end type d2bi_synth_params

    public :: d2bi_gpu_init, d2bi_gpu_finalize, d2bi_last_error, d2bi_version, d2bi_plan_image
    public :: d2bi_setup_tilt_sg, d2bi_setup_tilt_ur, d2bi_setup_tilt_ng, d2bi_rotate_to_standard_frame
    public :: d2bi_aperture_points, d2bi_aperture_setup, d2bi_wavelength, d2bi_velocity, d2bi_deviation_parameter
    public :: d2bi_gpu_atoms_upload, d2bi_gpu_atoms_free, d2bi_gpu_atoms_count
    public :: d2bi_gpu_image, d2bi_gpu_image_dev, d2bi_gpu_image_dev_out, d2bi_gpu_image_noalloc
    public :: d2bi_gpu_image_aperture, d2bi_gpu_image_batch, d2bi_gpu_image_batch_dev_out
    public :: d2bi_gpu_field, d2bi_gpu_linkcells, d2bi_gpu_lc_query, d2bi_gpu_blur, d2bi_synth_bcc
    public :: d2bi_write_png16, d2bi_error_message

    interface
        integer(c_int) function d2bi_gpu_init(device) bind(c, name="d2bi_gpu_init")
            import :: c_int
            integer(c_int), value :: device
        end function
        integer(c_int) function d2bi_gpu_finalize() bind(c, name="d2bi_gpu_finalize")
            import :: c_int
        end function
        type(c_ptr) function d2bi_last_error() bind(c, name="d2bi_last_error")
            import :: c_ptr
        end function
        integer(c_int) function d2bi_plan_image(rec, nAtoms, precAngle, nPrec, plan) &
                bind(c, name="d2bi_plan_image")
            import :: c_int, c_int32_t, c_int64_t, c_double, d2bi_record, d2bi_plan
            type(d2bi_record), intent(in)   :: rec
            integer(c_int64_t), value       :: nAtoms
            real(c_double), value           :: precAngle
            integer(c_int32_t), value       :: nPrec
            type(d2bi_plan), intent(out)    :: plan
        end function
        integer(c_int) function d2bi_setup_tilt_sg(rec, k, sg, mxy, mzz, d) &
                bind(c, name="d2bi_setup_tilt_sg")
            import :: c_int, c_int32_t, c_double, d2bi_record
            type(d2bi_record), intent(inout) :: rec
            real(c_double), intent(in)       :: k(3)
            real(c_double), value            :: sg
            integer(c_int32_t), value        :: mxy, mzz
            real(c_double), intent(in)       :: d(3)
        end function
        integer(c_int) function d2bi_setup_tilt_ur(rec, U, R, mxy, mzz, d) &
                bind(c, name="d2bi_setup_tilt_ur")
            import :: c_int, c_int32_t, c_double, d2bi_record
            type(d2bi_record), intent(inout) :: rec
            real(c_double), intent(in)       :: U(9), R(9)
            integer(c_int32_t), value        :: mxy, mzz
            real(c_double), intent(in)       :: d(3)
        end function
        integer(c_int) function d2bi_setup_tilt_ng(rec, k, ng, lattice, xifile, mxy, mzz, d) &
                bind(c, name="d2bi_setup_tilt_ng")
            import :: c_int, c_int32_t, c_double, c_char, d2bi_record
            type(d2bi_record), intent(inout) :: rec
            real(c_double), intent(in)       :: k(3)
            real(c_double), value            :: ng
            integer(c_int32_t), value        :: lattice
            character(kind=c_char), intent(in) :: xifile(*)
            integer(c_int32_t), value        :: mxy, mzz
            real(c_double), intent(inout)    :: d(3)
        end function
        integer(c_int) function d2bi_rotate_to_standard_frame(k, a_cell, hkl, U) &
                bind(c, name="d2bi_rotate_to_standard_frame")
            import :: c_int, c_double
            real(c_double), intent(in)  :: k(3), a_cell(9), hkl(3)
            real(c_double), intent(out) :: U(9)
        end function
        integer(c_int) function d2bi_gpu_atoms_upload(x, ld, nAtoms, atoms) &
                bind(c, name="d2bi_gpu_atoms_upload")
            import :: c_int, c_int64_t, c_double, c_ptr
            real(c_double), intent(in)  :: x(*)
            integer(c_int64_t), value   :: ld, nAtoms
            type(c_ptr), intent(out)    :: atoms
        end function
        integer(c_int) function d2bi_gpu_atoms_free(atoms) bind(c, name="d2bi_gpu_atoms_free")
            import :: c_int, c_ptr
            type(c_ptr), value :: atoms
        end function
        integer(c_int64_t) function d2bi_gpu_atoms_count(atoms) bind(c, name="d2bi_gpu_atoms_count")
            import :: c_int64_t, c_ptr
            type(c_ptr), value :: atoms
        end function
        ! computeD2BI_img with host positions x(ld, nAtoms), first three rows read
        integer(c_int) function d2bi_gpu_image(rec, x, ld, nAtoms, precAngle, nPrec, mode, img, stats) &
                bind(c, name="d2bi_gpu_image")
            import :: c_int, c_int32_t, c_int64_t, c_double, d2bi_record, d2bi_stats
            type(d2bi_record), intent(in)   :: rec
            real(c_double), intent(in)      :: x(*)
            integer(c_int64_t), value       :: ld, nAtoms
            real(c_double), value           :: precAngle
            integer(c_int32_t), value       :: nPrec, mode
            real(c_double), intent(out)     :: img(*)
            type(d2bi_stats), intent(out)   :: stats
        end function
        ! the image stays in HBM: d_img is a device address of Mx*My doubles
        integer(c_int) function d2bi_gpu_image_dev_out(rec, atoms, precAngle, nPrec, mode, d_img, stats) &
                bind(c, name="d2bi_gpu_image_dev_out")
            import :: c_int, c_int32_t, c_double, c_ptr, d2bi_record, d2bi_stats
            type(d2bi_record), intent(in)   :: rec
            type(c_ptr), value              :: atoms
            real(c_double), value           :: precAngle
            integer(c_int32_t), value       :: nPrec, mode
            type(c_ptr), value              :: d_img
            type(d2bi_stats), intent(out)   :: stats
        end function
        ! as d2bi_gpu_image_dev_out with no device allocation (workspace: a
        ! static block of D2BI_SCRATCH_BYTES per device)
        integer(c_int) function d2bi_gpu_image_noalloc(rec, atoms, precAngle, nPrec, mode, d_img, stats) &
                bind(c, name="d2bi_gpu_image_noalloc")
            import :: c_int, c_int32_t, c_double, c_ptr, d2bi_record, d2bi_stats
            type(d2bi_record), intent(in)   :: rec
            type(c_ptr), value              :: atoms
            real(c_double), value           :: precAngle
            integer(c_int32_t), value       :: nPrec, mode
            type(c_ptr), value              :: d_img
            type(d2bi_stats), intent(out)   :: stats
        end function
        ! voxel field; each output may be c_null_ptr (else c_loc of an array)
        integer(c_int) function d2bi_gpu_field(rec, x, ld, nAtoms, mode, phase_re_im, rho, count, r2min) &
                bind(c, name="d2bi_gpu_field")
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr, d2bi_record
            type(d2bi_record), intent(in)   :: rec
            real(c_double), intent(in)      :: x(*)
            integer(c_int64_t), value       :: ld, nAtoms
            integer(c_int32_t), value       :: mode
            type(c_ptr), value              :: phase_re_im, rho, count, r2min
        end function
        ! findGdotR's link cells; each output may be c_null_ptr
        integer(c_int) function d2bi_gpu_linkcells(rec, x, ld, nAtoms, cell_start, ids, frac, gdotr, cap, &
                nPoints) bind(c, name="d2bi_gpu_linkcells")
            import :: c_int, c_int64_t, c_double, c_ptr, d2bi_record
            type(d2bi_record), intent(in)   :: rec
            real(c_double), intent(in)      :: x(*)
            integer(c_int64_t), value       :: ld, nAtoms
            type(c_ptr), value              :: cell_start, ids, frac, gdotr
            integer(c_int64_t), value       :: cap
            integer(c_int64_t), intent(out) :: nPoints
        end function
        ! LinkCell3D_ctor0 + add1 + neighbourList1b for nQ query points
        integer(c_int) function d2bi_gpu_lc_query(a, Nx, Ny, Nz, p, nP, q, nQ, rc, maxNeigh, n, id, r2) &
                bind(c, name="d2bi_gpu_lc_query")
            import :: c_int, c_int32_t, c_int64_t, c_double
            real(c_double), intent(in)      :: a(9)
            integer(c_int32_t), value       :: Nx, Ny, Nz
            real(c_double), intent(in)      :: p(*)
            integer(c_int64_t), value       :: nP
            real(c_double), intent(in)      :: q(*)
            integer(c_int64_t), value       :: nQ
            real(c_double), value           :: rc
            integer(c_int32_t), value       :: maxNeigh
            integer(c_int32_t), intent(out) :: n(*), id(*)
            real(c_double), intent(out)     :: r2(*)
        end function
        integer(c_int) function d2bi_aperture_points(n, xy) bind(c, name="d2bi_aperture_points")
            import :: c_int, c_int32_t, c_double
            integer(c_int32_t), value   :: n
            real(c_double), intent(out) :: xy(*)
        end function
        integer(c_int) function d2bi_aperture_setup(rec, precAngle, nPrec, apertureAngle, nAperture, gshift, aa) &
                bind(c, name="d2bi_aperture_setup")
            import :: c_int, c_int32_t, c_double, d2bi_record
            type(d2bi_record), intent(in) :: rec
            real(c_double), value         :: precAngle, apertureAngle
            integer(c_int32_t), value     :: nPrec, nAperture
            real(c_double), intent(out)   :: gshift(*), aa(*)
        end function
        real(c_double) function d2bi_wavelength(V) bind(c, name="d2bi_wavelength")
            import :: c_double
            real(c_double), value :: V
        end function
        real(c_double) function d2bi_velocity(V) bind(c, name="d2bi_velocity")
            import :: c_double
            real(c_double), value :: V
        end function
        real(c_double) function d2bi_deviation_parameter(E_keV, g_lab) bind(c, name="d2bi_deviation_parameter")
            import :: c_double
            real(c_double), value      :: E_keV
            real(c_double), intent(in) :: g_lab(3)
        end function
        integer(c_int) function d2bi_version() bind(c, name="d2bi_version")
            import :: c_int
        end function
        integer(c_int) function d2bi_synth_bcc(p, x, cap, nAtoms, box) bind(c, name="d2bi_synth_bcc")
            import :: c_int, c_int64_t, c_double, d2bi_synth_params
            type(d2bi_synth_params), intent(in) :: p
            real(c_double), intent(out)         :: x(*)
            integer(c_int64_t), value           :: cap
            integer(c_int64_t), intent(out)     :: nAtoms
            real(c_double), intent(out)         :: box(3)
        end function
        integer(c_int) function d2bi_gpu_image_dev(rec, atoms, precAngle, nPrec, mode, img, stats) &
                bind(c, name="d2bi_gpu_image_dev")
            import :: c_int, c_int32_t, c_double, c_ptr, d2bi_record, d2bi_stats
            type(d2bi_record), intent(in)   :: rec
            type(c_ptr), value              :: atoms
            real(c_double), value           :: precAngle
            integer(c_int32_t), value       :: nPrec, mode
            real(c_double), intent(out)     :: img(*)
            type(d2bi_stats), intent(out)   :: stats
        end function
        integer(c_int) function d2bi_gpu_image_aperture(rec, x, ld, nAtoms, precAngle, nPrec, apertureAngle, &
                nAperture, mode, img, stats) bind(c, name="d2bi_gpu_image_aperture")
            import :: c_int, c_int32_t, c_int64_t, c_double, d2bi_record, d2bi_stats
            type(d2bi_record), intent(in)   :: rec
            real(c_double), intent(in)      :: x(*)
            integer(c_int64_t), value       :: ld, nAtoms
            real(c_double), value           :: precAngle, apertureAngle
            integer(c_int32_t), value       :: nPrec, nAperture, mode
            real(c_double), intent(out)     :: img(*)
            type(d2bi_stats), intent(out)   :: stats
        end function
        integer(c_int) function d2bi_gpu_image_batch(recs, atoms, nRec, precAngle, nPrec, mode, img, stats) &
                bind(c, name="d2bi_gpu_image_batch")
            import :: c_int, c_int32_t, c_double, c_ptr, d2bi_record, d2bi_stats
            type(d2bi_record), intent(in)   :: recs(*)
            type(c_ptr), intent(in)         :: atoms(*)
            integer(c_int32_t), value       :: nRec, mode
            real(c_double), intent(in)      :: precAngle(*)
            integer(c_int32_t), intent(in)  :: nPrec(*)
            real(c_double), intent(out)     :: img(*)
            type(d2bi_stats), intent(out)   :: stats
        end function
        ! the images stay in HBM: d_img is a device address (e.g. from hipMalloc)
        integer(c_int) function d2bi_gpu_image_batch_dev_out(recs, atoms, nRec, precAngle, nPrec, mode, d_img, &
                stats) bind(c, name="d2bi_gpu_image_batch_dev_out")
            import :: c_int, c_int32_t, c_double, c_ptr, d2bi_record, d2bi_stats
            type(d2bi_record), intent(in)   :: recs(*)
            type(c_ptr), intent(in)         :: atoms(*)
            integer(c_int32_t), value       :: nRec, mode
            real(c_double), intent(in)      :: precAngle(*)
            integer(c_int32_t), intent(in)  :: nPrec(*)
            type(c_ptr), value              :: d_img
            type(d2bi_stats), intent(out)   :: stats
        end function
        integer(c_int) function d2bi_gpu_blur(fin, nx, ny, t, fout) bind(c, name="d2bi_gpu_blur")
            import :: c_int, c_int32_t, c_double
            real(c_double), intent(in)  :: fin(*)
            integer(c_int32_t), value   :: nx, ny
            real(c_double), value       :: t
            real(c_double), intent(out) :: fout(*)
        end function
        integer(c_int) function d2bi_write_png16(path, width, height, img) bind(c, name="d2bi_write_png16")
            import :: c_int, c_char, c_float
            character(kind=c_char), intent(in) :: path(*)
            integer(c_int), value              :: width, height
            real(c_float), intent(in)          :: img(*)
        end function
        integer(c_size_t) function c_strlen(s) bind(c, name="strlen")
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function
    end interface

contains

    !   the library's last error message as a Fortran string
    function d2bi_error_message() result(msg)
        character(len=:), allocatable :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: s(:)
        integer :: n, i
        p = d2bi_last_error()
        if (.not. c_associated(p)) then
            msg = ""
            return
        end if
        n = int(c_strlen(p))
        call c_f_pointer(p, s, [n])
        allocate(character(len=n) :: msg)
        do i = 1, n
            msg(i:i) = s(i)
        end do
    end function d2bi_error_message

end module d2bi_gpu_mod
