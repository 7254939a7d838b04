"""The Fortran wrappers that INTEGRATION.md gives a maintainer of the
reference (computeD2BI_img, fill_record, neighbourLists_gpu,
findGdotR_linkcells) compile with flang against simulated_tem_amd/host's
d2bi_gpu_mod and link against libd2bi_gpu.so.  The reference's private
types are stood in for by types with the fields the wrappers read
(Lib_DynamicalTwoBeamImaging.f90:64-90, SimpleSupercell's getA/getNx/...).
CPU only: the program is linked, not run."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "simulated_tem_amd", "host")
LIBDIR = os.path.join(ROOT, "simulated_tem_amd", "_lib")
FLANG = shutil.which("flang")

pytestmark = pytest.mark.skipif(FLANG is None, reason="flang not found")

STANDINS = """
module integration_check
    use iso_fortran_env, only: real64
    use iso_c_binding
    implicit none
    type SimpleSupercell
        real(kind=real64), dimension(3, 3) :: a = 0
        integer :: nx = 1, ny = 1, nz = 1
    end type SimpleSupercell
    type DynamicalTwoBeamImaging
        real(kind=real64), dimension(3) :: k, g
        real(kind=real64) :: a0
        type(SimpleSupercell) :: super
        real(kind=real64) :: sg
        integer :: mxy, mzz
        real(kind=real64) :: xi0, xig, E
        logical :: useDensityField
        type(SimpleSupercell) :: gksuper
        real(kind=real64), dimension(3) :: delta
        integer :: nuvw
        integer, dimension(3, 125) :: uvw
        real(kind=real64), dimension(3, 3) :: a_cell, U, R
        logical :: test = .false.
    end type DynamicalTwoBeamImaging
contains
    function getA(s) result(a)
        type(SimpleSupercell), intent(in) :: s
        real(kind=real64), dimension(3, 3) :: a
        a = s%a
    end function getA
    integer function getNx(s)
        type(SimpleSupercell), intent(in) :: s
        getNx = s%nx
    end function getNx
    integer function getNy(s)
        type(SimpleSupercell), intent(in) :: s
        getNy = s%ny
    end function getNy
    integer function getNz(s)
        type(SimpleSupercell), intent(in) :: s
        getNz = s%nz
    end function getNz
@WRAPPERS@
end module integration_check

program link_integration
    use integration_check
    implicit none
    print *, "linked"
end program link_integration
"""


def integration_blocks():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```fortran\n(.*?)```", text, re.S)
    return [b for b in blocks if b.lstrip().startswith("subroutine")]


def test_integration_wrappers_compile_and_link(tmp_path):
    blocks = integration_blocks()
    names = sorted(re.findall(r"^\s*subroutine\s+(\w+)", "\n".join(blocks), re.M))
    assert names == ["computeD2BI_img", "fill_record", "findGdotR_linkcells", "neighbourLists_gpu"]
    subprocess.run(["make", "-s", "-C", HOST, "build/d2bi_gpu_mod.o"], check=True, timeout=300)
    src = tmp_path / "integration_check.f90"
    src.write_text(STANDINS.replace("@WRAPPERS@", "\n".join(blocks)))
    lib = os.path.join(LIBDIR, "libd2bi_gpu.so")
    if not os.path.exists(lib):
        pytest.skip("libd2bi_gpu.so not built")
    exe = str(tmp_path / "link_integration")
    r = subprocess.run([FLANG, "-O1", "-I", os.path.join(HOST, "build"), "-J", str(tmp_path), "-o", exe, str(src),
                        os.path.join(HOST, "build", "d2bi_gpu_mod.o"), "-L" + LIBDIR, "-ld2bi_gpu",
                        "-Wl,-rpath-link,/opt/rocm/lib", "-Wl,--no-as-needed"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    # every library entry point the wrappers call is one the library exports
    nm = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\b(d2bi_\w+)$", nm, re.M))
    called = set(re.findall(r"\b(d2bi_gpu_\w+)\s*\(", "\n".join(blocks)))
    assert called and called <= exported, called - exported


def test_fortran_module_declares_every_header_entry_point():
    """d2bi_gpu_mod has a bind(c) interface for every function of
    include/d2bi_gpu.h (d2bi_write_png16, the host's PNG writer, besides)."""
    hdr = open(os.path.join(ROOT, "include", "d2bi_gpu.h")).read()
    declared = set(re.findall(r"^[a-z_0-9\* ]+?\b(d2bi_\w+)\(", hdr, re.M))
    mod = open(os.path.join(HOST, "d2bi_gpu_mod.f90")).read()
    bound = set(re.findall(r'bind\(c, name="(d2bi_\w+)"\)', mod))
    public = set(w for line in re.findall(r"^\s*public\s*::(.*)$", mod, re.M) for w in re.findall(r"\w+", line))
    assert declared and declared <= bound, declared - bound
    assert declared <= public, declared - public
