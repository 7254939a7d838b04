"""The Fortran host's input readers (simulated_tem_amd/host/d2bi_host_io.f90)
on the CPU, through a small probe program compiled with flang:

  LAMMPS write_data, orthogonal and triclinic ("xy xz yz" tilt factors:
      a = [[lx, xy, xz], [0, ly, yz], [0, 0, lz]], Lib_XYZFiles.f90:928-948)
  extended .xyz Lattice="..." with 9 values (columns) or 3 (cuboid) (:950-966)
  PARCAS .xyz "boxsize" (the first three values, cuboid) (:968-983)
  .xyz with "supercell nx ny nz" and three rows of cell vectors in the
      comment lines (:995-1050)
  parse of the number tokens (ParseRealArray1 / ParseReal,
      NBAX_StringTokenizers.f90:273-425): exact for binary fractions, within
      one ulp otherwise, tokens that do not begin with a number skipped
  wrapPBC0 (Lib_SimpleSupercells.f90:992-1021) in a triclinic cell: positions
      inside stay, lattice translations map back, cell coordinates in [0, n)
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "simulated_tem_amd", "host")
FLANG = shutil.which("flang")

pytestmark = pytest.mark.skipif(FLANG is None, reason="flang not found")


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    subprocess.run(["make", "-s", "-C", HOST, "build/d2bi_host_io.o"], check=True, timeout=300)
    exe = str(tmp_path_factory.mktemp("probe") / "host_io_probe")
    subprocess.run([FLANG, "-O2", "-I", os.path.join(HOST, "build"), "-o", exe,
                    os.path.join(ROOT, "tests", "host_io_probe.f90"), os.path.join(HOST, "build", "d2bi_host_io.o")],
                   check=True, timeout=300)
    return exe


def read(probe, path):
    out = subprocess.run([probe, "atoms", str(path)], capture_output=True, text=True, timeout=60, check=True)
    lines = out.stdout.split("\n")
    if lines[0].strip() != "T":
        return None
    n = int(lines[1])
    a = np.array([float(v) for v in " ".join(lines[2:5]).split()]).reshape(3, 3).T      # columns
    xw = np.array([[float(v) for v in ln.split()] for ln in lines[5:5 + n]]).reshape(n, 6)
    return a, xw[:, :3], xw[:, 3:]


def parse(probe, s):
    out = subprocess.run([probe, "parse", s], capture_output=True, text=True, timeout=60, check=True)
    lines = out.stdout.split()
    return [float(v) for v in lines[1:1 + int(lines[0])]]


X = np.array([[0.5, 1.25, 2.0], [3.75, 0.125, 7.5], [9.0, 10.5, 1.0]])


def lammps(path, lo, hi, tilt=None, x=X):
    with open(path, "w") as f:
        f.write("LAMMPS data file via write_data\n\n%d atoms\n1 atom types\n\n" % len(x))
        for i, ax in enumerate("xyz"):
            f.write("%.10f %.10f %slo %shi\n" % (lo[i], hi[i], ax, ax))
        if tilt is not None:
            f.write("%.10f %.10f %.10f xy xz yz\n" % tuple(tilt))
        f.write("\nMasses\n\n1 183.84\n\nAtoms # atomic\n\n")
        for i, p in enumerate(x):
            f.write("%d 1 %.17g %.17g %.17g 0 0 0\n" % (i + 1, p[0], p[1], p[2]))


def xyz(path, desc, comments=(), x=X):
    with open(path, "w") as f:
        for c in comments:
            f.write(c + "\n")
        f.write("%d\n%s\n" % (len(x), desc))
        for p in x:
            f.write("W %.17g %.17g %.17g 1 0.5\n" % tuple(p))


def test_lammps_orthogonal_and_triclinic(probe, tmp_path):
    lammps(tmp_path / "o.lammps", (-1.0, 0.0, 2.0), (11.0, 12.5, 14.0))
    a, x, _ = read(probe, tmp_path / "o.lammps")
    assert np.array_equal(a, np.diag([12.0, 12.5, 12.0]))
    assert np.array_equal(x, X)                       # file order, not shifted by xlo
    lammps(tmp_path / "t.lammps", (0.0, 0.0, 0.0), (12.0, 12.5, 14.0), tilt=(1.5, -0.75, 2.25))
    a, x, _ = read(probe, tmp_path / "t.lammps")
    want = np.array([[12.0, 0.0, 0.0], [1.5, 12.5, 0.0], [-0.75, 2.25, 14.0]]).T    # columns a, b, c
    assert np.array_equal(a, want)
    # no box (all zero): not a supercell
    lammps(tmp_path / "z.lammps", (0.0, 0.0, 0.0), (0.0, 0.0, 0.0))
    assert read(probe, tmp_path / "z.lammps") is None


def test_xyz_lattice_9_and_3_values(probe, tmp_path):
    xyz(tmp_path / "l9.xyz", 'Lattice="12.5 0.5 0.0 -1.0 13.0 0.25 0.0 0.0 14.0" Properties=species:S:1:pos:R:3 pbc="T T T"')
    a, x, _ = read(probe, tmp_path / "l9.xyz")
    assert np.array_equal(a, np.array([[12.5, 0.5, 0.0], [-1.0, 13.0, 0.25], [0.0, 0.0, 14.0]]).T)
    assert np.array_equal(x, X)
    xyz(tmp_path / "l3.xyz", 'Lattice="12.5 13.0 14.0" Properties=species:S:1:pos:R:3')
    a, _, _ = read(probe, tmp_path / "l3.xyz")
    assert np.array_equal(a, np.diag([12.5, 13.0, 14.0]))
    # neither 3 nor 9 numbers: refused
    xyz(tmp_path / "l4.xyz", 'Lattice="12.5 13.0 14.0 15.0" Properties=species:S:1:pos:R:3')
    assert read(probe, tmp_path / "l4.xyz") is None


def test_parcas_boxsize(probe, tmp_path):
    xyz(tmp_path / "p.xyz", "Frame number 0 0.0 fs boxsize 20.25 21.5 22.75 4")
    a, _, _ = read(probe, tmp_path / "p.xyz")
    assert np.array_equal(a, np.diag([20.25, 21.5, 22.75]))
    xyz(tmp_path / "p2.xyz", "Frame number 0 fs boxsize 20.25 21.5")
    assert read(probe, tmp_path / "p2.xyz") is None
    xyz(tmp_path / "none.xyz", "no cell information here")
    assert read(probe, tmp_path / "none.xyz") is None


def test_commented_supercell(probe, tmp_path):
    rows = ["# supercell 2 3 4", "# 3.0 0.0 0.0", "# 0.5 3.0 0.0", "# 0.0 0.25 3.0"]
    xyz(tmp_path / "s.xyz", "W position_x position_y position_z", comments=rows)
    a, _, _ = read(probe, tmp_path / "s.xyz")
    rows_m = np.array([[3.0, 0.0, 0.0], [0.5, 3.0, 0.0], [0.0, 0.25, 3.0]])   # a(kk, :) = line kk
    assert np.array_equal(a, rows_m * np.array([2, 3, 4])[None, :])          # column j times n_j
    xyz(tmp_path / "s9.xyz", "W position_x position_y position_z",
        comments=["% super cell 1 1 2", "% 3.0 0.0 0.0 0.0 3.0 0.0 0.0 0.0 3.5"])
    a, _, _ = read(probe, tmp_path / "s9.xyz")
    assert np.array_equal(a, np.diag([3.0, 3.0, 7.0]))


def test_number_tokens(probe):
    assert parse(probe, '12.5 0.25 -1.5e2 1d1 .5" -.5') == [12.5, 0.25, -150.0, 10.0, 0.5, -0.5]
    assert parse(probe, "Properties=species:S:1:pos:R:3 pbc=T 7") == [7.0]
    assert parse(probe, "1,2;3\t4") == [1.0, 2.0, 3.0, 4.0]
    for s in ("3.61", "3.1652", "74.731", "0.79502"):
        (v,) = parse(probe, s)
        assert abs(v - float(s)) <= np.spacing(float(s))
    assert len(parse(probe, " ".join(["1"] * 12))) == 9        # at most nine


def test_wrap_pbc_triclinic(probe, tmp_path):
    a = np.array([[12.0, 0.0, 0.0], [1.5, 12.5, 0.0], [-0.75, 2.25, 14.0]]).T
    rng = np.random.default_rng(4)
    f = rng.uniform(0.02, 0.98, size=(40, 3))
    inside = f @ a.T
    shifts = rng.integers(-2, 3, size=(40, 3))
    moved = inside + shifts @ a.T
    x = np.vstack([inside, moved])
    lammps(tmp_path / "w.lammps", (0.0, 0.0, 0.0), (12.0, 12.5, 14.0), tilt=(1.5, -0.75, 2.25), x=x)
    a_r, xr, xw = read(probe, tmp_path / "w.lammps")
    assert np.array_equal(a_r, a)
    assert np.allclose(xw[:40], inside, rtol=0, atol=1e-12)
    assert np.allclose(xw[40:], inside, rtol=0, atol=1e-11)
    frac = np.linalg.solve(a / 3.0, xw.T).T             # voxel-supercell coordinates, n = 3
    assert np.all(frac > -1e-12) and np.all(frac < 3.0 + 1e-12)
