"""The Fortran d2bi host (simulated_tem_amd/host): command line, input and
output files.  Error paths run on the CPU; imaging runs on the GPU."""
import os
import re
import subprocess
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "simulated_tem_amd", "host", "d2bi")

needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="d2bi host not built (flang missing)")


def run(args, cwd):
    return subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=300)


def write_lammps(path, x, box, tilt=None):
    with open(path, "w") as f:
        f.write("LAMMPS data file written by the d2bi host tests\n\n%d atoms\n1 atom types\n\n" % len(x))
        for i, ax in enumerate("xyz"):
            f.write("%.10f %.10f %slo %shi\n" % (box[i][0], box[i][1], ax, ax))
        if tilt is not None:
            f.write("%.10f %.10f %.10f xy xz yz\n" % tuple(tilt))
        f.write("\nMasses\n\n1 183.84\n\nAtoms # atomic\n\n")
        for i, p in enumerate(x):
            f.write("%d 1 %.16g %.16g %.16g 0 0 0\n" % (i + 1, p[0], p[1], p[2]))


def write_xi(path, xi0, xig, hkl):
    with open(path, "w") as f:
        f.write("# extinction distances for the d2bi host tests\n")
        f.write(" element          g        T (K)      V (V)         Re(U_g)         Im(U_g)            xi_g\n")
        f.write("      W    0   0   0     300.000  200000.000     37.83095547      1.11167480 %15.8f\n" % xi0)
        f.write("      W    1   1   0     300.000  200000.000     18.91115463      1.07133212    227.56411986\n")
        f.write("      W  %3d %3d %3d     300.000  200000.000     13.94499223      1.03823539 %15.8f\n"
                % (hkl[0], hkl[1], hkl[2], xig))


@needs_exe
def test_usage_and_errors(tmp_path):
    r = run([], tmp_path)
    assert r.returncode != 0 and "usage" in r.stdout
    r = run(["-f", "missing.lammps", "-a0", "3.1652", "-g", "2,0,0", "-k", "0,0,1"], tmp_path)
    assert r.returncode != 0 and "could not find input file" in (r.stdout + r.stderr)
    (tmp_path / "a.lammps").write_text("x")
    r = run(["-f", "a.lammps", "-a0", "3.1652", "-g", "2,0,0", "-k", "0,0,1", "-ng", "1"], tmp_path)
    assert r.returncode != 0 and "needs -xifile" in (r.stdout + r.stderr)
    r = run(["-f", "a.lammps", "-a0", "3.1652", "-g", "2,0,0", "-k", "0,0,1", "-ng", "1", "-sg", "0.01"],
            tmp_path)
    assert r.returncode != 0 and "both n_g and s_g" in (r.stdout + r.stderr)
    r = run(["-f", "a.lammps", "-a0", "3.1652", "-g", "2,0,0", "-k", "0,0,1", "-alterg"], tmp_path)
    assert r.returncode != 0 and "not part of this build" in (r.stdout + r.stderr)
    r = run(["-f", "a.lammps", "-a0", "3.1652", "-g", "2,0,0", "-k", "0,0,1", "-xifile", "x.dat",
             "-xi0", "100"], tmp_path)
    assert r.returncode != 0 and "xi_0 or xi_g" in (r.stdout + r.stderr)
    r = run(["-f", "a.lammps", "-a0", "3.1652", "-g", "2,0,0,1", "-k", "0,0,1"], tmp_path)
    assert r.returncode != 0 and "g-vectors" in (r.stdout + r.stderr)


def _read_png16(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w = 8, b"", None
    while pos < len(data):
        n = int.from_bytes(data[pos:pos + 4], "big")
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc = int.from_bytes(data[pos + 8 + n:pos + 12 + n], "big")
        assert zlib.crc32(typ + body) & 0xffffffff == crc
        if typ == b"IHDR":
            w, h = int.from_bytes(body[:4], "big"), int.from_bytes(body[4:8], "big")
            assert body[8] == 16 and body[9] == 0
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 2 * w)
    assert np.all(rows[:, 0] == 0)
    px = rows[:, 1:].reshape(h, w, 2).astype(np.int64)
    return px[:, :, 0] * 256 + px[:, :, 1]


def d2bi_script_gate(got):
    """testD2BI.sh's pass test: int(max |(x - target)/(1/256) + 0.5|) == 0
    over min, max and avg, i.e. x - target in (-1.5/256, 0.5/256)."""
    acc = 0.0039
    worst = max(abs((a - b) / acc + 0.5) for a, b in zip(got, (0.000028, 0.230127, 0.081501)))
    return int(worst) == 0


@needs_exe
@pytest.mark.gpu
def test_testd2bi_ng_run_through_host(tmp_path, golden_dir):
    """testD2BI.sh itself: -lattice bcc -k 0,0,1 -g 2,0,0 -xifile <W table>
    -ng 1 -o "" on the W loop.  The tilt comes from this build's
    (g,ng) search; the printed minmaxavg must pass the script's gate."""
    w = np.load(os.path.join(golden_dir, "w_loop_golden.npz"))
    write_lammps(tmp_path / "w.lammps", w["x"], w["box"])
    xi = os.path.join(golden_dir, "W_V200_T300.dat")
    for mode in ("fp64", "fast"):
        r = run(["-f", "w.lammps", "-a0", "3.1652", "-lattice", "bcc", "-k", "0,0,1", "-g", "2,0,0",
                 "-xifile", xi, "-ng", "1", "-o", "", "-mode", mode], tmp_path)
        assert r.returncode == 0, r.stdout + r.stderr
        m = re.search(r"minmaxavg d2bimg\s+(\S+)\s+(\S+)\s+(\S+)", r.stdout)
        got = [float(v) for v in m.groups()]
        print("[testD2BI -ng 1] %s minmaxavg %s" % (mode, got), flush=True)
        assert d2bi_script_gate(got), got


@needs_exe
@pytest.mark.gpu
def test_golden_run_through_host(tmp_path, golden_dir):
    """testD2BI.sh's run, with the golden image's rotation given by -U/-R:
    minmaxavg within 1/256 of 0.000028/0.230127/0.081501, .dat within 1e-4
    of the golden image, .png within a few grey levels of the golden .png
    (Data/test/...ng1.00.png, decoded into the fixture)."""
    w = np.load(os.path.join(golden_dir, "w_loop_golden.npz"))
    write_lammps(tmp_path / "w.lammps", w["x"], w["box"])
    write_xi(tmp_path / "xi.dat", w["xi"][0], w["xi"][1], (2, 0, 0))
    U = ",".join("%.6f" % v for v in w["rotation_cm"])
    d = ",".join("%.3f" % v for v in w["space"])
    for mode in ("fp64", "fast"):
        r = run(["-f", "w.lammps", "-a0", "3.1652", "-lattice", "bcc", "-k", "0,0,1", "-g", "2,0,0",
                 "-xifile", "xi.dat", "-U", U, "-R", "1,0,0,0,1,0,0,0,1", "-d", d, "-o", "out_" + mode,
                 "-mode", mode], tmp_path)
        assert r.returncode == 0, r.stdout + r.stderr
        m = re.search(r"minmaxavg d2bimg\s+(\S+)\s+(\S+)\s+(\S+)", r.stdout)
        got = [float(v) for v in m.groups()]
        for a, b in zip(got, (0.000028, 0.230127, 0.081501)):
            assert abs(a - b) <= 0.0039
        dats = [f for f in os.listdir(tmp_path) if f.startswith("out_" + mode) and f.endswith(".dat")]
        assert len(dats) == 1
        lines = open(tmp_path / dats[0]).read().splitlines()
        assert lines[0].startswith("# d2bi v.")
        h = next(i for i, ln in enumerate(lines) if not ln.startswith("#"))
        assert any(ln.startswith("# rotation") for ln in lines[:h])
        nx, ny = (int(v) for v in lines[h].split())
        vals = [float(ln[i:i + 12]) for ln in lines[h + 1:] for i in range(0, len(ln), 12)
                if ln[i:i + 12].strip()]
        img = np.array(vals).reshape(ny, nx)
        assert img.shape == w["image"].shape
        assert np.max(np.abs(img - w["image"])) < 1e-4
        png = _read_png16(tmp_path / (dats[0][:-4] + ".png"))
        assert png.shape == img.shape
        assert png.min() == 0 and png.max() >= 65534     # normalised output spans the range
        # the reference's own PNG of this run (blurred by 2.5 A, normalised):
        # the oracle's image of it is within 3 grey levels of 65536
        dpng = np.abs(png - w["png"].astype(np.int64))
        print("[golden png] %s: max %d levels, mean %.3f" % (mode, dpng.max(), dpng.mean()), flush=True)
        assert dpng.max() <= (4 if mode == "fp64" else 6)


@needs_exe
@pytest.mark.gpu
def test_tomography_series_batched_equals_sequential(tmp_path, golden_dir):
    """-nTomoAngle 3 on the W loop: the series imaged as one batch (one
    captured graph for the three tilts) writes the same .dat files, byte for
    byte, as the tilts imaged one at a time (-nobatch), in both modes; the
    tilts differ from each other."""
    w = np.load(os.path.join(golden_dir, "w_loop_golden.npz"))
    write_lammps(tmp_path / "w.lammps", w["x"], w["box"])
    write_xi(tmp_path / "xi.dat", w["xi"][0], w["xi"][1], (2, 0, 0))
    base = ["-f", "w.lammps", "-a0", "3.1652", "-lattice", "bcc", "-k", "0,0,1", "-g", "2,0,0",
            "-xifile", "xi.dat", "-sg", "0.02", "-nTomoAngle", "3", "-tomoAngle", "2", "-nopng"]
    for mode in ("fp64", "fast"):
        outs = {}
        for tag, extra in (("batch", []), ("seq", ["-nobatch"])):
            r = run(base + ["-o", tag + "_" + mode, "-mode", mode] + extra, tmp_path)
            assert r.returncode == 0, r.stdout + r.stderr
            dats = sorted(f for f in os.listdir(tmp_path) if f.startswith(tag + "_" + mode) and f.endswith(".dat"))
            assert len(dats) == 3, dats
            outs[tag] = [open(tmp_path / f).read() for f in dats]
            mm = re.findall(r"minmaxavg d2bimg\s+(\S+)\s+(\S+)\s+(\S+)", r.stdout)
            assert len(mm) == 3
        for a, b in zip(outs["batch"], outs["seq"]):
            assert a.splitlines()[1:] == b.splitlines()[1:]       # first line names the file
        assert outs["batch"][0].splitlines()[-5:] != outs["batch"][2].splitlines()[-5:]
        print("[tomography] %s: batch == sequential for 3 tilts" % mode, flush=True)


def _read_dat(path):
    lines = open(path).read().splitlines()
    h = next(i for i, ln in enumerate(lines) if not ln.startswith("#"))
    nx, ny = (int(v) for v in lines[h].split())
    vals = [float(ln[i:i + 12]) for ln in lines[h + 1:] for i in range(0, len(ln), 12) if ln[i:i + 12].strip()]
    return np.array(vals).reshape(ny, nx)


@needs_exe
@pytest.mark.gpu
def test_triclinic_lammps_run_through_host(tmp_path, golden_dir):
    """A triclinic LAMMPS input (the W loop sheared by S, upper
    unitriangular; tilt factors on the "xy xz yz" line): the host reads
    the tilted supercell, wraps the atoms into it and writes the image of
    the library's Python binding on the same supercell, to the .dat's 7
    decimals."""
    from simulated_tem_amd.imaging import DynamicalTwoBeamImaging, input_supercell, lattice_a0
    S = np.array([[1.0, 0.04, 0.02], [0.0, 1.0, 0.03], [0.0, 0.0, 1.0]])
    w = np.load(os.path.join(golden_dir, "w_loop_golden.npz"))
    box = w["box"][:, 1] - w["box"][:, 0]
    lo = np.zeros(3)
    A = S @ np.diag(box)
    x = (w["x"] - w["box"][:, 0]) @ S.T
    write_lammps(tmp_path / "t.lammps", x, list(zip(lo, box)), tilt=(A[0, 1], A[0, 2], A[1, 2]))
    write_xi(tmp_path / "xi.dat", w["xi"][0], w["xi"][1], (2, 0, 0))
    U = ",".join("%.6f" % v for v in w["rotation_cm"])
    d = ",".join("%.3f" % v for v in w["space"])
    r = run(["-f", "t.lammps", "-a0", "3.1652", "-lattice", "bcc", "-k", "0,0,1", "-g", "2,0,0", "-xifile", "xi.dat",
             "-U", U, "-R", "1,0,0,0,1,0,0,0,1", "-d", d, "-o", "tri", "-mode", "fp64", "-nopng"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    dats = [f for f in os.listdir(tmp_path) if f.startswith("tri") and f.endswith(".dat")]
    assert len(dats) == 1
    got = _read_dat(tmp_path / dats[0])
    a0 = lattice_a0(3.1652)
    # the supercell and positions as the file holds them
    r10 = lambda v: float("%.10f" % v)
    A_file = np.array([[r10(box[0]), 0, 0], [r10(A[0, 1]), r10(box[1]), 0], [r10(A[0, 2]), r10(A[1, 2]), r10(box[2])]]).T
    x = np.array([[float("%.16g" % v) for v in p] for p in x])
    sa, sn = input_supercell(A_file, 4, a0)
    im = DynamicalTwoBeamImaging(sa, sn, a0, (0, 0, 1), (2, 0, 0), 200.0, 4, 4)
    im.setConventionalUnitCell(a0 * np.eye(3))
    im.setVoxelSupercellSize([float(v) for v in d.split(",")])
    im.setSupercellTilt(np.array([float(v) for v in U.split(",")]).reshape(3, 3).T, np.eye(3))
    im.setExtinctionDistances(w["xi"][0], w["xi"][1])
    im.setUseDensityField(False)
    want = im.computeD2BI_img(x, mode="fp64")
    e = np.max(np.abs(got - want))
    print("[triclinic host] max |dat - image| %.3e" % e, flush=True)
    assert got.shape == want.shape
    assert e <= 6e-8

