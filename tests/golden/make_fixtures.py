"""Extract the reference's own test data into small fixtures.

Run in the build container only (it reads the reference checkout as text):

    python tests/golden/make_fixtures.py /root/reference

Writes, next to this script:
  linkcell_kat.json     the 200 positions and expected outputs of
                        test/testLib_LinkCell3D.f90:66-290
  w_loop_golden.npz     Data/test/test.i111_loop_W.lammps atom positions and
                        box, the golden image + header of
                        Data/test/...g2_0_0.k0.00_0.00_1.00_ng1.00.dat, and
                        the 16-bit grey levels of the golden
                        Data/test/...g2_0_0.k0.00_0.00_1.00_ng1.00.png
  aperture_points.json  the table of n = 1..25 points in the unit disk that
                        evenlySpacedPointsInCircle returns (src/evenlySpaced.h),
                        as the decimal literals written there
The fixtures are data (inputs and expected outputs); no reference source
text is stored.
"""
import json
import os
import re
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def linkcell_kat(ref):
    src = open(os.path.join(ref, "test", "testLib_LinkCell3D.f90")).read()
    # positions: the reshape( (/ ... /),(/ 3,200 /) ) literal
    m = re.search(r"xx\s*=\s*reshape\(\s*\(/(.*?)/\)\s*,\s*\(/\s*3\s*,\s*200\s*/\)", src, re.S)
    nums = [float(v) for v in re.findall(r"([-+]?\d+\.\d+)d0", m.group(1))]
    assert len(nums) == 600, len(nums)
    pos = np.array(nums).reshape(200, 3)
    # expected output lines: the output0 array of strings
    m = re.search(r"output0\s*=\s*\(/(.*?)/\)", src, re.S)
    lines = [s.strip() for s in re.findall(r'"([^"]*)"', m.group(1))]
    assert len(lines) == 20, len(lines)
    neigh = []
    for ln in lines[6:19]:
        t = ln.split()
        # neigh <n> id <id> dx <dx1> <dx2> <dx3> |dx| <d>
        neigh.append({"id": int(t[3]), "dx": [float(t[5]), float(t[6]), float(t[7])],
                      "d": float(t[9])})
    exc_self = [int(v) for v in lines[19].split()[3:]]
    return {
        "source": "test/testLib_LinkCell3D.f90:66-290",
        "a_cell": [[1, 0, 0], [0, 1, 0], [0, 0, 1]],
        "N": [6, 4, 5],
        "positions": pos.tolist(),
        "n_atoms": int(lines[0].split()[-1]),
        "max_count": int(lines[1].split()[-1]),
        "max_storage": int(lines[2].split()[-1]),
        "max_neigh": int(lines[3].split()[-1]),
        "atom100": [float(v) for v in lines[4].split()[3:6]],
        "nearest": {"x": [float(v) for v in lines[5].split()[2:5]],
                    "dx": [float(v) for v in lines[5].split()[6:9]],
                    "d": float(lines[5].split()[10])},
        "neigh_r1": neigh,
        "neigh_exc_self": exc_self,
        "expected_lines": lines,
    }


def w_loop(ref):
    lam = os.path.join(ref, "Data", "test", "test.i111_loop_W.lammps")
    lines = open(lam).read().splitlines()
    n_atoms = int(lines[2].split()[0])
    box = []
    for ln in lines[:12]:
        for ax in ("xlo", "ylo", "zlo"):
            if ax in ln:
                lo, hi = (float(v) for v in ln.split()[:2])
                box.append((lo, hi))
    start = next(i for i, ln in enumerate(lines) if ln.startswith("Atoms")) + 2
    xyz = np.array([[float(v) for v in lines[start + i].split()[2:5]] for i in range(n_atoms)])
    ids = np.array([int(lines[start + i].split()[0]) for i in range(n_atoms)], np.int32)
    dat = os.path.join(ref, "Data", "test", "test.i111_loop_W.lammps.g2_0_0.k0.00_0.00_1.00_ng1.00.dat")
    dl = open(dat).read().splitlines()
    hdr = {}
    for ln in dl[:11]:
        key = ln[1:].split()[0] if ln.startswith("#") else None
        hdr[ln] = key
    g_line = dl[2].split()
    xi = [float(v) for v in dl[5].split()[3:5]]
    sg = float(dl[7].split()[-1])
    space = [float(v) for v in dl[8].split()[-3:]]
    scale = float(dl[9].split()[-1])
    rot = [float(v) for v in dl[10].split()[2:11]]
    nx, ny = (int(v) for v in dl[11].split())
    vals = []
    for ln in dl[12:]:
        vals.extend(float(ln[i:i + 12]) for i in range(0, len(ln), 12) if ln[i:i + 12].strip())
    img = np.array(vals).reshape(ny, nx)        # written x fastest: row = iy
    png = read_png16(dat[:-4] + ".png")
    assert png.shape == img.shape
    return dict(
        x=xyz, ids=ids, box=np.array(box), modg=float(g_line[-1]), xi=np.array(xi), sg=sg,
        space=np.array(space), scale=scale, rotation_cm=np.array(rot), image=img,
        header="\n".join(dl[:11]), png=png.astype(np.uint16))


def read_png16(path):
    """Grey levels of a 16-bit greyscale, non-interlaced PNG as an (h, w)
    array of int64 (row 0 first): the IDAT stream inflated and each row's
    filter (none, sub, up, average, Paeth; 2 bytes per pixel) undone, as
    the PNG specification defines them."""
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat = 8, b""
    while pos < len(data):
        n = int.from_bytes(data[pos:pos + 4], "big")
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        if typ == b"IHDR":
            w, h = int.from_bytes(body[:4], "big"), int.from_bytes(body[4:8], "big")
            assert body[8] == 16 and body[9] == 0 and body[12] == 0, "16-bit grey, not interlaced"
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    stride, bpp = 2 * w, 2
    prev = [0] * stride
    rows = []
    for y in range(h):
        f = raw[y * (stride + 1)]
        line = raw[y * (stride + 1) + 1:(y + 1) * (stride + 1)]
        cur = [0] * stride
        for i in range(stride):
            a = cur[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            if f == 0:
                p = 0
            elif f == 1:
                p = a
            elif f == 2:
                p = b
            elif f == 3:
                p = (a + b) // 2
            else:
                pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            cur[i] = (line[i] + p) & 255
        rows.append(cur)
        prev = cur
    v = np.array(rows, np.int64).reshape(h, w, 2)
    return v[:, :, 0] * 256 + v[:, :, 1]


def aperture_points(ref):
    """xKnown(2, 25, 25) of src/evenlySpaced.h: the literal list fills x(1:2,
    point, n) first index fastest; point k of the n-point set for k <= n.
    The literals have no kind suffix (default real), so they are kept as the
    decimal strings written there."""
    src = open(os.path.join(ref, "src", "evenlySpaced.h")).read()
    maxk = int(re.search(r"MAXKNOWN\s*=\s*(\d+)", src).group(1))
    m = re.search(r"reshape\(\s*\(/(.*?)/\)\s*,\s*\(/\s*2\s*,\s*MAXKNOWN\s*,\s*MAXKNOWN\s*/\)", src, re.S)
    lits = re.findall(r"[-+]?\d+\.\d+", m.group(1))
    assert len(lits) == 2 * maxk * maxk, len(lits)
    pts = {}
    for n in range(1, maxk + 1):
        base = 2 * maxk * (n - 1)
        pts[str(n)] = [[lits[base + 2 * k], lits[base + 2 * k + 1]] for k in range(n)]
    return {"source": "src/evenlySpaced.h (xKnown, MAXKNOWN = %d)" % maxk,
            "kind": "default real (FP32) literals", "points": pts}


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(HERE, "linkcell_kat.json"), "w") as f:
        json.dump(linkcell_kat(ref), f, indent=1)
    w = w_loop(ref)
    np.savez_compressed(os.path.join(HERE, "w_loop_golden.npz"), **w)
    with open(os.path.join(HERE, "aperture_points.json"), "w") as f:
        json.dump(aperture_points(ref), f, indent=0)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
