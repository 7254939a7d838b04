"""Seeded synthetic workloads (the reference reads atom files; it has no
generator).  Atoms come from d2bi_synth_bcc (C++ in the library): a bcc
crystal with Gaussian thermal noise, Eshelby point-dilatation centres
u = A r/|r|^3 (minimum image, |u| capped at U_CAP per centre) and spherical
voids.  Conditions are set up the way d2bi's driver does it for -sg
(d2bi.f90:1145-1394): input supercell voxelisation, optional -zpad surface,
setSupercellTilt0 for the requested s_g, and tomography tilts
R = Rx(theta) R0 with setSupercellTilt2 (d2bi.f90:1352-1381).

Workloads (sizes and flags):
  c0  bcc a0=3.1652, 20^3 cells (16,000 sites), noise 0.05 A, 4 centres
      A=2 A^3, 1 void r=8 A, periodic; g=110, k=001, s_g=0, M=4,
      -d 63.304,63.304,63.304 -> 80x80x80 voxels
  c1  200x200x100 cells (8.0M sites), 256 centres A~U[1,5], 32 voids
      r~U[5,20], -zpad, g=110, s_g=0.01, -d 633.04,633.04,316.52
      -> 800x800x400 voxels
  c2  160x160x400 cells (20.5M sites), as c1 distributions, g=200,
      s_g in {-0.02,0,0.02}, -density, -d 506.432,506.432,1266.08
      -> 640x640x1600 voxels
  c3  c1 atoms, nPrecAngle=32 (0.005 rad) x nTomoAngle=21 (20 deg)
      x g in {110, 011}: 42 voxel fields, 1344 integrations
  c4  64 seeded configurations: sites log-uniform in [1e4, 4e6], box aspect
      ratios U[0.5, 2], 0-128 centres, 0-16 voids, 1-4 distinct g from
      {110, 1-10, 200, 211} (zone axis 001, or 01-1 for 211), s_g ~ U[0, 0.02], 1-8
      precession points, M in {3, 4, 6}, image space = the periodic box
"""
import ctypes as C
import math

import numpy as np

from . import abi
from .abi import D2biSynthParams, check, ptr
from .imaging import DynamicalTwoBeamImaging, input_supercell, lattice_a0

U_CAP = 0.5                     # cap on |u| from one Eshelby centre (A)
THETA_MAX = 5.0                 # Lib_DiffractionConditions.f90:26 (deg)
PI = 3.141592653589790

CONFIGS = {
    "c0": dict(a0=3.1652, cells=(20, 20, 20), noise=0.05, n_centres=4, A=(2.0, 2.0),
               n_voids=1, r=(8.0, 8.0), zpad=False, g=[(1, 1, 0)], k=(0, 0, 1), sg=[0.0],
               xi0=103.9, xig=207.5, V=200.0, M=4, d=(63.304, 63.304, 63.304), density=False,
               n_prec=1, prec_angle=0.005, n_tomo=1, tomo_angle=20.0, seed=20240401),
    "c1": dict(a0=3.1652, cells=(200, 200, 100), noise=0.05, n_centres=256, A=(1.0, 5.0),
               n_voids=32, r=(5.0, 20.0), zpad=True, g=[(1, 1, 0)], k=(0, 0, 1), sg=[0.01],
               xi0=103.9, xig=207.5, V=200.0, M=4, d=(633.04, 633.04, 316.52), density=False,
               n_prec=1, prec_angle=0.005, n_tomo=1, tomo_angle=20.0, seed=20240402),
    "c2": dict(a0=3.1652, cells=(160, 160, 400), noise=0.05, n_centres=256, A=(1.0, 5.0),
               n_voids=32, r=(5.0, 20.0), zpad=False, g=[(2, 0, 0)], k=(0, 0, 1),
               sg=[-0.02, 0.0, 0.02], xi0=103.9, xig=207.5, V=200.0, M=4,
               d=(506.432, 506.432, 1266.08), density=True, n_prec=1, prec_angle=0.005,
               n_tomo=1, tomo_angle=20.0, seed=20240403),
}
CONFIGS["c3"] = dict(CONFIGS["c1"], g=[(1, 1, 0), (0, 1, 1)], n_prec=32, prec_angle=0.005,
                     n_tomo=21, tomo_angle=20.0)


def c4_configs(seed=20240404, n=64, max_sites=4.0e6):
    """The uneven batch: a list of workload dicts (each possibly with several
    g-vectors)."""
    rng = np.random.default_rng(seed)
    out = []
    gset = [((1, 1, 0), (0, 0, 1)), ((1, -1, 0), (0, 0, 1)), ((2, 0, 0), (0, 0, 1)),
            ((2, 1, 1), (0, 1, -1))]
    for i in range(n):
        sites = math.exp(rng.uniform(math.log(1.0e4), math.log(max_sites)))
        ar = rng.uniform(0.5, 2.0, size=2)
        cells3 = sites / 2.0
        nz = max(3, int(round((cells3 / (ar[0] * ar[1])) ** (1.0 / 3.0))))
        nx = max(3, int(round(ar[0] * nz)))
        ny = max(3, int(round(ar[1] * nz)))
        ng = int(rng.integers(1, 5))
        picks = [gset[j] for j in rng.choice(4, size=ng, replace=False)]
        a0 = 3.1652
        box = (a0 * nx, a0 * ny, a0 * nz)
        out.append(dict(a0=a0, cells=(nx, ny, nz), noise=0.05, n_centres=int(rng.integers(0, 129)),
                        A=(1.0, 5.0), n_voids=int(rng.integers(0, 17)), r=(5.0, 20.0), zpad=False,
                        gk=picks, sg=[float(rng.uniform(0.0, 0.02))], xi0=103.9, xig=207.5, V=200.0,
                        M=int(rng.choice([3, 4, 6])), d=box, density=bool(rng.integers(0, 2)),
                        n_prec=int(rng.integers(1, 9)), prec_angle=0.005, n_tomo=1, tomo_angle=20.0,
                        seed=seed * 1000 + i))
    return out


def c4_conditions(cfgs):
    """Per configuration: (x, [(imaging, precAngle, nPrec), ...])."""
    out = []
    for cfg in cfgs:
        x, box = make_atoms(cfg)
        conds = []
        for g, k in cfg["gk"]:
            c = dict(cfg, g=[g], k=k)
            x2, cc = conditions(c, x, box)
            conds.extend(cc)
        out.append((x2, conds))
    return out


def scaled(name, cells=None, d=None):
    """A copy of a workload with a different crystal size (tests use small
    versions of the large workloads)."""
    cfg = dict(CONFIGS[name])
    if cells is not None:
        cfg["cells"] = tuple(cells)
        f = [c / c0 for c, c0 in zip(cells, CONFIGS[name]["cells"])]
        cfg["d"] = tuple(dd * ff for dd, ff in zip(CONFIGS[name]["d"], f)) if d is None else d
        vol = f[0] * f[1] * f[2]
        cfg["n_centres"] = max(1, int(round(cfg["n_centres"] * vol)))
        cfg["n_voids"] = max(0, int(round(cfg["n_voids"] * vol)))
    return cfg


def matmul3(A, B):
    """Fortran matmul of 3x3 matrices (k ascending, no fused operations)."""
    Cm = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            Cm[i, j] = (A[i, 0] * B[0, j] + A[i, 1] * B[1, j]) + A[i, 2] * B[2, j]
    return Cm


def make_atoms(cfg):
    """(x (nAtoms,3), box lengths (3,)) for a workload."""
    p = D2biSynthParams()
    p.a0 = cfg["a0"]
    p.nx, p.ny, p.nz = cfg["cells"]
    p.n_centres = cfg["n_centres"]
    p.A_min, p.A_max = cfg["A"]
    p.u_cap = U_CAP
    p.n_voids = cfg["n_voids"]
    p.r_min, p.r_max = cfg["r"]
    p.noise = cfg["noise"]
    p.offset = (C.c_double * 3)(0.0, 0.0, 0.0)
    p.seed = cfg["seed"]
    cap = 2 * p.nx * p.ny * p.nz
    x = np.zeros((cap, 3))
    n = C.c_int64()
    box = np.zeros(3)
    check(abi.load().d2bi_synth_bcc(C.byref(p), ptr(x), cap, C.byref(n), ptr(box)))
    return x[:n.value].copy(), box


def apply_zpad(x, a_super, d0, theta):
    """-zpad (d2bi.f90:1052-1068): stretch the supercell along z and centre
    the atoms so the foil has free surfaces; returns (x, a_super, d)."""
    x = x.copy()
    a_super = a_super.copy()
    d = list(d0)
    len1 = float(np.linalg.norm(a_super[2, :]))
    len2 = max(float(np.linalg.norm(a_super[0, :])), float(np.linalg.norm(a_super[1, :])))
    if d[0] is not None:
        len2 = max(len2, d[0] / math.cos(theta))
    if d[1] is not None:
        len2 = max(len2, d[1] / math.cos(theta))
    pad = len1 * math.cos(theta) + len2 * math.sin(theta)
    x[:, 2] = x[:, 2] + (pad - len1) * float(np.float32(1.05)) / 2     # 1.05 is a default real
    a_super[2, :] = a_super[2, :] * (pad * 1.05 / len1)
    if d[2] is None or d[2] == 0:
        d[2] = pad
    return x, a_super, d


def conditions(cfg, x, box):
    """The imaging conditions of a workload as d2bi would set them up:
    returns (x, list of (DynamicalTwoBeamImaging, precAngle, nPrec))."""
    a0 = lattice_a0(cfg["a0"])
    a_super = np.diag(np.asarray(box, dtype=np.float64))
    d = list(cfg["d"])
    if cfg["zpad"]:
        if cfg["n_tomo"] > 1:
            theta = (cfg["tomo_angle"] + THETA_MAX) * PI / 180.0 + cfg["prec_angle"]
        else:
            theta = THETA_MAX * PI / 180.0 + cfg["prec_angle"]
        x, a_super, d = apply_zpad(x, a_super, d, theta)
    super_a, super_n = input_supercell(a_super, cfg["M"], a0)
    out = []
    for g in cfg["g"]:
        for sg in cfg["sg"]:
            im = DynamicalTwoBeamImaging(super_a, super_n, a0, cfg["k"], g, cfg["V"], cfg["M"],
                                         cfg["M"])
            im.setConventionalUnitCell(a0 * np.eye(3))
            im.setVoxelSupercellSize(d)
            im.setDeviationParameter(sg)
            im.setSupercellTilt()
            im.setExtinctionDistances(cfg["xi0"], cfg["xig"])
            im.setUseDensityField(cfg["density"])
            if cfg["n_tomo"] <= 1:
                out.append((im, cfg["prec_angle"], cfg["n_prec"]))
                continue
            U0, R0 = im.getRotation()
            nt = cfg["n_tomo"]
            for n in range(nt):
                th = cfg["tomo_angle"] - 2 * cfg["tomo_angle"] * n / (nt - 1)
                th = th * 3.141592654 / 180.0
                Rx = np.array([[1.0, 0.0, 0.0], [0.0, math.cos(th), -math.sin(th)],
                               [0.0, math.sin(th), math.cos(th)]])
                imt = DynamicalTwoBeamImaging(super_a, super_n, a0, cfg["k"], g, cfg["V"],
                                              cfg["M"], cfg["M"])
                imt.setConventionalUnitCell(a0 * np.eye(3))
                imt.setVoxelSupercellSize(d)
                imt.setSupercellTilt(U0, matmul3(Rx, R0))
                imt.setExtinctionDistances(cfg["xi0"], cfg["xig"])
                imt.setUseDensityField(cfg["density"])
                out.append((imt, cfg["prec_angle"], cfg["n_prec"]))
    return x, out
