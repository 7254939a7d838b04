"""Python mirror of the reference's imaging interface, on the HIP library.

  DynamicalTwoBeamImaging   type + procedures of Lib_DynamicalTwoBeamImaging.f90
                            (ctor :156, setters :245-345, setSupercellTilt0/2
                            :567/:1048, findImageParameters :450,
                            computeD2BI_img :1563)
  LinkCell3D                LinkCell3D_ctor0 / add1 / neighbourList1b
                            (Lib_LinkCell3D.f90:192, 680, 1515)
  gaussianBlurImage         d2bi.f90:1543
  input_supercell           the voxelised input supercell d2bi builds from
                            the file's box (d2bi.f90:920-929, 986-992)

Images are returned as (My, Mx) arrays: img[iy, ix] is d2big2(ix, iy), the
memory order of the reference's column-major array.  Every compute call
runs on the GPU through libd2bi_gpu.so and raises if the library is not
built or no device is visible.
"""
import ctypes as C
import math

import numpy as np

from . import abi
from .abi import (D2biRecord, D2biPlan, D2biStats, D2BI_MODE_FP64, D2BI_MODE_FAST, check,
                  f64, mat_cm, mat_from_cm, ptr, iptr)

MODES = {"fp64": D2BI_MODE_FP64, "fast": D2BI_MODE_FAST, D2BI_MODE_FP64: D2BI_MODE_FP64,
         D2BI_MODE_FAST: D2BI_MODE_FAST}

_device = None


def init(device=0):
    """Select the HIP device (d2bi_gpu_init)."""
    global _device
    L = abi.load()
    if _device != device:
        check(L.d2bi_gpu_init(int(device)))
        _device = device


def finalize():
    global _device
    if _device is not None:
        abi.load().d2bi_gpu_finalize()
        _device = None


def _ensure():
    if _device is None:
        init(0)


class Atoms:
    """Atom positions resident in HBM (d2bi_gpu_atoms_upload)."""

    def __init__(self, x):
        _ensure()
        x = f64(x)
        if x.ndim != 2 or x.shape[1] != 3:
            raise ValueError("x must be (nAtoms, 3)")
        h = C.c_void_p()
        check(abi.load().d2bi_gpu_atoms_upload(ptr(x), 3, len(x), C.byref(h)))
        self.h = h
        self.n = len(x)

    def __del__(self):
        if getattr(self, "h", None) and _device is not None:
            abi.load().d2bi_gpu_atoms_free(self.h)
            self.h = None


class DeviceBuffer:
    """A float64 array in HBM, allocated with the HIP runtime the library
    itself links (a second HIP runtime in the process, such as the one
    bundled with a torch wheel, could not open the device).  Used to keep an
    image in HBM (computeD2BI_img(..., out=buffer))."""

    def __init__(self, shape):
        _ensure()
        self.shape = tuple(int(v) for v in shape)
        self.nbytes = 8 * int(np.prod(self.shape))
        self._hip = _hip_runtime()
        p = C.c_void_p()
        rc = self._hip.hipMalloc(C.byref(p), C.c_size_t(max(self.nbytes, 8)))
        if rc != 0:
            raise MemoryError("hipMalloc failed (%d)" % rc)
        self.ptr = p.value

    def numpy(self):
        """Copy to a host array."""
        out = np.empty(self.shape)
        rc = self._hip.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(self.ptr), C.c_size_t(self.nbytes), 2)
        if rc != 0:
            raise RuntimeError("hipMemcpy failed (%d)" % rc)
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and _device is not None:
            self._hip.hipFree(C.c_void_p(self.ptr))
            self.ptr = None


_hip = None


def _hip_runtime():
    """libamdhip64 as already loaded by libd2bi_gpu.so."""
    global _hip
    if _hip is None:
        abi.load()
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _hip


def input_supercell(a_super, M, a0):
    """Voxelised input supercell of d2bi (d2bi.f90:986-992): N_i =
    max(3, floor(M * |A_i| / a0)), cell vectors A_i / N_i."""
    a_super = np.asarray(a_super, dtype=np.float64)
    n = []
    for i in range(3):
        col = a_super[:, i]
        n.append(max(3, int(math.floor(M * float(np.sqrt(np.dot(col, col))) / a0))))
    a = np.stack([a_super[:, i] / n[i] for i in range(3)], axis=1)
    return a, n


def lattice_a0(a0in):
    """a0 of a cubic lattice from the -a0 values (d2bi.f90:976-983)."""
    a = list(a0in) if np.ndim(a0in) else [a0in] * 3
    return ((a[0] + a[1]) + a[2]) / 3


# Lib_Lattices.f90:65-69
LATTICE_CODES = {"sc": 1, "bcc": 2, "fcc": 3, "hcp": 4}


class DynamicalTwoBeamImaging:
    """DynamicalTwoBeamImaging_ctor0(super,latt,a0,k,g,E,mxy,mzz) and its
    setters.  super is given as (cell vectors a[3x3] with columns as vectors,
    repeats n[3]), as SimpleSupercell_ctor(a,Nx,Ny,Nz) holds it; latt is the
    lattice name (it only matters to the (g,ng) tilt search)."""

    def __init__(self, super_a, super_n, a0, k, g, E=200.0, mxy=4, mzz=4, latt="bcc"):
        self.rec = D2biRecord()
        r = self.rec
        r.a0 = a0
        r.super_a = mat_cm(super_a)
        r.super_n = (C.c_int32 * 3)(*[int(v) for v in super_n])
        r.g = (C.c_double * 3)(*[float(v) for v in g])
        r.E = E
        r.xi0, r.xig = 103.9, 207.5            # DynamicalTwoBeamImaging_null :133
        r.U = mat_cm(np.eye(3))
        r.R = mat_cm(np.eye(3))
        r.use_density = 1                      # null ctor sets .true. (:134); d2bi overrides
        r.a_cell = mat_cm(a0 * np.eye(3))
        r.nuvw = 1
        self.k = np.asarray(k, dtype=np.float64)
        self.mxy, self.mzz = int(mxy), int(mzz)
        self.sg_target = float("nan")
        self.d = np.zeros(3)
        self.latt = LATTICE_CODES.get(str(latt).lower(), 1)

    # --- setters (Lib_DynamicalTwoBeamImaging.f90:245-345)
    def setConventionalUnitCell(self, a_cell):
        self.rec.a_cell = mat_cm(a_cell)

    def setVoxelSupercellSize(self, d):
        d = np.asarray(d, dtype=np.float64)
        if np.any(d <= 0):
            raise NotImplementedError("voxel cuboid sides must be given (> 0)")
        self.d = d

    def setDeviationParameter(self, sg):
        self.sg_target = float(sg)

    def setExtinctionDistances(self, xi0, xig):
        self.rec.xi0, self.rec.xig = float(xi0), float(xig)

    def setUseDensityField(self, use):
        self.rec.use_density = 1 if use else 0

    def setSupercellTilt(self, U=None, R=None, ng=None, extinction_distances_filename=None):
        """setSupercellTilt0 (no arguments: tilt for the deviation parameter),
        setSupercellTilt2(U,R) (previously determined rotation) or
        setSupercellTilt1a(ng, extinction_distances_filename) (the (g,ng)
        tilt search, without atom positions).  The search uses the
        extinction distances the record holds at the time (the d2bi driver
        sets the table values only afterwards, d2bi.f90:1265)."""
        L = abi.load()
        d = (C.c_double * 3)(*self.d)
        if ng is not None:
            k = (C.c_double * 3)(*self.k)
            check(L.d2bi_setup_tilt_ng(C.byref(self.rec), k, float(ng), self.latt,
                                       str(extinction_distances_filename).encode(), self.mxy, self.mzz, d))
            self.d = np.array(list(d))
        elif U is None and R is None:
            k = (C.c_double * 3)(*self.k)
            check(L.d2bi_setup_tilt_sg(C.byref(self.rec), k, self.sg_target, self.mxy, self.mzz, d))
        else:
            check(L.d2bi_setup_tilt_ur(C.byref(self.rec), mat_cm(U), mat_cm(R), self.mxy, self.mzz, d))

    # --- getters
    def getDeviationParameter(self):
        return self.rec.sg

    def getRotation(self):
        return mat_from_cm(self.rec.U), mat_from_cm(self.rec.R)

    def plan(self, n_atoms, prec_angle=0.0, n_prec=1):
        p = D2biPlan()
        check(abi.load().d2bi_plan_image(C.byref(self.rec), int(n_atoms), float(prec_angle),
                                         int(n_prec), C.byref(p)))
        return p

    def aperture(self, precessionAngle, nPrecessionAngle, apertureAngle, nAperturePoints):
        """(gshift (nApp,3), aa (nPrec, nApp)): the g-vector for g.r at each
        aperture point and 2 pi s_g per (precession point, aperture point)
        (Lib_DynamicalTwoBeamImaging.f90:1655-1667, 1805-1836)."""
        gs = np.zeros(3 * nAperturePoints)
        aa = np.zeros(nAperturePoints * nPrecessionAngle)
        check(abi.load().d2bi_aperture_setup(C.byref(self.rec), float(precessionAngle),
                                             int(nPrecessionAngle), float(apertureAngle),
                                             int(nAperturePoints), ptr(gs), ptr(aa)))
        return gs.reshape(nAperturePoints, 3), aa.reshape(nPrecessionAngle, nAperturePoints)

    def findImageParameters(self):
        """(aperpx, xmax[3], modg, Mx, My, Mz) (:450-482)."""
        p = self.plan(1)
        return p.aperpx, np.array(list(p.xmax)), p.modg, p.Mx, p.My, p.Mz

    # --- the hot path
    def computeD2BI_img(self, x, precessionAngle=0.0, nPrecessionAngle=1, mode="fast",
                        stats=None, out=None, apertureAngle=None, nAperturePoints=None):
        """Image of one condition, (My, Mx) array.  x: (nAtoms,3) array or Atoms.
        out: optional DeviceBuffer of shape (My, Mx) that receives the image
        in HBM (x must then be Atoms); it is returned.
        apertureAngle (radians) and nAperturePoints are computeD2BI_img's
        optional aperture sampling (Lib_DynamicalTwoBeamImaging.f90:1563,
        1655-1667): the image is then the mean over the aperture points,
        each with its own g.r (x must be an (nAtoms,3) array)."""
        _ensure()
        L = abi.load()
        p = self.plan(1)
        st = stats if stats is not None else D2biStats()
        m = MODES[mode]
        if nAperturePoints is not None:
            if out is not None or isinstance(x, Atoms):
                raise TypeError("aperture sampling takes host atom arrays and returns a host image")
            img = np.zeros((p.My, p.Mx))
            x = f64(x)
            check(L.d2bi_gpu_image_aperture(C.byref(self.rec), ptr(x), 3, len(x), float(precessionAngle),
                                            int(nPrecessionAngle),
                                            float(apertureAngle if apertureAngle is not None else 0.0),
                                            int(nAperturePoints), m, ptr(img), C.byref(st)))
            return img
        if out is not None:
            if not isinstance(x, Atoms):
                raise TypeError("a device output needs device-resident Atoms")
            if not isinstance(out, DeviceBuffer) or out.shape != (p.My, p.Mx):
                raise ValueError("out must be a DeviceBuffer of shape (%d, %d)" % (p.My, p.Mx))
            check(L.d2bi_gpu_image_dev_out(C.byref(self.rec), x.h, float(precessionAngle),
                                           int(nPrecessionAngle), m, C.c_void_p(out.ptr), C.byref(st)))
            return out
        img = np.zeros((p.My, p.Mx))
        if isinstance(x, Atoms):
            check(L.d2bi_gpu_image_dev(C.byref(self.rec), x.h, float(precessionAngle),
                                       int(nPrecessionAngle), m, ptr(img), C.byref(st)))
        else:
            x = f64(x)
            check(L.d2bi_gpu_image(C.byref(self.rec), ptr(x), 3, len(x), float(precessionAngle),
                                   int(nPrecessionAngle), m, ptr(img), C.byref(st)))
        return img

    def field(self, x, mode="fp64", diag=True):
        """Voxel field for tests: (phase[Mz,My,Mx] complex, rho, count, r2min).
        diag=False (FAST) takes the field from the imaging kernel itself and
        returns None for count and r2min."""
        _ensure()
        p = self.plan(len(x))
        n = p.Mx * p.My * p.Mz
        ph = np.zeros(2 * n)
        rho = np.zeros(n)
        cnt = np.zeros(n, np.int32) if diag else None
        r2m = np.zeros(n) if diag else None
        x = f64(x)
        check(abi.load().d2bi_gpu_field(C.byref(self.rec), ptr(x), 3, len(x), MODES[mode], ptr(ph),
                                        ptr(rho), iptr(cnt) if diag else None,
                                        ptr(r2m) if diag else None))
        shp = (p.Mz, p.My, p.Mx)
        return ((ph[0::2] + 1j * ph[1::2]).reshape(shp), rho.reshape(shp),
                cnt.reshape(shp) if diag else None, r2m.reshape(shp) if diag else None)

    def linkcells(self, x):
        """findGdotR's link cells: (cell_start, ids, frac, gdotr)."""
        _ensure()
        x = f64(x)
        p = self.plan(len(x))
        ncell = p.Ex * p.Ey * p.Ez
        cs = np.zeros(ncell + 1, np.int32)
        n = C.c_int64()
        L = abi.load()
        check(L.d2bi_gpu_linkcells(C.byref(self.rec), ptr(x), 3, len(x), iptr(cs), None, None,
                                   None, 0, C.byref(n)))
        cap = max(1, n.value)
        ids = np.zeros(cap, np.int32)
        frac = np.zeros(3 * cap)
        gd = np.zeros(cap)
        check(L.d2bi_gpu_linkcells(C.byref(self.rec), ptr(x), 3, len(x), iptr(cs), iptr(ids),
                                   ptr(frac), ptr(gd), cap, C.byref(n)))
        k = n.value
        return cs, ids[:k], frac[:3 * k].reshape(k, 3), gd[:k]


def batch_image_size(conditions):
    """Doubles the images of a batch occupy (sum of Mx*My)."""
    total = 0
    for c in conditions:
        p = c[0].plan(1)
        total += p.Mx * p.My
    return total


def computeD2BI_img_batch(conditions, mode="fast", stats=None, out=None):
    """Batched conditions: list of (DynamicalTwoBeamImaging, Atoms, precAngle,
    nPrec).  Returns the list of images; with out (a DeviceBuffer of
    batch_image_size(conditions) doubles) the images stay in HBM, one after
    another, and out is returned."""
    _ensure()
    n = len(conditions)
    recs = (D2biRecord * n)(*[c[0].rec for c in conditions])
    hs = (C.c_void_p * n)(*[c[1].h for c in conditions])
    pa = np.array([c[2] for c in conditions], dtype=np.float64)
    npr = np.array([c[3] for c in conditions], dtype=np.int32)
    shapes = []
    for c in conditions:
        p = c[0].plan(1)
        shapes.append((p.My, p.Mx))
    st = stats if stats is not None else D2biStats()
    if out is not None:
        total = sum(a * b for a, b in shapes)
        if not isinstance(out, DeviceBuffer) or int(np.prod(out.shape)) != total:
            raise ValueError("out must be a DeviceBuffer of %d doubles" % total)
        check(abi.load().d2bi_gpu_image_batch_dev_out(recs, hs, n, ptr(pa), iptr(npr), MODES[mode],
                                                      C.c_void_p(out.ptr), C.byref(st)))
        return out
    img = np.zeros(sum(a * b for a, b in shapes))
    check(abi.load().d2bi_gpu_image_batch(recs, hs, n, ptr(pa), iptr(npr), MODES[mode], ptr(img),
                                          C.byref(st)))
    out, off = [], 0
    for a, b in shapes:
        out.append(img[off:off + a * b].reshape(a, b))
        off += a * b
    return out


class LinkCell3D:
    """LinkCell3D_ctor0(a,nMax,Nx,Ny,Nz) + add1(x) + neighbourList1b, on the GPU.
    Points get ids 1..n in the order given."""

    def __init__(self, a, Nx, Ny, Nz):
        self.a = np.asarray(a, dtype=np.float64)
        self.N = (int(Nx), int(Ny), int(Nz))
        self.x = np.zeros((0, 3))

    def add(self, x):
        self.x = np.concatenate([self.x, np.asarray(x, dtype=np.float64).reshape(-1, 3)])

    def neighbourList(self, q, rc, max_neigh=512):
        """For each query point: (ids, r2) in the reference's traversal order."""
        _ensure()
        q = f64(np.asarray(q, dtype=np.float64).reshape(-1, 3))
        nq = len(q)
        n = np.zeros(nq, np.int32)
        ids = np.zeros(max(1, nq * max_neigh), np.int32)
        r2 = np.zeros(max(1, nq * max_neigh))
        p = f64(self.x)
        check(abi.load().d2bi_gpu_lc_query(mat_cm(self.a), *self.N, ptr(p), len(p), ptr(q), nq,
                                           float(rc), int(max_neigh), iptr(n), iptr(ids), ptr(r2)))
        if np.any(n > max_neigh):
            raise abi.D2biError("neighbour list longer than max_neigh")
        return [(ids[i * max_neigh:i * max_neigh + n[i]].copy(),
                 r2[i * max_neigh:i * max_neigh + n[i]].copy()) for i in range(nq)]


def gaussianBlurImage(f_in, t):
    """d2bi.f90:1543-1624 on the GPU; f_in is (Ny, Nx)."""
    _ensure()
    f = f64(f_in)
    ny, nx = f.shape
    out = np.zeros_like(f)
    check(abi.load().d2bi_gpu_blur(ptr(f), nx, ny, float(t), ptr(out)))
    return out


def evenlySpacedPointsInCircle(n):
    """n points in the unit disk, (n, 2): the reference's known arrangements
    for n <= 25, the sunflower placement beyond
    (Lib_EvenlySpacedPointsInCircle.f90:29-100)."""
    xy = np.zeros(2 * int(n))
    check(abi.load().d2bi_aperture_points(int(n), ptr(xy)))
    return xy.reshape(int(n), 2)


def wavelength(V):
    return abi.load().d2bi_wavelength(float(V))


def velocity(V):
    return abi.load().d2bi_velocity(float(V))
