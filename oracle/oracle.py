"""ctypes front end of the FP64 CPU oracle (oracle/d2bi_oracle.c).

Test infrastructure: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker of the HIP library.
The product package never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libd2bi_oracle.so")
_lib = None

# The record/plan layouts are shared with include/d2bi_gpu.h; the product
# package defines the same ctypes structures, imported here to avoid drift.
from simulated_tem_amd.abi import D2biRecord, D2biPlan  # noqa: E402

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_wavelength.restype = C.c_double
        L.orc_wavelength.argtypes = [C.c_double]
        L.orc_velocity.restype = C.c_double
        L.orc_velocity.argtypes = [C.c_double]
        L.orc_det3.restype = C.c_double
        L.orc_det3.argtypes = [_dp]
        L.orc_inverse3.argtypes = [_dp, _dp]
        L.orc_recip.argtypes = [_dp, _dp]
        L.orc_deviation_parameter.restype = C.c_double
        L.orc_deviation_parameter.argtypes = [C.c_double, _dp]
        L.orc_estimate_max_count.restype = C.c_int
        L.orc_estimate_max_count.argtypes = [C.c_double, C.c_int]
        L.orc_plan.restype = C.c_int
        L.orc_plan.argtypes = [C.POINTER(D2biRecord), C.c_int64, C.c_double, C.c_int,
                               C.POINTER(D2biPlan)]
        L.orc_lc_new.restype = C.c_void_p
        L.orc_lc_new.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_lc_free.argtypes = [C.c_void_p]
        L.orc_lc_clear.argtypes = [C.c_void_p]
        L.orc_lc_add.restype = C.c_int
        L.orc_lc_add.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_lc_neighbours.restype = C.c_int
        L.orc_lc_neighbours.argtypes = [C.c_void_p, _dp, C.c_double, _ip, _dp, _dp]
        L.orc_lc_count.restype = C.c_int
        L.orc_lc_count.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.orc_lc_storage.restype = C.c_int
        L.orc_lc_storage.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.orc_lc_nmax.restype = C.c_int
        L.orc_lc_nmax.argtypes = [C.c_void_p]
        L.orc_lc_flatten.restype = C.c_int64
        L.orc_lc_flatten.argtypes = [C.c_void_p, _ip, _ip, _dp]
        L.orc_rho.restype = C.c_double
        L.orc_rho.argtypes = [C.c_double, C.c_double]
        L.orc_integrate_column.restype = C.c_double
        L.orc_integrate_column.argtypes = [C.c_double] * 4 + [_dp, _dp, _dp, C.c_int]
        L.orc_intensity_at_depth.restype = C.c_double
        L.orc_intensity_at_depth.argtypes = [C.c_double] * 4
        L.orc_prepare.restype = C.c_void_p
        L.orc_prepare.argtypes = [C.POINTER(D2biRecord), _dp, C.c_int64, C.c_int64, C.c_double,
                                  C.c_int, _ip]
        L.orc_state_free.argtypes = [C.c_void_p]
        L.orc_state_npoints.restype = C.c_int64
        L.orc_state_npoints.argtypes = [C.c_void_p]
        L.orc_state_lc.restype = C.c_void_p
        L.orc_state_lc.argtypes = [C.c_void_p]
        L.orc_state_gdotr.restype = _dp
        L.orc_state_gdotr.argtypes = [C.c_void_p]
        L.orc_state_plan.restype = C.POINTER(D2biPlan)
        L.orc_state_plan.argtypes = [C.c_void_p]
        L.orc_field.argtypes = [C.c_void_p, C.POINTER(D2biRecord), C.c_int, C.c_int, C.c_int,
                                C.c_int, _dp, _dp, _dp, _ip, _dp, C.c_int]
        L.orc_image_window.argtypes = [C.c_void_p, C.POINTER(D2biRecord), C.c_int, C.c_int,
                                       C.c_int, C.c_int, _dp, C.c_int]
        L.orc_compute_img.restype = C.c_int
        L.orc_compute_img.argtypes = [C.POINTER(D2biRecord), _dp, C.c_int64, C.c_int64,
                                      C.c_double, C.c_int, _dp, C.c_int]
        L.orc_blur.argtypes = [_dp, C.c_int, C.c_int, C.c_double, _dp]
        L.orc_sunflower.argtypes = [C.c_int, _dp]
        L.orc_aperture_setup.restype = C.c_int
        L.orc_aperture_setup.argtypes = [C.POINTER(D2biRecord), C.c_double, C.c_int, C.c_double, C.c_int,
                                         _dp, _dp, _dp]
        L.orc_find_gdotr_aperture.restype = C.c_int64
        L.orc_find_gdotr_aperture.argtypes = [C.POINTER(D2biRecord), C.POINTER(D2biPlan), _dp, C.c_int64,
                                              C.c_int64, C.c_void_p, _dp, C.c_int, _dp, C.c_int64]
        L.orc_compute_img_aperture.restype = C.c_int
        L.orc_compute_img_aperture.argtypes = [C.POINTER(D2biRecord), _dp, C.c_int64, C.c_int64, C.c_double,
                                               C.c_int, C.c_double, C.c_int, _dp, _dp, C.c_int]
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def wavelength(V):
    return lib().orc_wavelength(V)


def velocity(V):
    return lib().orc_velocity(V)


def det3(M):
    M = _f64(np.asarray(M).T)  # column-major
    return lib().orc_det3(_p(M))


def inverse3(M):
    Mc = _f64(np.asarray(M).T)
    N = np.zeros(9)
    lib().orc_inverse3(_p(Mc), _p(N))
    return N.reshape(3, 3).T.copy()


def deviation_parameter(E_keV, g_lab):
    g = _f64(g_lab)
    return lib().orc_deviation_parameter(E_keV, _p(g))


def estimate_max_count(n_cells, n_total):
    return lib().orc_estimate_max_count(float(n_cells), int(n_total))


def plan(rec, n_atoms, prec_angle=0.0, n_prec=1):
    p = D2biPlan()
    rc = lib().orc_plan(C.byref(rec), int(n_atoms), float(prec_angle), int(n_prec), C.byref(p))
    if rc:
        raise ValueError("orc_plan failed")
    return p


class LinkCells:
    """Oracle link cells: LinkCell3D_ctor0 + add0 + neighbourList1b."""

    def __init__(self, a, nx, ny, nz, nmax):
        self.a = _f64(np.asarray(a, dtype=np.float64).T)
        self.h = lib().orc_lc_new(_p(self.a), nx, ny, nz, nmax)
        self.shape = (nx, ny, nz)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_lc_free(self.h)
            self.h = None

    def add(self, ident, x):
        x = _f64(x)
        if lib().orc_lc_add(self.h, int(ident), _p(x)):
            raise RuntimeError("link cell overflow beyond 16x capacity")

    def neighbours(self, x, rc, with_dx=False):
        nx, ny, nz = self.shape
        cap = 27 * 16 * 64 + 1024
        ids = np.zeros(cap, np.int32)
        r2 = np.zeros(cap)
        dx = np.zeros(3 * cap) if with_dx else None
        x = _f64(x)
        n = lib().orc_lc_neighbours(self.h, _p(x), float(rc), _p(ids, _ip), _p(r2),
                                    _p(dx) if with_dx else None)
        if with_dx:
            return ids[:n].copy(), r2[:n].copy(), dx[:3 * n].reshape(n, 3).copy()
        return ids[:n].copy(), r2[:n].copy()

    def count(self, ix, iy, iz):
        return lib().orc_lc_count(self.h, ix, iy, iz)

    def storage(self, ix, iy, iz):
        return lib().orc_lc_storage(self.h, ix, iy, iz)

    def nmax(self):
        return lib().orc_lc_nmax(self.h)

    def flatten(self, n_points):
        nx, ny, nz = self.shape
        cs = np.zeros(nx * ny * nz + 1, np.int32)
        ids = np.zeros(max(n_points, 1), np.int32)
        frac = np.zeros(3 * max(n_points, 1))
        n = lib().orc_lc_flatten(self.h, _p(cs, _ip), _p(ids, _ip), _p(frac))
        return cs, ids[:n], frac[:3 * n].reshape(n, 3)


def lc_flatten_handle(h, shape, n_points):
    nx, ny, nz = shape
    cs = np.zeros(nx * ny * nz + 1, np.int32)
    ids = np.zeros(max(n_points, 1), np.int32)
    frac = np.zeros(3 * max(n_points, 1))
    n = lib().orc_lc_flatten(h, _p(cs, _ip), _p(ids, _ip), _p(frac))
    return cs, ids[:n], frac[:3 * n].reshape(n, 3)


def rho(r2_min, sig):
    return lib().orc_rho(float(r2_min), float(sig))


def integrate_column(a, b0, bg, dz, x, rho_):
    """x: complex array of nz+1 samples; rho_: real array of nz+1 samples."""
    xr = _f64(np.real(x))
    xi = _f64(np.imag(x))
    rh = _f64(rho_)
    return lib().orc_integrate_column(a, b0, bg, dz, _p(xr), _p(xi), _p(rh), len(xr) - 1)


def intensity_at_depth(xi0, xig, sg, z):
    return lib().orc_intensity_at_depth(xi0, xig, sg, z)


def _atoms_colmajor(x):
    """x: (nAtoms,3) array -> contiguous column-major (3,nAtoms) buffer, ld=3."""
    x = _f64(x)
    return x, 3


class Condition:
    """One imaging condition prepared by the oracle: plan, link cells, g.r."""

    def __init__(self, rec, x, prec_angle=0.0, n_prec=1):
        self.rec = rec
        self.x, ld = _atoms_colmajor(x)
        err = np.zeros(1, np.int32)
        self.h = lib().orc_prepare(C.byref(rec), _p(self.x), ld, len(self.x), float(prec_angle),
                                   int(n_prec), _p(err, _ip))
        if not self.h:
            raise ValueError("orc_prepare failed (%d)" % err[0])
        if err[0]:
            lib().orc_state_free(self.h)
            self.h = None
            raise RuntimeError("oracle binning failed (%d)" % err[0])
        self.plan = D2biPlan.from_buffer_copy(lib().orc_state_plan(self.h).contents)
        self.n_points = lib().orc_state_npoints(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_state_free(self.h)
            self.h = None

    def gdotr(self):
        ptr = lib().orc_state_gdotr(self.h)
        return np.ctypeslib.as_array(ptr, shape=(self.n_points,)).copy()

    def cells(self):
        p = self.plan
        return lc_flatten_handle(lib().orc_state_lc(self.h), (p.Ex, p.Ey, p.Ez), self.n_points)

    def field(self, window=None, nthreads=1):
        p = self.plan
        ix0, ix1, iy0, iy1 = window or (0, p.Mx, 0, p.My)
        n = (ix1 - ix0) * (iy1 - iy0) * p.Mz
        xr = np.zeros(n)
        xi = np.zeros(n)
        rh = np.zeros(n)
        cnt = np.zeros(n, np.int32)
        r2m = np.zeros(n)
        lib().orc_field(self.h, C.byref(self.rec), ix0, ix1, iy0, iy1, _p(xr), _p(xi), _p(rh),
                        _p(cnt, _ip), _p(r2m), int(nthreads))
        shp = (p.Mz, iy1 - iy0, ix1 - ix0)
        return ((xr + 1j * xi).reshape(shp), rh.reshape(shp), cnt.reshape(shp),
                r2m.reshape(shp))

    def image(self, window=None, nthreads=1):
        p = self.plan
        ix0, ix1, iy0, iy1 = window or (0, p.Mx, 0, p.My)
        img = np.zeros((iy1 - iy0) * (ix1 - ix0))
        lib().orc_image_window(self.h, C.byref(self.rec), ix0, ix1, iy0, iy1, _p(img),
                               int(nthreads))
        return img.reshape(iy1 - iy0, ix1 - ix0)


def compute_img(rec, x, prec_angle=0.0, n_prec=1, nthreads=1):
    """Whole image, shape (My, Mx) (row = iy, column = ix)."""
    return Condition(rec, x, prec_angle, n_prec).image(nthreads=nthreads)


_APERTURE_FIXTURE = os.path.join(os.path.dirname(_HERE), "tests", "golden", "aperture_points.json")


def aperture_points(n):
    """evenlySpacedPointsInCircle (Lib_EvenlySpacedPointsInCircle.f90:29-100):
    (n, 2).  n <= 25: the reference's table (the fixture, FP32 literals);
    otherwise the sunflower restatement."""
    n = int(n)
    if n <= 25:
        import json
        pts = json.load(open(_APERTURE_FIXTURE))["points"][str(n)]
        return np.array([[float(np.float32(a)), float(np.float32(b))] for a, b in pts])
    xy = np.zeros(2 * n)
    lib().orc_sunflower(n, _p(xy))
    return xy.reshape(n, 2)


def aperture_setup(rec, prec_angle, n_prec, ap_angle, n_app):
    """(gshift (nApp,3), aa (nPrec, nApp)) of computeD2BI_img's aperture
    sampling (Lib_DynamicalTwoBeamImaging.f90:1655-1667, 1805-1836)."""
    xy = _f64(aperture_points(n_app).reshape(-1))
    gs = np.zeros(3 * n_app)
    aa = np.zeros(n_app * n_prec)
    if lib().orc_aperture_setup(C.byref(rec), float(prec_angle), int(n_prec), float(ap_angle), int(n_app),
                                _p(xy), _p(gs), _p(aa)):
        raise ValueError("orc_aperture_setup failed")
    return gs.reshape(n_app, 3), aa.reshape(n_prec, n_app)


def gdotr_aperture(rec, x, gshift):
    """findGdotR_aperture (:1269-1396) for arbitrary g-vectors gshift (k,3):
    (n_points, k), rows in point order ii."""
    x, ld = _atoms_colmajor(x)
    plan = D2biPlan()
    if lib().orc_plan(C.byref(rec), len(x), 0.0, 1, C.byref(plan)):
        raise ValueError("orc_plan failed")
    gs = _f64(np.asarray(gshift, float).reshape(-1))
    k = len(gs) // 3
    cap = rec.nuvw * len(x) + 1
    out = np.zeros(cap * k)
    h = lib().orc_lc_new(_p(_f64(np.array(plan.a_col))), plan.Ex, plan.Ey, plan.Ez, plan.nMax)
    try:
        n = lib().orc_find_gdotr_aperture(C.byref(rec), C.byref(plan), _p(x), ld, len(x), h, _p(gs), k,
                                          _p(out), cap)
    finally:
        lib().orc_lc_free(h)
    if n < 0:
        raise RuntimeError("oracle binning failed (%d)" % n)
    return out.reshape(k, cap)[:, :n].T.copy()


def compute_img_aperture(rec, x, prec_angle, n_prec, ap_angle, n_app, nthreads=1):
    """computeD2BI_img with apertureAngle and nAperturePoints: (My, Mx)."""
    x, ld = _atoms_colmajor(x)
    plan = D2biPlan()
    if lib().orc_plan(C.byref(rec), len(x), float(prec_angle), int(n_prec), C.byref(plan)):
        raise ValueError("orc_plan failed")
    xy = _f64(aperture_points(n_app).reshape(-1))
    img = np.zeros(plan.Mx * plan.My)
    rc = lib().orc_compute_img_aperture(C.byref(rec), _p(x), ld, len(x), float(prec_angle), int(n_prec),
                                        float(ap_angle), int(n_app), _p(xy), _p(img), int(nthreads))
    if rc:
        raise RuntimeError("oracle aperture image failed (%d)" % rc)
    return img.reshape(plan.My, plan.Mx)


def blur(img, t):
    """img: (Ny, Nx) array (x fastest in memory)."""
    img = _f64(img)
    ny, nx = img.shape
    out = np.zeros_like(img)
    lib().orc_blur(_p(img), nx, ny, float(t), _p(out))
    return out
