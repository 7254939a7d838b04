"""ctypes view of the C-ABI in include/d2bi_gpu.h and the loader of
libd2bi_gpu.so.

The shared library is built in-tree (simulated_tem_amd/_lib/libd2bi_gpu.so)
by simulated_tem_amd.build.  Loading fails loudly when it is missing: there
is no CPU fallback behind this interface.
"""
import ctypes as C
import os

import numpy as np

D2BI_MAX_UVW = 125
D2BI_MAX_PREC = 256

D2BI_MODE_FP64 = 0
D2BI_MODE_FAST = 1

ERRORS = {0: "ok", -1: "bad argument", -2: "HIP error", -3: "capacity exceeded",
          -4: "not initialised", -5: "no device"}

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libd2bi_gpu.so")


class D2biRecord(C.Structure):
    """type DynamicalTwoBeamImaging fields read by computeD2BI_img
    (Lib_DynamicalTwoBeamImaging.f90:64-90)."""
    _fields_ = [
        ("a0", C.c_double),
        ("a_cell", C.c_double * 9),
        ("g", C.c_double * 3),
        ("E", C.c_double),
        ("xi0", C.c_double),
        ("xig", C.c_double),
        ("U", C.c_double * 9),
        ("R", C.c_double * 9),
        ("delta", C.c_double * 3),
        ("super_a", C.c_double * 9),
        ("super_n", C.c_int32 * 3),
        ("nuvw", C.c_int32),
        ("uvw", C.c_int32 * (3 * D2BI_MAX_UVW)),
        ("gk_a", C.c_double * 9),
        ("gk_n", C.c_int32 * 3),
        ("use_density", C.c_int32),
        ("sg", C.c_double),
    ]


class D2biPlan(C.Structure):
    """computeD2BI_img prologue (Lib_DynamicalTwoBeamImaging.f90:1635-1836)."""
    _fields_ = [
        ("Mx", C.c_int32), ("My", C.c_int32), ("Mz", C.c_int32),
        ("deltaM", C.c_int32), ("nMax", C.c_int32),
        ("Ex", C.c_int32), ("Ey", C.c_int32), ("Ez", C.c_int32),
        ("nPrec", C.c_int32), ("pad0", C.c_int32),
        ("sigma", C.c_double), ("i2s2", C.c_double), ("rc2", C.c_double),
        ("a_col", C.c_double * 9), ("ia_col", C.c_double * 9),
        ("lc_delta", C.c_double * 3), ("RU", C.c_double * 9), ("gg", C.c_double * 3),
        ("xmax", C.c_double * 3), ("aperpx", C.c_double), ("modg", C.c_double),
        ("dz", C.c_double), ("b0", C.c_double), ("bg", C.c_double),
        ("aa", C.c_double * D2BI_MAX_PREC),
    ]


class D2biStats(C.Structure):
    _fields_ = [
        ("n_points", C.c_int64), ("n_guard", C.c_int64),
        ("max_per_cell", C.c_int32), ("n_cells", C.c_int32),
        ("ms_bin", C.c_double), ("ms_voxel", C.c_double),
        ("ms_integrate", C.c_double), ("ms_total", C.c_double),
        ("branch_bytes", C.c_int64), ("n_branches", C.c_int32), ("pad0", C.c_int32),
    ]


class D2biSynthParams(C.Structure):
    _fields_ = [
        ("a0", C.c_double),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("n_centres", C.c_int32),
        ("A_min", C.c_double), ("A_max", C.c_double), ("u_cap", C.c_double),
        ("n_voids", C.c_int32), ("pad0", C.c_int32),
        ("r_min", C.c_double), ("r_max", C.c_double),
        ("noise", C.c_double),
        ("offset", C.c_double * 3),
        ("seed", C.c_uint64),
    ]


# Every symbol include/d2bi_gpu.h declares, for the export check.
EXPORTED = [
    "d2bi_gpu_init", "d2bi_gpu_finalize", "d2bi_last_error", "d2bi_version",
    "d2bi_plan_image", "d2bi_setup_tilt_sg", "d2bi_setup_tilt_ur", "d2bi_setup_tilt_ng",
    "d2bi_rotate_to_standard_frame", "d2bi_wavelength", "d2bi_velocity",
    "d2bi_deviation_parameter", "d2bi_aperture_points", "d2bi_aperture_setup",
    "d2bi_gpu_atoms_upload", "d2bi_gpu_atoms_free", "d2bi_gpu_atoms_count",
    "d2bi_gpu_image", "d2bi_gpu_image_aperture", "d2bi_gpu_image_dev", "d2bi_gpu_image_dev_out",
    "d2bi_gpu_image_noalloc", "d2bi_gpu_image_batch", "d2bi_gpu_image_batch_dev_out", "d2bi_gpu_field",
    "d2bi_gpu_linkcells", "d2bi_gpu_lc_query", "d2bi_gpu_blur", "d2bi_synth_bcc",
]

_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class D2biError(RuntimeError):
    pass


def load():
    """Load libd2bi_gpu.so (no GPU calls are made by loading)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libd2bi_gpu.so is not built (%s); run "
                          "python -m simulated_tem_amd.build" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    L.d2bi_gpu_init.argtypes = [C.c_int]
    L.d2bi_last_error.restype = C.c_char_p
    L.d2bi_plan_image.argtypes = [C.POINTER(D2biRecord), C.c_int64, C.c_double, C.c_int32,
                                  C.POINTER(D2biPlan)]
    L.d2bi_setup_tilt_sg.argtypes = [C.POINTER(D2biRecord), _dp, C.c_double, C.c_int32,
                                     C.c_int32, _dp]
    L.d2bi_setup_tilt_ur.argtypes = [C.POINTER(D2biRecord), _dp, _dp, C.c_int32, C.c_int32, _dp]
    L.d2bi_setup_tilt_ng.argtypes = [C.POINTER(D2biRecord), _dp, C.c_double, C.c_int32, C.c_char_p,
                                     C.c_int32, C.c_int32, _dp]
    L.d2bi_rotate_to_standard_frame.argtypes = [_dp, _dp, _dp, _dp]
    L.d2bi_wavelength.restype = C.c_double
    L.d2bi_wavelength.argtypes = [C.c_double]
    L.d2bi_velocity.restype = C.c_double
    L.d2bi_velocity.argtypes = [C.c_double]
    L.d2bi_deviation_parameter.restype = C.c_double
    L.d2bi_deviation_parameter.argtypes = [C.c_double, _dp]
    L.d2bi_aperture_points.argtypes = [C.c_int32, _dp]
    L.d2bi_aperture_setup.argtypes = [C.POINTER(D2biRecord), C.c_double, C.c_int32, C.c_double,
                                      C.c_int32, _dp, _dp]
    L.d2bi_gpu_atoms_upload.argtypes = [_dp, C.c_int64, C.c_int64, C.POINTER(C.c_void_p)]
    L.d2bi_gpu_atoms_free.argtypes = [C.c_void_p]
    L.d2bi_gpu_atoms_count.restype = C.c_int64
    L.d2bi_gpu_atoms_count.argtypes = [C.c_void_p]
    L.d2bi_gpu_image.argtypes = [C.POINTER(D2biRecord), _dp, C.c_int64, C.c_int64, C.c_double,
                                 C.c_int32, C.c_int32, _dp, C.POINTER(D2biStats)]
    L.d2bi_gpu_image_aperture.argtypes = [C.POINTER(D2biRecord), _dp, C.c_int64, C.c_int64, C.c_double,
                                          C.c_int32, C.c_double, C.c_int32, C.c_int32, _dp,
                                          C.POINTER(D2biStats)]
    L.d2bi_gpu_image_dev.argtypes = [C.POINTER(D2biRecord), C.c_void_p, C.c_double, C.c_int32,
                                     C.c_int32, _dp, C.POINTER(D2biStats)]
    L.d2bi_gpu_image_dev_out.argtypes = [C.POINTER(D2biRecord), C.c_void_p, C.c_double, C.c_int32,
                                         C.c_int32, C.c_void_p, C.POINTER(D2biStats)]
    L.d2bi_gpu_image_noalloc.argtypes = [C.POINTER(D2biRecord), C.c_void_p, C.c_double, C.c_int32,
                                         C.c_int32, C.c_void_p, C.POINTER(D2biStats)]
    L.d2bi_gpu_image_batch.argtypes = [C.POINTER(D2biRecord), C.POINTER(C.c_void_p), C.c_int32,
                                       _dp, _ip, C.c_int32, _dp, C.POINTER(D2biStats)]
    L.d2bi_gpu_image_batch_dev_out.argtypes = [C.POINTER(D2biRecord), C.POINTER(C.c_void_p), C.c_int32,
                                               _dp, _ip, C.c_int32, C.c_void_p, C.POINTER(D2biStats)]
    L.d2bi_gpu_field.argtypes = [C.POINTER(D2biRecord), _dp, C.c_int64, C.c_int64, C.c_int32,
                                 _dp, _dp, _ip, _dp]
    L.d2bi_gpu_linkcells.argtypes = [C.POINTER(D2biRecord), _dp, C.c_int64, C.c_int64, _ip, _ip,
                                     _dp, _dp, C.c_int64, _i64p]
    L.d2bi_gpu_lc_query.argtypes = [_dp, C.c_int32, C.c_int32, C.c_int32, _dp, C.c_int64, _dp,
                                    C.c_int64, C.c_double, C.c_int32, _ip, _ip, _dp]
    L.d2bi_gpu_blur.argtypes = [_dp, C.c_int32, C.c_int32, C.c_double, _dp]
    L.d2bi_synth_bcc.argtypes = [C.POINTER(D2biSynthParams), _dp, C.c_int64, _i64p, _dp]
    _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = load().d2bi_last_error()
        raise D2biError("%s (%d): %s" % (ERRORS.get(rc, "error"), rc,
                                         msg.decode() if msg else ""))
    return rc


def ptr(a, t=_dp):
    return a.ctypes.data_as(t)


def iptr(a):
    return a.ctypes.data_as(_ip)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def mat_cm(m):
    """3x3 numpy matrix -> 9 doubles column-major (Fortran order)."""
    return (C.c_double * 9)(*np.asarray(m, dtype=np.float64).T.reshape(9))


def mat_from_cm(a):
    return np.array(list(a), dtype=np.float64).reshape(3, 3).T.copy()
