"""ctypes binding of libmsfec_hip.so (the C-ABI declared in include/msfec.h).

The library is built in-tree next to this file by ``msfec_amd.build.build()``.
There is no fallback: if the shared object is missing or fails to load, every
entry point raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libmsfec_hip.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)

# C enums (include/msfec.h)
OK, ERR_INVALID, ERR_HIP, ERR_STATE, ERR_UNSUPPORTED, ERR_CELL = 0, 1, 2, 3, 4, 5
HOST, DEVICE = 0, 1
COEF_SCALAR_CELL, COEF_TENSOR_QP = 0, 1
RHS_CONST, RHS_CELL, RHS_QP = 0, 1, 2
REACTION_NONE, REACTION_CONST, REACTION_CELL = 0, 1, 2
SOLVER_BAND_FP64, SOLVER_BAND_FP64_MFMA, SOLVER_BAND_MIXED, SOLVER_CG, SOLVER_AUTO = 0, 1, 2, 3, 4
CELL_OK, CELL_NOT_CONVERGED, CELL_NOT_SPD, CELL_NONFINITE = 0, 1, 2, 3
FLAG_TIMING = 1

#: every symbol include/msfec.h declares
EXPORTED = (
    "msfec_q_create", "msfec_q_set_coefficient", "msfec_q_set_rhs", "msfec_q_set_reaction",
    "msfec_q_compute", "msfec_q_synchronize", "msfec_q_get_element", "msfec_q_get_basis",
    "msfec_q_get_cell_status", "msfec_q_reconstruct", "msfec_q_get_info",
    "msfec_q_kernel_times", "msfec_q_destroy", "msfec_coarse_solve", "msfec_q_adjoint_gradient",
    "msfec_q_get_cg_iterations", "msfec_q_compute_noalloc",
    "msfec_ned_create", "msfec_ned_set_coefficients", "msfec_ned_compute",
    "msfec_ned_synchronize", "msfec_ned_get_element", "msfec_ned_get_basis",
    "msfec_ned_get_cell_status", "msfec_ned_kernel_times", "msfec_ned_destroy",
    "msfec_rt_create", "msfec_rt_set_coefficients", "msfec_rt_compute", "msfec_rt_synchronize",
    "msfec_rt_get_element", "msfec_rt_get_basis", "msfec_rt_reconstruct",
    "msfec_rt_get_cell_status", "msfec_rt_kernel_times", "msfec_rt_destroy",
    "msfec_nrt_create", "msfec_nrt_set_coefficients", "msfec_nrt_compute",
    "msfec_nrt_synchronize", "msfec_nrt_get_element", "msfec_nrt_get_basis",
    "msfec_nrt_reconstruct", "msfec_nrt_get_cell_status", "msfec_nrt_get_iterations",
    "msfec_nrt_kernel_times", "msfec_nrt_destroy",
    "msfec_rt3_create", "msfec_rt3_set_coefficients", "msfec_rt3_compute",
    "msfec_rt3_synchronize", "msfec_rt3_get_element", "msfec_rt3_get_basis",
    "msfec_rt3_get_cell_status", "msfec_rt3_get_iterations", "msfec_rt3_kernel_times",
    "msfec_rt3_destroy",
    "msfec_qu_create", "msfec_qu_offsets", "msfec_qu_set_coefficient", "msfec_qu_set_rhs",
    "msfec_qu_set_reaction", "msfec_qu_tune", "msfec_qu_level_info", "msfec_qu_level_handle",
    "msfec_qu_compute", "msfec_qu_synchronize", "msfec_qu_get_element", "msfec_qu_get_basis",
    "msfec_qu_get_cell_status", "msfec_qu_coarse_solve", "msfec_qu_destroy",
    "msfec_last_error", "msfec_version",
)


class Params(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("n_refine_local", ctypes.c_int),
        ("coef_mode", ctypes.c_int),
        ("rhs_mode", ctypes.c_int),
        ("rhs_const", ctypes.c_double),
        ("reaction_mode", ctypes.c_int),
        ("reaction_const", ctypes.c_double),
        ("solver", ctypes.c_int),
        ("flags", ctypes.c_int),
        ("max_refine", ctypes.c_int),
        ("cg_tol", ctypes.c_double),
        ("cg_max_iter", ctypes.c_int),
        ("band_window", ctypes.c_int),
    ]


class NedParams(ctypes.Structure):
    _fields_ = [
        ("n_refine_local", ctypes.c_int),
        ("full_rhs", ctypes.c_int),
        ("flags", ctypes.c_int),
        ("solver", ctypes.c_int),
        ("max_refine", ctypes.c_int),
    ]


class RtParams(ctypes.Structure):
    _fields_ = [
        ("n_refine_local", ctypes.c_int),
        ("flags", ctypes.c_int),
        ("coef_mode", ctypes.c_int),
    ]


class Rt3Params(ctypes.Structure):
    _fields_ = [
        ("n_refine_local", ctypes.c_int),
        ("flags", ctypes.c_int),
        ("max_iter", ctypes.c_int),
        ("rtol", ctypes.c_double),
    ]


class NrtParams(ctypes.Structure):
    _fields_ = [
        ("n_refine_local", ctypes.c_int),
        ("flags", ctypes.c_int),
        ("solver", ctypes.c_int),
        ("max_iter", ctypes.c_int),
        ("rtol", ctypes.c_double),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("n_cells", ctypes.c_int),
        ("dim", ctypes.c_int),
        ("n_refine_local", ctypes.c_int),
        ("n_basis", ctypes.c_int),
        ("n_fine_nodes", ctypes.c_int),
        ("n_fine_cells", ctypes.c_int),
        ("n_interior", ctypes.c_int),
        ("bandwidth", ctypes.c_int),
        ("band_window", ctypes.c_int),
        ("coef_count", ctypes.c_size_t),
        ("rhs_count", ctypes.c_size_t),
        ("reaction_count", ctypes.c_size_t),
        ("device_bytes", ctypes.c_size_t),
    ]


class MsfecError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"msfec status {status}: {msg}")
        self.status = status


_lib = None


def load(path: str | None = None):
    """Load (once) and return the ctypes library object."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with msfec_amd.build.build() (hipcc, gfx950); "
            "there is no CPU fallback")
    L = ctypes.CDLL(path)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    L.msfec_q_create.argtypes = [ctypes.POINTER(Params), ctypes.c_int, vp, ctypes.c_int,
                                 ctypes.POINTER(vp)]
    L.msfec_q_set_coefficient.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    L.msfec_q_set_rhs.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    L.msfec_q_set_reaction.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    L.msfec_q_compute.argtypes = [vp, vp]
    L.msfec_q_synchronize.argtypes = [vp]
    L.msfec_q_get_element.argtypes = [vp, vp, vp, ctypes.c_int]
    L.msfec_q_get_basis.argtypes = [vp, vp, ctypes.c_int]
    L.msfec_q_get_cell_status.argtypes = [vp, ip]
    L.msfec_q_reconstruct.argtypes = [vp, vp, ctypes.c_int, vp, ctypes.c_int]
    L.msfec_q_get_info.argtypes = [vp, ctypes.POINTER(Info)]
    L.msfec_q_kernel_times.argtypes = [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int, ip]
    L.msfec_q_destroy.argtypes = [vp]
    L.msfec_q_destroy.restype = None
    L.msfec_coarse_solve.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_int, ip, dp]
    L.msfec_q_adjoint_gradient.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_int, dp, vp, ctypes.c_int]
    L.msfec_q_get_cg_iterations.argtypes = [vp, ip]
    L.msfec_ned_create.argtypes = [ctypes.POINTER(NedParams), ctypes.c_int, vp, ctypes.c_int,
                                   ctypes.POINTER(vp)]
    L.msfec_ned_set_coefficients.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int]
    L.msfec_ned_compute.argtypes = [vp, vp]
    L.msfec_ned_synchronize.argtypes = [vp]
    L.msfec_ned_get_element.argtypes = [vp, vp, vp, ctypes.c_int]
    L.msfec_ned_get_basis.argtypes = [vp, vp, ctypes.c_int]
    L.msfec_ned_get_cell_status.argtypes = [vp, ip]
    L.msfec_ned_kernel_times.argtypes = [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int, ip]
    L.msfec_ned_destroy.argtypes = [vp]
    L.msfec_rt_create.argtypes = [ctypes.POINTER(RtParams), ctypes.c_int, vp, ctypes.c_int,
                                  ctypes.POINTER(vp)]
    L.msfec_rt_set_coefficients.argtypes = [vp, vp, vp, ctypes.c_int]
    L.msfec_rt_compute.argtypes = [vp, vp]
    L.msfec_rt_synchronize.argtypes = [vp]
    L.msfec_rt_get_element.argtypes = [vp, vp, vp, ctypes.c_int]
    L.msfec_rt_get_basis.argtypes = [vp, vp, ctypes.c_int]
    L.msfec_rt_reconstruct.argtypes = [vp, vp, ctypes.c_int, vp, vp, ctypes.c_int]
    L.msfec_rt_get_cell_status.argtypes = [vp, ip]
    L.msfec_rt_kernel_times.argtypes = [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int, ip]
    L.msfec_rt_destroy.argtypes = [vp]
    L.msfec_nrt_create.argtypes = [ctypes.POINTER(NrtParams), ctypes.c_int, vp, ctypes.c_int,
                                   ctypes.POINTER(vp)]
    L.msfec_nrt_set_coefficients.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int]
    L.msfec_nrt_compute.argtypes = [vp, vp]
    L.msfec_nrt_synchronize.argtypes = [vp]
    L.msfec_nrt_get_element.argtypes = [vp, vp, vp, ctypes.c_int]
    L.msfec_nrt_get_basis.argtypes = [vp, vp, ctypes.c_int]
    L.msfec_nrt_reconstruct.argtypes = [vp, vp, ctypes.c_int, vp, vp, ctypes.c_int]
    L.msfec_nrt_get_cell_status.argtypes = [vp, ip]
    L.msfec_nrt_get_iterations.argtypes = [vp, ip]
    L.msfec_rt3_create.argtypes = [ctypes.POINTER(Rt3Params), ctypes.c_int, vp, ctypes.c_int,
                                   ctypes.POINTER(vp)]
    L.msfec_rt3_set_coefficients.argtypes = [vp, vp, vp, ctypes.c_int]
    L.msfec_rt3_compute.argtypes = [vp, vp]
    L.msfec_rt3_synchronize.argtypes = [vp]
    L.msfec_rt3_get_element.argtypes = [vp, vp, vp, ctypes.c_int]
    L.msfec_rt3_get_basis.argtypes = [vp, vp, ctypes.c_int]
    L.msfec_rt3_get_cell_status.argtypes = [vp, ip]
    L.msfec_rt3_get_iterations.argtypes = [vp, ip]
    L.msfec_rt3_kernel_times.argtypes = [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int, ip]
    L.msfec_rt3_destroy.argtypes = [vp]
    L.msfec_nrt_kernel_times.argtypes = [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_int, ip]
    L.msfec_nrt_destroy.argtypes = [vp]
    sp = ctypes.POINTER(ctypes.c_size_t)
    L.msfec_qu_create.argtypes = [ctypes.POINTER(Params), ctypes.c_int, vp, ctypes.c_int, ip,
                                  ctypes.POINTER(vp)]
    L.msfec_qu_offsets.argtypes = [vp, sp, sp]
    L.msfec_qu_set_coefficient.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    L.msfec_qu_set_rhs.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    L.msfec_qu_set_reaction.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    L.msfec_qu_tune.argtypes = [vp]
    L.msfec_qu_level_info.argtypes = [vp, ctypes.c_int, ip, ip, ip, ctypes.POINTER(ctypes.c_float)]
    L.msfec_qu_level_handle.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
    L.msfec_qu_compute.argtypes = [vp, vp]
    L.msfec_qu_synchronize.argtypes = [vp]
    L.msfec_qu_get_element.argtypes = [vp, vp, vp, ctypes.c_int]
    L.msfec_qu_get_basis.argtypes = [vp, vp, ctypes.c_int]
    L.msfec_qu_get_cell_status.argtypes = [vp, ip]
    L.msfec_qu_coarse_solve.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_int, ip,
                                        ctypes.POINTER(ctypes.c_double)]
    L.msfec_qu_destroy.argtypes = [vp]
    L.msfec_qu_destroy.restype = None
    L.msfec_last_error.restype = ctypes.c_char_p
    L.msfec_version.restype = ctypes.c_char_p
    for name in EXPORTED:
        fn = getattr(L, name)
        if name not in ("msfec_q_destroy", "msfec_qu_destroy", "msfec_last_error", "msfec_version"):
            fn.restype = ctypes.c_int
    _lib = L
    return L


def check(status: int):
    if status != OK:
        raise MsfecError(status, load().msfec_last_error().decode(errors="replace"))
