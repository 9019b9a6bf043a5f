"""ctypes binding of ``liberlbm.so`` (the HIP library's C ABI, include/erlbm.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liberlbm.so")

ERLBM_OK, ERLBM_ERR_ARG, ERLBM_ERR_HIP, ERLBM_ERR_STATE, ERLBM_ERR_NOMEM, ERLBM_ERR_UNSUPPORTED = range(6)
COL_BGK, COL_MRT, COL_REG, COL_EXT_REG = 1, 2, 3, 4
FP64, FP32_SHIFTED, FP16_SHIFTED, FP32_SHIFTED_F32 = 0, 1, 2, 4
STATE, POST_COLLISION = 0, 1
ARITH_EXACT, ARITH_MOMENT = 0, 1
STOP_DENSITY, STOP_ENERGY = 1, 2

_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_i, _d, _ll = C.c_int, C.c_double, C.c_longlong

INIT_REST, INIT_TAYLOR_GREEN = 0, 1
NOALLOC_MAX_JOBS = 16384   # ERLBM_NOALLOC_MAX_JOBS
FP_AUTO = 3


class Job(C.Structure):
    """erlbm_job (include/erlbm.h)."""
    _fields_ = [("tau", C.c_double), ("tau_nh", C.c_double), ("u0", C.c_double), ("rho0", C.c_double),
                ("energy_factor", C.c_double), ("dens_limit", C.c_double), ("n_steps", C.c_longlong),
                ("check_every", C.c_int), ("init", C.c_int), ("collision", C.c_int), ("precision", C.c_int)]


class JobResult(C.Structure):
    """erlbm_job_result (include/erlbm.h)."""
    _fields_ = [("unstable", C.c_int), ("steps", C.c_int), ("done", C.c_int), ("precision", C.c_int),
                ("e0", C.c_double), ("energy", C.c_double), ("margin", C.c_double)]


class Request(C.Structure):
    """erlbm_request (include/erlbm.h)."""
    _fields_ = [("kind", C.c_int), ("i", C.c_int), ("j", C.c_int), ("n_i", C.c_int), ("n_j", C.c_int),
                ("L", C.c_int), ("collision", C.c_int), ("precision", C.c_int), ("tau_nh", C.c_double),
                ("t_final", C.c_double)]


class Response(C.Structure):
    """erlbm_response (include/erlbm.h)."""
    _fields_ = [("id", C.c_longlong), ("kind", C.c_int), ("status", C.c_int), ("unstable", C.c_int),
                ("precision", C.c_int), ("steps", C.c_longlong), ("value", C.c_double * 3),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]


JOB_DTYPE = np.dtype([("tau", "f8"), ("tau_nh", "f8"), ("u0", "f8"), ("rho0", "f8"), ("energy_factor", "f8"),
                      ("dens_limit", "f8"), ("n_steps", "i8"), ("check_every", "i4"), ("init", "i4"),
                      ("collision", "i4"), ("precision", "i4")])
RESULT_DTYPE = np.dtype([("unstable", "i4"), ("steps", "i4"), ("done", "i4"), ("precision", "i4"), ("e0", "f8"),
                         ("energy", "f8"), ("margin", "f8")])
assert JOB_DTYPE.itemsize == C.sizeof(Job) and RESULT_DTYPE.itemsize == C.sizeof(JobResult)


class ErlbmError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed with code {code}: {msg}")
        self.code = code


# (name, restype, argtypes) of every C-ABI entry point declared in include/erlbm.h
SIGNATURES = {
    "erlbm_last_error": (C.c_char_p, []),
    "erlbm_device_count": (_i, [C.POINTER(_i)]),
    "erlbm_lattice_create": (_i, [_i, _i, _i, _i, _vp, _i, _vp, C.POINTER(_vp)]),
    "erlbm_lattice_destroy": (_i, [_vp]),
    "erlbm_lattice_fluid_count": (_i, [_vp, C.POINTER(_i)]),
    "erlbm_set_params": (_i, [_vp, _d, _d, _d, _i]),
    "erlbm_set_zouhe_inlet": (_i, [_vp, _i, _d]),
    "erlbm_set_outlet_rescale": (_i, [_vp, _i, _d]),
    "erlbm_set_density_check": (_i, [_vp, _d]),
    "erlbm_set_body_force": (_i, [_vp, _d, _d, _d]),
    "erlbm_set_arith": (_i, [_vp, _i]),
    "erlbm_autotune_candidates": (_i, [_i, _i, _i, _i, _i, _i, _ip, _i, C.POINTER(_i)]),
    "erlbm_upload_aos": (_i, [_vp, _dp]),
    "erlbm_download_aos": (_i, [_vp, _i, _dp]),
    "erlbm_step": (_i, [_vp, _i, _i, _d]),
    "erlbm_step_timed": (_i, [_vp, _i, C.POINTER(C.c_float)]),
    "erlbm_status": (_i, [_vp, C.POINTER(_ll), C.POINTER(_i)]),
    "erlbm_synchronize": (_i, [_vp]),
    "erlbm_kinetic_energy": (_i, [_vp, _i, _i, C.POINTER(_d)]),
    "erlbm_density_monitor": (_i, [_vp, _i, _d, C.POINTER(_d), C.POINTER(_i)]),
    "erlbm_hbm_copy_bench": (_i, [_ll, _i, C.POINTER(C.c_float)]),
    "erlbm_taylor_green_error": (_i, [_i, _i, _d, _d, _ll, _dp, _dp]),
    "erlbm_server_create": (_i, [_ll, _i, _i, _vp, C.POINTER(_vp)]),
    "erlbm_server_destroy": (_i, [_vp]),
    "erlbm_server_submit": (_i, [_vp, C.POINTER(Request), C.POINTER(_ll)]),
    "erlbm_server_pump": (_i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
    "erlbm_server_poll": (_i, [_vp, C.POINTER(Response), _i, C.POINTER(_i)]),
    "erlbm_server_clock": (_i, [_vp, C.POINTER(_d)]),
    "erlbm_server_memory": (_i, [_vp, C.POINTER(_ll), C.POINTER(_ll)]),
    "erlbm_batch_create": (_i, [_i, _i, _vp, _i, _i, _vp, C.POINTER(_vp)]),
    "erlbm_batch_destroy": (_i, [_vp]),
    "erlbm_batch_run": (_i, [_vp, _vp, _i, _vp]),
    "erlbm_batch_run_noalloc": (_i, [_vp, _vp, _i, _vp]),
    "erlbm_batch_set_launch": (_i, [_vp, _i, _i]),
    "erlbm_batch_set_chunk": (_i, [_vp, _i]),
    "erlbm_lattice_autotune": (_i, [_vp, _i, C.c_char_p, _i, C.POINTER(C.c_float)]),
    "erlbm_step_shape": (_i, [_vp, C.POINTER(_i), C.POINTER(_i), C.c_char_p, _i]),
    "erlbm_velocity_field": (_i, [_vp, _i, _vp]),
    "erlbm_vorticity_field": (_i, [_vp, _i, _vp]),
    "erlbm_enstrophy": (_i, [_vp, _i, C.POINTER(C.c_double)]),
    "erlbm_write_velocity_vtk": (_i, [C.c_char_p, _i, _i, _i, _vp]),
    "erlbm_write_vorticity_vtk": (_i, [C.c_char_p, _i, _i, _i, _vp]),
    "erlbm_read_velocity_vtk": (_i, [C.c_char_p, _vp, C.c_longlong, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)]),
    "erlbm_vtk_round": (_i, [_vp, C.c_longlong]),
    "erlbm_velocity_l2_error": (_i, [_vp, _i, _i, _vp, _i, _i, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "erlbm_write_tg_map_vtk": (_i, [C.c_char_p, _i, _i, _i, C.c_double, _vp]),
    "erlbm_write_obstacle_map_vtk": (_i, [C.c_char_p, _i, _i, C.c_double, _vp]),
    "erlbm_batch_autotune": (_i, [_vp, _vp, _i, C.c_longlong, _vp, _vp, _vp]),
    "erlbm_tg_map_jobs": (_i, [_i, _i, _i, _d, _i, _i, _i, _i, _i, _vp]),
    "erlbm_obstacle_map_jobs": (_i, [_i, _i, _d, _i, _i, _i, _i, _i, _vp]),
    "erlbm_dist_eq": (_i, [_i, _d, _d, _d, _d, _dp]),
    "erlbm_init_taylor_green": (_i, [_i, _i, _d, _d, _i, _i, _dp]),
    "erlbm_init_double_shear": (_i, [_i, _d, _dp]),
    "erlbm_init_tg_map": (_i, [_i, _d, _d, _dp]),
    "erlbm_tg_map_tables": (_i, [_i, _dp, _dp]),
    "erlbm_obstacle_mask": (_i, [_ip]),
    "erlbm_tg_map_point_params": (_i, [_i, _i, _i, _i, _i, _i, _d, C.POINTER(_d), C.POINTER(_d), C.POINTER(_d),
                                       C.POINTER(_i)]),
    "erlbm_obstacle_map_point_params": (_i, [_i, _i, _i, _i, _d, C.POINTER(_d), C.POINTER(_d), C.POINTER(_d),
                                             C.POINTER(_i)]),
}


def lib() -> C.CDLL:
    """The loaded library; raises if it was not built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m erlbm_amd.build` "
                              "(the HIP path has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, where: str) -> None:
    if rc != ERLBM_OK:
        msg = lib().erlbm_last_error()
        raise ErlbmError(rc, where, msg.decode() if msg else "")


def hbm_copy_bench(nbytes: int, iters: int = 20) -> float:
    """Milliseconds per launch of a plain copy of nbytes (read nbytes + write nbytes)."""
    ms = C.c_float(0)
    check(lib().erlbm_hbm_copy_bench(nbytes, iters, C.byref(ms)), "erlbm_hbm_copy_bench")
    return ms.value


def device_count() -> int:
    n = C.c_int(0)
    check(lib().erlbm_device_count(C.byref(n)), "erlbm_device_count")
    return n.value
