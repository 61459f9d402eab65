"""ctypes binding of libmsm_hip.so (the C-ABI declared in include/msm_hip.h).

The shared library is built in-tree by ``__graft_entry__.build()``.  There is no fallback: if the
library is missing, or no HIP device is present when a device entry point is called, the call
raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmsm_hip.so")

MSM_FS_TYPE = 0
MSM_EAM_TYPE = 1

STATU_ACTIVE = 1
STATU_FIXPOSX, STATU_FIXPOSY, STATU_FIXPOSZ = 2, 4, 8
STATU_FIXVELX, STATU_FIXVELY, STATU_FIXVELZ = 16, 32, 64
STATU_OUTOFBOX = 65536
STATU_REFLECT = 131072
STATU_TRANSMIT = 262144
STATU_PASSBOUND = 524288

ERRORS = {
    0: "MSM_OK",
    1: "MSM_ERR_ARG",
    2: "MSM_ERR_HIP",
    3: "MSM_ERR_NOMEM",
    4: "MSM_ERR_NLIST_OVERFLOW",
    5: "MSM_ERR_OUTOFBOX",
    6: "MSM_ERR_EMBD_RANGE",
    7: "MSM_ERR_STATE",
}

PAIRFN = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double))

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p

# name -> (restype, argtypes); every function declared in include/msm_hip.h
SIGNATURES = {
    "msm_strerror": (C.c_char_p, [C.c_int]),
    "msm_table_pairwise": (C.c_int, [C.c_int, C.c_double, _vp, _vp, C.c_double, _dp, _dp, _dp, _dp, _dp, C.c_int]),
    "msm_table_embd": (C.c_int, [C.c_int, C.c_double, _vp, _dp, _dp, C.c_int]),
    "msm_pot_wfs_nn": (None, [_dp, _dp, _dp]),
    "msm_pot_wfs_ne": (None, [_dp, _dp, _dp]),
    "msm_pot_ackland_ww_nn": (None, [_dp, _dp, _dp]),
    "msm_pot_ackland_ww_ne": (None, [_dp, _dp, _dp]),
    "msm_emb_fs": (None, [_dp, _dp, _dp]),
    "msm_pot_tbsma_nn": (None, [C.c_int, _dp, _dp, _dp]),
    "msm_pot_tbsma_ne": (None, [C.c_int, _dp, _dp, _dp]),
    "msm_pot_tbsma_cucu_nn": (None, [_dp, _dp, _dp]),
    "msm_pot_tbsma_cucu_ne": (None, [_dp, _dp, _dp]),
    "msm_pot_tbsma_nini_nn": (None, [_dp, _dp, _dp]),
    "msm_pot_tbsma_nini_ne": (None, [_dp, _dp, _dp]),
    "msm_pot_tbsma_cuni_nn": (None, [_dp, _dp, _dp]),
    "msm_pot_tbsma_cuni_ne": (None, [_dp, _dp, _dp]),
    "msm_drand32": (C.c_double, [_ip]),
    "msm_thermalize_mc": (C.c_int, [C.c_int, _ip, _dp, C.c_double, C.c_double, _ip, _dp]),
    "msm_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
    "msm_destroy": (C.c_int, [_vp]),
    "msm_last_error": (C.c_char_p, [_vp]),
    "msm_set_force_table": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _ip,
                                      _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_double, _ip, _dp, _dp]),
    "msm_set_ctrl": (C.c_int, [_vp, C.c_int, _dp, _dp, C.c_int, _dp]),
    "msm_set_boxes": (C.c_int, [_vp, C.c_int, _ip, _dp, _dp, _ip]),
    "msm_set_atoms": (C.c_int, [_vp, C.c_int, _dp, _dp, _ip, _ip]),
    "msm_get_atoms": (C.c_int, [_vp, _dp, _dp, _dp, _ip, _dp]),
    "msm_put_atoms": (C.c_int, [_vp, _dp, _dp]),
    "msm_check_timestep": (C.c_int, [_vp, C.c_double, C.c_double, _ip]),
    "msm_get_den": (C.c_int, [_vp, _dp]),
    "msm_get_epot": (C.c_int, [_vp, _dp]),
    "msm_get_box_epot": (C.c_int, [_vp, _dp]),
    "msm_build_neighbors": (C.c_int, [_vp]),
    "msm_get_neighbors": (C.c_int, [_vp, _ip, _ip, C.c_int]),
    "msm_calc_force": (C.c_int, [_vp]),
    "msm_calc_density": (C.c_int, [_vp]),
    "msm_add_forces": (C.c_int, [_vp, _vp, C.c_int]),
    "msm_add_epot": (C.c_int, [_vp, _vp, C.c_int]),
    "msm_put_dis": (C.c_int, [_vp, _vp]),
    "msm_thermalize": (C.c_int, [_vp, C.c_double, C.c_double, C.c_ulonglong]),
    "msm_calc_epot": (C.c_int, [_vp]),
    "msm_predictor": (C.c_int, [_vp, C.c_double]),
    "msm_correction": (C.c_int, [_vp, C.c_double]),
    "msm_run": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_double]),
    "msm_sync": (C.c_int, [_vp]),
    "msm_list_stats": (C.c_int, [_vp, C.POINTER(C.c_longlong), _ip]),
    "msm_set_option": (C.c_int, [_vp, C.c_int, C.c_int]),
    "msm_get_option": (C.c_int, [_vp, C.c_int, _ip]),
    "msm_kernel_times": (C.c_int, [_vp, _dp, C.POINTER(C.c_longlong), C.c_int, C.c_int]),
    "msm_autotune": (C.c_int, [_vp, C.c_int, _ip, _dp]),
    "msm_calc_ptensor": (C.c_int, [_vp, _dp]),
    "msm_calc_ptensor_noalloc": (C.c_int, [_vp, _dp]),
    "msm_calc_atomic_stress": (C.c_int, [_vp, _dp]),
    "msm_damping": (C.c_int, [_vp]),
    "msm_calc_ekin": (C.c_int, [_vp, C.c_double, _dp, _dp]),
    "msm_scale_velocities": (C.c_int, [_vp, C.c_double, C.c_double]),
    "msm_dyn_damp": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, _ip, _dp]),
    "msm_steepest": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                               C.c_double, _ip, _dp, _dp]),
    "msm_cg": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _ip, _dp]),
    "msm_adjoint": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "msm_set_spline_table": (C.c_int, [_vp, _vp, C.c_int, C.c_int]),
    "msm_fm_set_targets": (C.c_int, [_vp, _dp, _dp, C.c_double]),
    "msm_fm_loss_grad": (C.c_int, [_vp, _dp, _vp, C.c_int, C.c_int]),
}

MSM_OPT_TRACK_DIS = 1
MSM_OPT_KERNEL_TIMING = 2
MSM_OPT_NLIST_KERNEL = 3
MSM_OPT_FORCE_PATH = 4
MSM_OPT_COMPACT_LIST = 5
MSM_OPT_CHUNK_SWEEP = 15
MSM_OPT_TABLE_LDS = 6
MSM_OPT_LT_DENSITY_BLOCK = 7
MSM_OPT_LT_DENSITY_WGS = 8
MSM_OPT_LT_FORCE_BLOCK = 9
MSM_OPT_LT_FORCE_WGS = 10
MSM_OPT_CELL_BRICK = 11
MSM_OPT_BRICK_LANES = 12
MSM_OPT_PASS_KERNEL = 13
MSM_OPT_FORCE_LANES = 14
PASS_KERNELS = ["bricks", "global", "lanes", "lane_pair", "lds_table", "brick_table", "staged"]
KCLASS = ["density", "force", "rebuild", "predictor", "nlist", "adjoint"]

_lib = None


class MSMError(RuntimeError):
    """A libmsm_hip.so entry point returned a non-zero MSM_ERR_* code."""

    def __init__(self, code: int, where: str, msg: str = ""):
        self.code = code
        super().__init__(f"{where}: {ERRORS.get(code, code)} {msg}".rstrip())


def lib() -> C.CDLL:
    """Load libmsm_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def builtin_fn(name: str):
    """Address of a built-in potential callback, for msm_table_pairwise/msm_table_embd."""
    return C.cast(getattr(lib(), name), _vp).value


def check(rc: int, where: str, handle=None):
    if rc != 0:
        msg = ""
        if handle is not None:
            m = lib().msm_last_error(handle)
            msg = m.decode() if m else ""
        raise MSMError(rc, where, msg)
