"""ctypes declarations of include/mplbm.h and the loader of libmplbm.so.

The loader never falls back to anything: if the shared object is missing or
fails to load, the import of the step API raises.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmplbm.so")

OK, ERR_ARG, ERR_HIP, ERR_STATE, ERR_NOMEM, ERR_UNSUPPORTED = range(6)
FP32, FP64, FP16S, BF16S = range(4)
KERNEL_AUTO, KERNEL_FUSED, KERNEL_TWOPASS, KERNEL_COMPACT = range(4)
FACE_PERIODIC, FACE_PRESSURE = 0, 1
MATH_REFERENCE, MATH_FAST = 0, 1
BB_WRITE, BB_PULL = 0, 1
STEP_WRITE_MACRO = 1
ROUND_NEAREST, ROUND_STOCHASTIC = 0, 1
SHIFT_GLOBAL, SHIFT_NODE = 0, 1

STORAGE_NAMES = {"fp32": FP32, "fp64": FP64, "fp16s": FP16S, "bf16s": BF16S}
STORAGE_BYTES = {FP32: 4, FP64: 8, FP16S: 2, BF16S: 2}


class Config(ctypes.Structure):
    _fields_ = [
        ("lx", ctypes.c_int64), ("ly", ctypes.c_int64), ("lz", ctypes.c_int64),
        ("storage", ctypes.c_int32), ("kernel", ctypes.c_int32),
        ("tau", ctypes.c_double), ("G", ctypes.c_double),
        ("carn_star", ctypes.c_int32), ("math", ctypes.c_int32),
        ("T_Tc", ctypes.c_double), ("rhol_spinodal", ctypes.c_double), ("rhog_spinodal", ctypes.c_double),
        ("rho_inject_period", ctypes.c_int64), ("rho_increment", ctypes.c_double),
        ("rhos", ctypes.c_double), ("inject_position", ctypes.c_double),
        ("face_type", ctypes.c_int32 * 6), ("face_rho", ctypes.c_double * 6),
        ("body_force", ctypes.c_double * 3),
        ("rho_ref", ctypes.c_double), ("dev_scale", ctypes.c_double),
        ("tile_y", ctypes.c_int32), ("tile_z", ctypes.c_int32),
        ("seg_x", ctypes.c_int32), ("plane_pad", ctypes.c_int32),
        ("bounce_back", ctypes.c_int32), ("block_threads", ctypes.c_int32),
        ("rounding", ctypes.c_int32), ("shift", ctypes.c_int32),
    ]


class Stats64(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "mass", "mom_x", "mom_y", "mom_z", "fluid_nodes", "liquid_nodes", "vapor_nodes",
        "liquid_rho_sum", "vapor_rho_sum", "liquid_p_sum", "vapor_p_sum",
        "rho_min", "rho_max", "p_min", "p_max", "nonfinite")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Info(ctypes.Structure):
    _fields_ = [
        ("fluid_nodes", ctypes.c_int64), ("cells", ctypes.c_int64),
        ("storage", ctypes.c_int32), ("kernel", ctypes.c_int32),
        ("tile_y", ctypes.c_int32), ("tile_z", ctypes.c_int32), ("seg_x", ctypes.c_int32),
        ("bytes_per_pop", ctypes.c_int32),
        ("rho_ref", ctypes.c_double), ("dev_scale", ctypes.c_double),
        ("device_bytes", ctypes.c_int64),
    ]


class Clusters(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "min_rho", "max_rho", "min_p", "max_p", "avg_liq_press", "avg_vap_press", "saturation",
        "avg_liq_rho", "avg_vap_rho")] + [
        ("liq_cluster_count", ctypes.c_int64), ("gas_cluster_count", ctypes.c_int64)] + [
        (n, ctypes.c_double) for n in ("cluster_order", "wet_coord_num", "ratio_liquid_solid_boundary",
                                       "ratio_liquid_vapor_boundary", "cluster_radius")] + [
        ("fluid_count", ctypes.c_int64), ("liquid_count", ctypes.c_int64), ("vapor_count", ctypes.c_int64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


# every function include/mplbm.h declares, with its ctypes signature
VP = ctypes.c_void_p
SIGNATURES = {
    "mplbm_config_default": (None, [ctypes.POINTER(Config)]),
    "mplbm_create": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.POINTER(VP)]),
    "mplbm_destroy": (None, [VP]),
    "mplbm_last_error": (ctypes.c_char_p, [VP]),
    "mplbm_set_geometry": (ctypes.c_int, [VP, VP]),
    "mplbm_init_fields": (ctypes.c_int, [VP, VP]),
    "mplbm_init_fields_f64": (ctypes.c_int, [VP, VP]),
    "mplbm_step": (ctypes.c_int, [VP, ctypes.c_int64, ctypes.c_int32, ctypes.c_uint32, ctypes.c_int32,
                                  ctypes.c_uint32]),
    "mplbm_step_timed": (ctypes.c_int, [VP, ctypes.c_int64, ctypes.c_int32, ctypes.c_uint32, ctypes.c_int32,
                                        ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]),
    "mplbm_sync": (ctypes.c_int, [VP]),
    "mplbm_read_fields": (ctypes.c_int, [VP, VP, VP, VP]),
    "mplbm_read_fields_f64": (ctypes.c_int, [VP, VP, VP, VP]),
    "mplbm_read_populations_f64": (ctypes.c_int, [VP, VP]),
    "mplbm_write_populations_f64": (ctypes.c_int, [VP, VP]),
    "mplbm_save_checkpoint": (ctypes.c_int, [VP, ctypes.c_char_p, ctypes.c_int64]),
    "mplbm_load_checkpoint": (ctypes.c_int, [VP, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]),
    "mplbm_reduce": (ctypes.c_int, [VP, ctypes.POINTER(Stats64)]),
    "mplbm_set_solid_numbers": (ctypes.c_int, [VP, VP, ctypes.c_int64]),
    "mplbm_calc_cluster": (ctypes.c_int, [VP, ctypes.c_double, ctypes.c_double, ctypes.POINTER(Clusters), VP, VP]),
    "mplbm_place_spheres": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, VP, ctypes.c_int64,
                                           ctypes.c_int64, VP, VP]),
    "mplbm_place_spheres_noalloc": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, VP,
                                                   ctypes.c_int64, ctypes.c_int64, VP, VP]),
    "mplbm_set_tiling": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "mplbm_info_get": (ctypes.c_int, [VP, ctypes.POINTER(Info)]),
    "mplbm_version": (ctypes.c_char_p, []),
}

_lib = None


def load():
    """Load libmplbm.so (built by multiphase_lbm_amd.build); raise if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m multiphase_lbm_amd.build`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib
