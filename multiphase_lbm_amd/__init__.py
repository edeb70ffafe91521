"""multiphase_lbm_amd — MI355X-native D3Q19 Shan-Chen multiphase LBM step.

Python view of the C ABI in include/mplbm.h (libmplbm.so).  The device work
runs only in the HIP library; this module moves host arrays across the ABI.
Arrays use the reference layout [x][y][z] with z fastest
(include/lbm_model.cuh:18-22) and populations [x][y][z][19].
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi
from ._abi import (BF16S, FACE_PERIODIC, FACE_PRESSURE, FP16S, FP32, FP64, KERNEL_AUTO, KERNEL_COMPACT,
                   KERNEL_FUSED, KERNEL_TWOPASS, ROUND_NEAREST, ROUND_STOCHASTIC, SHIFT_GLOBAL, SHIFT_NODE,
                   STEP_WRITE_MACRO, STORAGE_BYTES, STORAGE_NAMES)

__all__ = ["Context", "MplbmError", "FP32", "FP64", "FP16S", "BF16S", "KERNEL_AUTO", "KERNEL_FUSED",
           "KERNEL_TWOPASS", "KERNEL_COMPACT", "FACE_PERIODIC", "FACE_PRESSURE", "ROUND_NEAREST",
           "ROUND_STOCHASTIC", "SHIFT_GLOBAL", "SHIFT_NODE", "default_config", "place_spheres"]

KERNEL_NAMES = {"auto": KERNEL_AUTO, "fused": KERNEL_FUSED, "twopass": KERNEL_TWOPASS, "compact": KERNEL_COMPACT}


class MplbmError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"mplbm status {status}: {msg}")
        self.status = status


def default_config() -> _abi.Config:
    cfg = _abi.Config()
    _abi.load().mplbm_config_default(ctypes.byref(cfg))
    return cfg


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def place_spheres(is_fluid, spheres, first_id=1, solid_number=None):
    """Voxelise spheres (rows x, y, z, R) on the device into is_fluid
    (uint8 [x][y][z], updated in place) and solid_number (int64 [x][y][z]
    or None), as the reference's place_sphere for each sphere in order
    (mplbm_place_spheres)."""
    lib = _abi.load()
    if is_fluid.dtype != np.uint8 or is_fluid.ndim != 3 or not is_fluid.flags.c_contiguous:
        raise ValueError("is_fluid must be a C-contiguous uint8 [x][y][z] array")
    if solid_number is not None and (solid_number.dtype != np.int64 or solid_number.shape != is_fluid.shape
                                     or not solid_number.flags.c_contiguous):
        raise ValueError("solid_number must be a C-contiguous int64 array of is_fluid's shape")
    sph = np.ascontiguousarray(np.asarray(spheres, np.float32).reshape(-1, 4))
    st = lib.mplbm_place_spheres(*[int(v) for v in is_fluid.shape], _ptr(sph), int(sph.shape[0]), int(first_id),
                                 _ptr(is_fluid), _ptr(solid_number))
    if st != 0:
        raise MplbmError(st, lib.mplbm_last_error(None).decode())


class Context:
    """One simulation context on the current HIP device.

    Keyword parameters are fields of mplbm_config (tau, G, carn_star, T_Tc,
    rhol_spinodal, rhog_spinodal, rho_inject_period, rho_increment, rhos,
    inject_position, face_type, face_rho, body_force, rho_ref, dev_scale,
    tile_y, tile_z, seg_x, math, rounding, shift).  storage "fp32fast" is FP32
    storage with math="fast"; "fp16sr"/"bf16sr" add rounding="stochastic" and
    "fp16sn"/"bf16sn" shift="node", "fp16srn"/"bf16srn" both, to the 16-bit
    forms."""

    def __init__(self, lx, ly, lz, storage="fp32", kernel="auto", **params):
        self.lib = _abi.load()
        cfg = default_config()
        cfg.lx, cfg.ly, cfg.lz = int(lx), int(ly), int(lz)
        if storage == "fp32fast":
            storage, params = "fp32", dict(params, math="fast")
        elif storage in ("fp16sr", "bf16sr"):
            storage, params = storage[:-1], dict(params, rounding="stochastic")
        elif storage in ("fp16sn", "bf16sn"):
            storage, params = storage[:-1], dict(params, shift="node")
        elif storage in ("fp16srn", "bf16srn"):
            storage, params = storage[:-2], dict(params, rounding="stochastic", shift="node")
        cfg.storage = STORAGE_NAMES[storage] if isinstance(storage, str) else int(storage)
        cfg.kernel = KERNEL_NAMES[kernel] if isinstance(kernel, str) else int(kernel)
        for k, v in params.items():
            if k in ("face_type", "face_rho", "body_force"):
                arr = getattr(cfg, k)
                for i, x in enumerate(v):
                    arr[i] = x
            elif k == "carn_star":
                cfg.carn_star = int(bool(v))
            elif k == "math":
                cfg.math = {"reference": 0, "fast": 1}[v] if isinstance(v, str) else int(v)
            elif k == "rounding":
                cfg.rounding = {"nearest": 0, "stochastic": 1}[v] if isinstance(v, str) else int(v)
            elif k == "shift":
                cfg.shift = {"global": 0, "node": 1}[v] if isinstance(v, str) else int(v)
            else:
                if not hasattr(cfg, k):
                    raise TypeError(f"unknown parameter {k}")
                setattr(cfg, k, v)
        self.cfg = cfg
        self.shape = (cfg.lx, cfg.ly, cfg.lz)
        h = ctypes.c_void_p()
        st = self.lib.mplbm_create(ctypes.byref(cfg), ctypes.byref(h))
        if st != _abi.OK:
            raise MplbmError(st, self.lib.mplbm_last_error(None).decode())
        self.h = h

    # -- plumbing ---------------------------------------------------------
    def _check(self, st):
        if st != _abi.OK:
            raise MplbmError(st, self.lib.mplbm_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.mplbm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- API ---------------------------------------------------------------
    def set_geometry(self, is_fluid):
        m = np.ascontiguousarray(np.asarray(is_fluid).astype(np.uint8))
        if m.shape != self.shape:
            raise ValueError(f"mask shape {m.shape} != {self.shape}")
        self._check(self.lib.mplbm_set_geometry(self.h, _ptr(m)))

    def init_fields(self, rho):
        rho = np.asarray(rho)
        if rho.shape != self.shape:
            raise ValueError(f"rho shape {rho.shape} != {self.shape}")
        if rho.dtype == np.float32:
            r = np.ascontiguousarray(rho)
            self._check(self.lib.mplbm_init_fields(self.h, _ptr(r)))
        else:
            r = np.ascontiguousarray(rho, dtype=np.float64)
            self._check(self.lib.mplbm_init_fields_f64(self.h, _ptr(r)))

    def step(self, step=0, nsteps=1, reverse=1, inject_type=0, write_macro=False):
        flags = STEP_WRITE_MACRO if write_macro else 0
        self._check(self.lib.mplbm_step(self.h, int(step), int(reverse), int(inject_type), int(nsteps), flags))

    def step_timed(self, step=0, nsteps=1, reverse=1, inject_type=0, write_macro=False) -> float:
        flags = STEP_WRITE_MACRO if write_macro else 0
        ms = ctypes.c_double()
        self._check(self.lib.mplbm_step_timed(self.h, int(step), int(reverse), int(inject_type), int(nsteps),
                                              flags, ctypes.byref(ms)))
        return ms.value

    def sync(self):
        self._check(self.lib.mplbm_sync(self.h))

    def read_fields(self, dtype=np.float64):
        """(rho, p, u) of the last macro write; u is [x][y][z][3]."""
        rho = np.empty(self.shape, dtype)
        p = np.empty(self.shape, dtype)
        u = np.empty(self.shape + (3,), dtype)
        fn = self.lib.mplbm_read_fields_f64 if np.dtype(dtype) == np.float64 else self.lib.mplbm_read_fields
        self._check(fn(self.h, _ptr(rho), _ptr(p), _ptr(u)))
        return rho, p, u

    def read_populations(self):
        f = np.empty(self.shape + (19,), np.float64)
        self._check(self.lib.mplbm_read_populations_f64(self.h, _ptr(f)))
        return f

    def write_populations(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        if f.shape != self.shape + (19,):
            raise ValueError("population array shape mismatch")
        self._check(self.lib.mplbm_write_populations_f64(self.h, _ptr(f)))

    def save_checkpoint(self, path, step):
        self._check(self.lib.mplbm_save_checkpoint(self.h, str(path).encode(), int(step)))

    def load_checkpoint(self, path) -> int:
        """Restore populations saved by save_checkpoint; returns the saved step."""
        step = ctypes.c_int64()
        self._check(self.lib.mplbm_load_checkpoint(self.h, str(path).encode(), ctypes.byref(step)))
        return step.value

    def reduce(self) -> dict:
        s = _abi.Stats64()
        self._check(self.lib.mplbm_reduce(self.h, ctypes.byref(s)))
        return s.as_dict()

    def set_solid_numbers(self, solid_number, solid_count):
        """Solid numbering ([x][y][z], 0 at fluid cells) for calc_cluster."""
        sn = None
        if solid_number is not None:
            sn = np.ascontiguousarray(solid_number, np.int64)
            if sn.shape != self.shape:
                raise ValueError(f"solid_number shape {sn.shape} != {self.shape}")
        self._check(self.lib.mplbm_set_solid_numbers(self.h, None if sn is None else _ptr(sn), int(solid_count)))

    def calc_cluster(self, rhol, rhog, want_labels=False):
        """calc_cluster statistics of the last macro write on the device;
        with want_labels also the LiqClusters / GasClusters arrays."""
        s = _abi.Clusters()
        liq = np.zeros(self.shape, np.int32) if want_labels else None
        gas = np.zeros(self.shape, np.int32) if want_labels else None
        self._check(self.lib.mplbm_calc_cluster(self.h, float(rhol), float(rhog), ctypes.byref(s),
                                                None if liq is None else _ptr(liq),
                                                None if gas is None else _ptr(gas)))
        if want_labels:
            return s.as_dict(), liq, gas
        return s.as_dict()

    def set_tiling(self, tile_y=0, tile_z=0, seg_x=0):
        self._check(self.lib.mplbm_set_tiling(self.h, int(tile_y), int(tile_z), int(seg_x)))

    def info(self) -> dict:
        i = _abi.Info()
        self._check(self.lib.mplbm_info_get(self.h, ctypes.byref(i)))
        return {n: getattr(i, n) for n, _ in i._fields_}
