"""ctypes binding of the oracle (liboracle.so, built from oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker of the HIP library.  Nothing in
the multiphase_lbm_amd package imports it.

State arrays use the reference layout: fields [lx][ly][lz] (z fastest),
populations [lx][ly][lz][19] (include/lbm_model.cuh:18-22).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# flat parameter vector indices (oracle.c ORC_P_*)
P_LX, P_LY, P_LZ, P_TAU, P_G, P_CARN_STAR, P_T_TC = 0, 1, 2, 3, 4, 5, 6
P_RHOL_SP, P_RHOG_SP, P_INJ_PERIOD, P_RHO_INC, P_RHOS, P_INJ_POS = 7, 8, 9, 10, 11, 12
P_FACE_TYPE, P_FACE_RHO, P_BODY, P_COUNT = 13, 19, 25, 28

EX = np.array([0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0])
EY = np.array([0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1])
EZ = np.array([0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1])


@dataclass
class Params:
    """Step parameters; defaults follow the reference constructor
    (src/multiphase_lbm.cu:35-80) except rhos, which it leaves unset."""
    lx: int
    ly: int
    lz: int
    tau: float = 1.0
    G: float = -1.0
    carn_star: bool = False
    T_Tc: float = 0.7
    rhol_spinodal: float = 0.2726
    rhog_spinodal: float = 0.0484
    rho_inject_period: int = 100
    rho_increment: float = 0.0
    rhos: float = 0.0
    inject_position: float = 1.0
    face_type: tuple = (0, 0, 0, 0, 0, 0)
    face_rho: tuple = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    body_force: tuple = (0.0, 0.0, 0.0)

    def vector(self) -> np.ndarray:
        p = np.zeros(P_COUNT, dtype=np.float64)
        p[P_LX], p[P_LY], p[P_LZ] = self.lx, self.ly, self.lz
        p[P_TAU], p[P_G], p[P_CARN_STAR], p[P_T_TC] = self.tau, self.G, float(self.carn_star), self.T_Tc
        p[P_RHOL_SP], p[P_RHOG_SP] = self.rhol_spinodal, self.rhog_spinodal
        p[P_INJ_PERIOD], p[P_RHO_INC] = self.rho_inject_period, self.rho_increment
        p[P_RHOS], p[P_INJ_POS] = self.rhos, self.inject_position
        p[P_FACE_TYPE:P_FACE_TYPE + 6] = self.face_type
        p[P_FACE_RHO:P_FACE_RHO + 6] = self.face_rho
        p[P_BODY:P_BODY + 3] = self.body_force
        return p


_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.c_void_p
        for s, ct in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
            getattr(L, f"orc_{s}_init").argtypes = [dp] * 8
            getattr(L, f"orc_{s}_run").argtypes = [dp] * 9 + [ctypes.c_long, ctypes.c_long, ctypes.c_int, ctypes.c_uint]
            getattr(L, f"orc_{s}_stream").argtypes = [dp] * 6
            getattr(L, f"orc_{s}_density").argtypes = [dp] * 6 + [ctypes.c_long, ctypes.c_int, ctypes.c_uint]
            getattr(L, f"orc_{s}_force").argtypes = [dp] * 4
            getattr(L, f"orc_{s}_collide").argtypes = [dp] * 6
            for fn in ("psi", "press"):
                f = getattr(L, f"orc_{s}_{fn}")
                f.argtypes = [dp, ct]
                f.restype = ct
        L.orc_exp.argtypes = [ctypes.c_double]
        L.orc_exp.restype = ctypes.c_double
        L.orc_tables.argtypes = [dp] * 6
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _suffix(dtype) -> str:
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


class State:
    """Oracle lattice state: two AoS lattices plus rho, p, psi, F, u."""

    def __init__(self, params: Params, is_fluid: np.ndarray, rho0: np.ndarray, dtype=np.float32):
        self.params = params
        self.prm = params.vector()
        self.dtype = np.dtype(dtype)
        shape = (params.lx, params.ly, params.lz)
        assert is_fluid.shape == shape and rho0.shape == shape
        self.mask = np.ascontiguousarray(is_fluid.astype(np.uint8))
        self.rho = np.ascontiguousarray(rho0.astype(self.dtype))
        self.psi = np.zeros(shape, self.dtype)
        self.p = np.zeros(shape, self.dtype)
        self.fA = np.zeros(shape + (19,), self.dtype)
        self.fB = np.zeros(shape + (19,), self.dtype)
        self.F = np.zeros(shape + (3,), self.dtype)
        self.u = np.zeros(shape + (3,), self.dtype)
        self.sfx = _suffix(self.dtype)
        getattr(lib(), f"orc_{self.sfx}_init")(
            _p(self.prm), _p(self.mask), _p(self.rho), _p(self.psi), _p(self.p),
            _p(self.fA), _p(self.fB), _p(self.F))
        self.step = 0

    def run(self, nsteps: int, reverse: int = 1, inject_type: int = 0):
        """Advance nsteps steps; afterwards fA holds the newest populations and
        u the shifted velocity of the last step."""
        getattr(lib(), f"orc_{self.sfx}_run")(
            _p(self.prm), _p(self.mask), _p(self.fA), _p(self.fB), _p(self.rho), _p(self.p),
            _p(self.psi), _p(self.F), _p(self.u), self.step, nsteps, reverse, inject_type)
        self.step += nsteps
        return self


def stream(params: Params, mask, fA, rho=None, F=None):
    """Stream + boundary only (do_streaming + treat_boundary)."""
    dt = fA.dtype
    fB = np.zeros_like(fA)
    shape = fA.shape[:3]
    rho = np.ones(shape, dt) if rho is None else np.ascontiguousarray(rho, dt)
    F = np.zeros(shape + (3,), dt) if F is None else np.ascontiguousarray(F, dt)
    getattr(lib(), f"orc_{_suffix(dt)}_stream")(
        _p(params.vector()), _p(np.ascontiguousarray(mask, np.uint8)), _p(np.ascontiguousarray(fA)),
        _p(fB), _p(rho), _p(F))
    return fB


def force(params: Params, mask, psi):
    dt = psi.dtype
    F = np.zeros(psi.shape + (3,), dt)
    getattr(lib(), f"orc_{_suffix(dt)}_force")(
        _p(params.vector()), _p(np.ascontiguousarray(mask, np.uint8)), _p(np.ascontiguousarray(psi)), _p(F))
    return F


def collide(params: Params, mask, fB, rho, F):
    """BGK with velocity shift in place on a copy of fB; returns (f*, u)."""
    dt = fB.dtype
    out = np.ascontiguousarray(fB).copy()
    u = np.zeros(fB.shape[:3] + (3,), dt)
    Fc = np.ascontiguousarray(F, dt).copy()
    getattr(lib(), f"orc_{_suffix(dt)}_collide")(
        _p(params.vector()), _p(np.ascontiguousarray(mask, np.uint8)), _p(out),
        _p(np.ascontiguousarray(rho, dt)), _p(Fc), _p(u))
    return out, u


def psi(params: Params, rho: float, dtype=np.float64) -> float:
    return getattr(lib(), f"orc_{_suffix(dtype)}_psi")(_p(params.vector()), rho)


def press(params: Params, rho: float, dtype=np.float64) -> float:
    return getattr(lib(), f"orc_{_suffix(dtype)}_press")(_p(params.vector()), rho)


def exp(x: float) -> float:
    return lib().orc_exp(x)


def tables():
    ex, ey, ez, opp = (np.zeros(19, np.int32) for _ in range(4))
    w32, w64 = np.zeros(19), np.zeros(19)
    lib().orc_tables(_p(ex), _p(ey), _p(ez), _p(opp), _p(w32), _p(w64))
    return ex, ey, ez, opp, w32, w64
