"""ctypes view of the host harness (include/mplbm_host.h, libmplbm_host.so):
input parser, sphere voxeliser, positions/sample/density loaders, seeded
packings, cluster statistics and the VTK writer of the reference program
(src/multiphase_lbm.cu), plus the whole CLI driver."""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(PKG, "libmplbm_host.so")
CLI = os.path.join(PKG, "multiphaselbm")

VP = ctypes.c_void_p
I64 = ctypes.c_int64


class Input(ctypes.Structure):
    _fields_ = [
        ("lx", I64), ("ly", I64), ("lz", I64),
        ("result_folder", ctypes.c_char * 256),
        ("num_threads", ctypes.c_int32),
        ("step", I64), ("step_max", I64),
        ("vtkPeriod", I64), ("vtrPeriod", I64), ("tecplotPeriod", I64), ("statPeriod", I64),
        ("save_vtk_ascii", ctypes.c_int32), ("save_vtr_binary", ctypes.c_int32),
        ("save_tecplot_ascii", ctypes.c_int32),
        ("rho0", ctypes.c_double), ("rhos", ctypes.c_double), ("G", ctypes.c_double),
        ("IniPerturbRate", ctypes.c_double), ("tau", ctypes.c_double), ("res", ctypes.c_double),
        ("rho_inject_period", I64), ("rho_increment", ctypes.c_double),
        ("inject_density", ctypes.c_int32), ("inject_type", ctypes.c_uint32),
        ("inject_position", ctypes.c_double), ("carn_star", ctypes.c_int32),
        ("T_Tc", ctypes.c_double), ("rhol_spinodal", ctypes.c_double), ("rhog_spinodal", ctypes.c_double),
        ("read_densities", ctypes.c_int32), ("densities_filename", ctypes.c_char * 256),
        ("hysteresis", ctypes.c_int32), ("max_saturation", ctypes.c_double), ("min_saturation", ctypes.c_double),
        ("erosion", ctypes.c_int32), ("target_density", ctypes.c_double),
        ("storage", ctypes.c_int32), ("kernel", ctypes.c_int32), ("math", ctypes.c_int32),
        ("face_type", ctypes.c_int32 * 6), ("face_rho", ctypes.c_double * 6),
        ("body_force", ctypes.c_double * 3), ("unknown_tokens", ctypes.c_int32),
        ("host_statistics", ctypes.c_int32), ("read_checkpoint", ctypes.c_int32),
        ("checkpointPeriod", I64), ("checkpoint_filename", ctypes.c_char * 256),
        ("rounding", ctypes.c_int32), ("device_voxeliser", ctypes.c_int32), ("shift", ctypes.c_int32),
        ("saturation_cycle", ctypes.c_int32),
    ]

    def as_dict(self):
        out = {}
        for n, _ in self._fields_:
            v = getattr(self, n)
            if isinstance(v, bytes):
                v = v.decode()
            elif hasattr(v, "_length_"):
                v = list(v)
            out[n] = v
        return out


class ClusterStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "min_rho", "max_rho", "min_p", "max_p", "avg_liq_press", "avg_vap_press", "saturation",
        "avg_liq_rho", "avg_vap_rho")] + [("liq_cluster_count", I64), ("gas_cluster_count", I64)] + [
        (n, ctypes.c_double) for n in ("cluster_order", "wet_coord_num", "ratio_liquid_solid_boundary",
                                       "ratio_liquid_vapor_boundary", "cluster_radius")] + [
        ("fluid_count", I64), ("liquid_count", I64), ("vapor_count", I64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


SIGNATURES = {
    "mplbm_host_last_error": (ctypes.c_char_p, []),
    "mplbm_host_input_default": (None, [ctypes.POINTER(Input)]),
    "mplbm_host_parse": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(Input)]),
    "mplbm_host_place_sphere": (None, [VP, VP, I64, I64, I64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_float, I64]),
    "mplbm_host_read_positions": (I64, [ctypes.c_char_p, ctypes.c_float, VP, VP, I64, I64, I64, I64]),
    "mplbm_host_read_positions_device": (I64, [ctypes.c_char_p, ctypes.c_float, VP, VP, I64, I64, I64, I64]),
    "mplbm_host_rho_init": (None, [VP, I64, I64, I64, ctypes.c_float, ctypes.c_float]),
    "mplbm_host_read_densities": (I64, [ctypes.c_char_p, VP, I64, I64, I64]),
    "mplbm_host_read_sample": (ctypes.c_int, [ctypes.c_char_p, VP, I64, I64, I64]),
    "mplbm_host_generate_packing": (I64, [VP, I64, I64, I64, ctypes.c_uint64, I64, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_int32, VP, I64]),
    "mplbm_host_write_positions": (ctypes.c_int, [ctypes.c_char_p, VP, I64]),
    "mplbm_host_cluster_stats": (ctypes.c_int, [VP, VP, VP, VP, I64, I64, I64, ctypes.c_double, ctypes.c_double,
                                                I64, VP, VP, ctypes.POINTER(ClusterStats)]),
    "mplbm_host_write_vtk": (ctypes.c_int, [ctypes.c_char_p, VP, VP, VP, VP, VP, VP, I64, I64, I64]),
    "mplbm_host_write_vtr": (ctypes.c_int, [ctypes.c_char_p, VP, VP, VP, VP, VP, VP, I64, I64, I64]),
    "mplbm_host_read_data": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(Input), VP, VP, I64,
                                            ctypes.POINTER(I64)]),
    "mplbm_host_run": (ctypes.c_int, [ctypes.c_char_p]),
}

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is not built: run `python -m multiphase_lbm_amd.build`")
        lib = ctypes.CDLL(HOST_LIB)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _err(msg):
    return RuntimeError(f"{msg}: {load().mplbm_host_last_error().decode()}")


def parse_input(path) -> dict:
    inp = Input()
    if load().mplbm_host_parse(str(path).encode(), ctypes.byref(inp)) != 0:
        raise _err("parse")
    return inp.as_dict()


def read_data(path):
    """The reference's read_data (tokens, domain, density initialisation,
    geometry, erosion): returns (input dict, is_fluid uint8 [x][y][z],
    initial rho float32 [x][y][z], number of eroded solid nodes)."""
    lib = load()
    inp = Input()
    if lib.mplbm_host_read_data(str(path).encode(), ctypes.byref(inp), None, None, 0, None) != 0:
        raise _err("read_data")
    shape = (inp.lx, inp.ly, inp.lz)
    fluid = np.empty(shape, np.uint8)
    rho = np.empty(shape, np.float32)
    eroded = I64()
    if lib.mplbm_host_read_data(str(path).encode(), ctypes.byref(inp), _p(fluid), _p(rho), fluid.size,
                                ctypes.byref(eroded)) != 0:
        raise _err("read_data")
    return inp.as_dict(), fluid, rho, eroded.value


def place_sphere(is_fluid, x, y, z, R, solid_number=None, solid_id=0):
    lx, ly, lz = is_fluid.shape
    assert is_fluid.dtype == np.uint8 and is_fluid.flags.c_contiguous
    load().mplbm_host_place_sphere(_p(is_fluid), _p(solid_number), lx, ly, lz, x, y, z, R, solid_id)


def read_positions(path, shape, res=1.0, is_fluid=None, solid_number=None, first_id=1, device=False):
    """read_positions; device=True voxelises on the GPU (mplbm_place_spheres)."""
    lx, ly, lz = shape
    if is_fluid is None:
        is_fluid = np.ones(shape, np.uint8)
    if solid_number is None:
        solid_number = np.zeros(shape, np.int64)
    fn = load().mplbm_host_read_positions_device if device else load().mplbm_host_read_positions
    n = fn(str(path).encode(), res, _p(is_fluid), _p(solid_number), lx, ly, lz, first_id)
    if n < 0:
        raise _err("read_positions")
    return is_fluid, solid_number, n


def rho_init(shape, rho0, ini_perturb_rate):
    rho = np.empty(shape, np.float32)
    load().mplbm_host_rho_init(_p(rho), *shape, rho0, ini_perturb_rate)
    return rho


def read_densities(path, shape):
    rho = np.zeros(shape, np.float32)
    n = load().mplbm_host_read_densities(str(path).encode(), _p(rho), *shape)
    if n < 0:
        raise _err("read_densities")
    return rho, n


def read_sample(path, shape):
    m = np.ones(shape, np.uint8)
    if load().mplbm_host_read_sample(str(path).encode(), _p(m), *shape) != 0:
        raise _err("read_sample")
    return m


def generate_packing(shape, seed, n_max, rmin, rmax, porosity, nonoverlap=False):
    m = np.empty(shape, np.uint8)
    cap = int(min(n_max, 10_000_000))
    sph = np.zeros((cap, 4), np.float64)
    n = load().mplbm_host_generate_packing(_p(m), *shape, seed, n_max, rmin, rmax, porosity, int(nonoverlap),
                                           _p(sph), cap)
    if n < 0:
        raise _err("generate_packing")
    return m, sph[:min(n, cap)]


def write_positions(path, spheres):
    s = np.ascontiguousarray(spheres, np.float64)
    if load().mplbm_host_write_positions(str(path).encode(), _p(s), len(s)) != 0:
        raise _err("write_positions")


def cluster_stats(is_fluid, rho, pres, solid_number=None, rhol=0.2726, rhog=0.0484, solid_count=0,
                  want_labels=False):
    shape = is_fluid.shape
    liq = np.zeros(shape, np.int64) if want_labels else None
    gas = np.zeros(shape, np.int64) if want_labels else None
    s = ClusterStats()
    m = np.ascontiguousarray(is_fluid, np.uint8)
    r = np.ascontiguousarray(rho, np.float32)
    p = np.ascontiguousarray(pres, np.float32)
    sn = None if solid_number is None else np.ascontiguousarray(solid_number, np.int64)
    if load().mplbm_host_cluster_stats(_p(m), _p(r), _p(p), _p(sn), *shape, rhol, rhog, solid_count, _p(liq),
                                       _p(gas), ctypes.byref(s)) != 0:
        raise _err("cluster_stats")
    out = s.as_dict()
    if want_labels:
        return out, liq, gas
    return out


def write_vtr(path, is_fluid, rho, pres, liq=None, gas=None, solid_number=None):
    shape = is_fluid.shape
    m = np.ascontiguousarray(is_fluid, np.uint8)
    r = np.ascontiguousarray(rho, np.float32)
    p = np.ascontiguousarray(pres, np.float32)
    li = None if liq is None else np.ascontiguousarray(liq, np.int64)
    ga = None if gas is None else np.ascontiguousarray(gas, np.int64)
    sn = None if solid_number is None else np.ascontiguousarray(solid_number, np.int64)
    if load().mplbm_host_write_vtr(str(path).encode(), _p(m), _p(r), _p(p), _p(li), _p(ga), _p(sn), *shape) != 0:
        raise _err("write_vtr")


def read_vtr(path):
    """Arrays of a file written by write_vtr, as [x][y][z] numpy arrays, and
    the coordinates."""
    import re
    data = open(path, "rb").read()
    head, _, rest = data.partition(b"<AppendedData encoding=\"raw\">")
    base = rest.index(b"_") + 1
    ext = [int(v) for v in re.search(rb'WholeExtent="([^"]+)"', head).group(1).split()]
    shape = (ext[1] + 1, ext[3] + 1, ext[5] + 1)
    out = {}
    for m in re.finditer(rb'<DataArray type="(\w+)" Name="(\w+)" format="appended" offset="(\d+)"', head):
        typ, name, off = m.group(1).decode(), m.group(2).decode(), int(m.group(3))
        n = int(np.frombuffer(rest, np.uint64, 1, base + off)[0])
        dt = {"Float32": np.float32, "Int64": np.int64, "Int32": np.int32}[typ]
        a = np.frombuffer(rest, dt, n // np.dtype(dt).itemsize, base + off + 8)
        if name in ("x", "y", "z"):
            out[name] = a.copy()
        else:
            out[name] = a.reshape(shape[::-1]).transpose(2, 1, 0).copy()  # x fastest in the file
    return out


def write_vtk(path, is_fluid, rho, pres, liq=None, gas=None, solid_number=None):
    shape = is_fluid.shape
    m = np.ascontiguousarray(is_fluid, np.uint8)
    r = np.ascontiguousarray(rho, np.float32)
    p = np.ascontiguousarray(pres, np.float32)
    if load().mplbm_host_write_vtk(str(path).encode(), _p(m), _p(r), _p(p), _p(liq), _p(gas), _p(solid_number),
                                   *shape) != 0:
        raise _err("write_vtk")
