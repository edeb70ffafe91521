"""The C-ABI libraries load on a machine without a GPU and export every
function their public headers declare (no compute calls here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(mplbm_\w+|transfer_parameter|multiphaselbm_\w+)\s*\(", text)
    return sorted(set(n for n in names if not n.endswith("_t")))


def _exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


@pytest.mark.parametrize("header,lib", [("mplbm.h", "libmplbm.so"), ("mplbm_host.h", "libmplbm_host.so"),
                                        ("iteration_host_compat.h", "libmplbm.so"),
                                        ("mplbm_ensemble.h", "libmplbm.so")])
def test_library_exports_every_declared_symbol(header, lib):
    path = os.path.join(ROOT, "multiphase_lbm_amd", lib)
    assert os.path.exists(path), f"{lib} not built"
    declared = _declared(header)
    assert declared, header
    missing = [n for n in declared if n not in _exported(path)]
    assert not missing, f"{lib} lacks {missing}"
    L = ctypes.CDLL(path)
    for n in declared:
        getattr(L, n)


def test_python_binding_covers_the_abi():
    from multiphase_lbm_amd import _abi, harness
    assert set(_abi.SIGNATURES) == set(_declared("mplbm.h"))
    assert set(harness.SIGNATURES) == set(_declared("mplbm_host.h"))
    from multiphase_lbm_amd import ensemble
    assert set(ensemble.SIGNATURES) == set(_declared("mplbm_ensemble.h")) - set(_declared("mplbm.h"))


def test_config_layout_matches_c():
    """The ctypes mirror of mplbm_config has the C struct's size (checked
    against the library's default-filled struct)."""
    from multiphase_lbm_amd import _abi
    assert ctypes.sizeof(_abi.Config) == 256 and _abi.Config.rounding.offset == 248
    cfg = _abi.Config()
    _abi.load().mplbm_config_default(ctypes.byref(cfg))
    assert (cfg.lx, cfg.ly, cfg.lz) == (400, 400, 400)
    assert cfg.tau == 1.0 and cfg.G == -1.0 and cfg.rho_inject_period == 100
    assert abs(cfg.rhol_spinodal - 0.2726) < 1e-15 and cfg.inject_position == 1.0


def test_create_reports_errors_without_exit():
    """Invalid configurations return a status and a message (the reference
    exits the process, include/iteration_host.h:6-15)."""
    import multiphase_lbm_amd as mp
    with pytest.raises(mp.MplbmError, match="tau"):
        mp.Context(8, 8, 8, tau=0.5)
    with pytest.raises(mp.MplbmError, match="faces"):
        mp.Context(8, 8, 8, face_type=(1, 0, 0, 0, 0, 0))
    with pytest.raises(mp.MplbmError, match="2\\^32"):
        mp.Context(2048, 2048, 2048)


def test_compact_tile_limits_report_errors():
    """A compact workgroup holds one interior node per thread; tiles beyond
    that or beyond the LDS budget are refused before any device work."""
    import multiphase_lbm_amd as mp
    with pytest.raises(mp.MplbmError, match="below the tile"):
        mp.Context(64, 64, 64, kernel="compact", tile_y=16, tile_z=32, block_threads=256)
    with pytest.raises(mp.MplbmError, match="block_threads must be"):
        mp.Context(64, 64, 64, kernel="compact", block_threads=300)
    with pytest.raises(mp.MplbmError, match="exceeds 512 interior nodes"):
        mp.Context(64, 64, 64, storage="fp64", kernel="compact", tile_y=24, tile_z=24)


def test_place_spheres_refuses_bad_input_before_device_work():
    """The device voxeliser checks its arguments on the host: a sphere whose
    float-stepped axis would not end or would not fit the per-axis sample
    buffer, non-finite values, and arrays of the wrong type."""
    import numpy as np
    import multiphase_lbm_amd as mp
    m = np.ones((8, 8, 8), np.uint8)
    with pytest.raises(mp.MplbmError, match="outside the voxeliser's range"):
        mp.place_spheres(m, [[4.0, 4.0, 4.0, 600.0]])
    with pytest.raises(mp.MplbmError, match="outside the voxeliser's range"):
        mp.place_spheres(m, [[4.0, float("nan"), 4.0, 2.0]])
    with pytest.raises(mp.MplbmError, match="outside the voxeliser's range"):
        mp.place_spheres(m, [[2.0 ** 24, 4.0, 4.0, 2.0]])
    with pytest.raises(ValueError):
        mp.place_spheres(m.astype(np.int32), [[4.0, 4.0, 4.0, 2.0]])
    with pytest.raises(ValueError):
        mp.place_spheres(m, [[4.0, 4.0, 4.0, 2.0]], solid_number=np.zeros((8, 8, 8), np.int32))
    assert (m == 1).all()
