"""The reference-named entry points (include/iteration_host_compat.h) on a
machine without a GPU: arity and error behaviour only.

The reference's cudaErrorCheck prints and exits (include/iteration_host.h:6-15);
the compat layer does the same by default and prints and returns after
mplbm_compat_set_abort(0).  Errors raised before any device work are used
here; the GPU test of the three calls is tests/test_gpu_compat.py."""
import ctypes
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "multiphase_lbm_amd", "libmplbm.so")
STATE_ERROR = 3  # MPLBM_ERR_STATE


def _lib():
    L = ctypes.CDLL(LIB)
    L.multiphaselbm_iteration.argtypes = [ctypes.c_long, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint]
    L.multiphaselbm_iteration.restype = None
    L.multiphaselbm_initialization.argtypes = [ctypes.c_void_p] * 7
    L.multiphaselbm_initialization.restype = None
    L.mplbm_compat_set_abort.argtypes = [ctypes.c_int]
    L.mplbm_compat_status.restype = ctypes.c_int
    return L


def test_errors_return_after_set_abort_0(capfd):
    L = _lib()
    L.mplbm_compat_set_abort(0)
    L.mplbm_compat_release()
    L.multiphaselbm_iteration(0, 10, 1, 0)
    assert L.mplbm_compat_status() == STATE_ERROR
    L.multiphaselbm_initialization(None, None, None, None, None, None, None)
    assert L.mplbm_compat_status() != 0
    err = capfd.readouterr().err
    assert "multiphaselbm_initialization must come first" in err and "null" in err
    L.mplbm_compat_set_abort(1)


def test_errors_exit_with_the_status_by_default():
    code = ("import ctypes; L = ctypes.CDLL(%r); "
            "L.multiphaselbm_iteration.argtypes = [ctypes.c_long, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint]; "
            "L.multiphaselbm_iteration(0, 10, 1, 0); print('not reached')" % LIB)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == STATE_ERROR
    assert "not reached" not in r.stdout
    assert "must come first" in r.stderr
