"""mplbm_place_spheres_noalloc (the sphere voxeliser that allocates nothing:
device-resident mask and solid numbers, scratch in a static device buffer)
against mplbm_place_spheres on host arrays, bit for bit, on two calls in a
row, the second on a smaller domain (the scratch keeps the first call's
sphere numbers, which the second must clear); oversized calls are refused
before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import multiphase_lbm_amd as mp  # noqa: E402
from multiphase_lbm_amd import _abi  # noqa: E402

H2D, D2H = 1, 2


class _Dev:
    """A device copy of a numpy array (the test's own allocation)."""

    hip = None

    def __init__(self, a):
        if _Dev.hip is None:
            _Dev.hip = ctypes.CDLL("libamdhip64.so")
        self.shape, self.dtype, self.nbytes = a.shape, a.dtype, a.nbytes
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(a.nbytes)) == 0
        assert self.hip.hipMemcpy(self.ptr, ctypes.c_void_p(a.ctypes.data), ctypes.c_size_t(a.nbytes), H2D) == 0

    def get(self):
        out = np.empty(self.shape, self.dtype)
        assert self.hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), self.ptr, ctypes.c_size_t(self.nbytes), D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _case(shape, nsph, seed):
    rng = np.random.default_rng(seed)
    # centres in the box and up to 3 cells beyond it (periodic wrap), radii 1-7.5
    centres = rng.uniform(-3, 3, (nsph, 3)) + rng.uniform(0, 1, (nsph, 3)) * np.array(shape)
    sph = np.concatenate([centres, rng.uniform(1.0, 7.5, (nsph, 1))], axis=1).astype(np.float32)
    fluid = (rng.random(shape) > 0.02).astype(np.uint8)      # some cells already solid
    sn = np.where(fluid == 0, 999, 0).astype(np.int64)       # and numbered
    return sph, fluid, sn


def _noalloc(shape, sph, fluid, sn, first_id):
    lib = _abi.load()
    df, ds = _Dev(fluid), (_Dev(sn) if sn is not None else None)
    try:
        st = lib.mplbm_place_spheres_noalloc(*shape, sph.ctypes.data, len(sph), first_id, df.ptr,
                                             ds.ptr if ds else None)
        assert st == 0, lib.mplbm_last_error(None).decode()
        return df.get(), (ds.get() if ds else None)
    finally:
        df.free()
        if ds:
            ds.free()


@pytest.mark.parametrize("numbers", [True, False])
def test_noalloc_matches_place_spheres_twice(numbers):
    for call, (shape, nsph, first_id) in enumerate([((48, 40, 36), 60, 1), ((20, 18, 22), 9, 500)]):
        print(f"case noalloc call {call} {shape} spheres {nsph} numbers {numbers}", flush=True)
        sph, fluid, sn = _case(shape, nsph, 7 + call)
        ref_f, ref_sn = fluid.copy(), sn.copy() if numbers else None
        mp.place_spheres(ref_f, sph, first_id, ref_sn)
        got_f, got_sn = _noalloc(shape, sph, fluid, sn if numbers else None, first_id)
        assert (ref_f == 0).sum() > (fluid == 0).sum()  # the spheres covered cells
        assert np.array_equal(got_f, ref_f)
        if numbers:
            assert np.array_equal(got_sn, ref_sn)
            assert (ref_sn >= first_id).any() and (ref_sn == 999).any()


def test_noalloc_refuses_calls_beyond_the_scratch():
    lib = _abi.load()
    sph = np.array([[4.0, 4.0, 4.0, 2.0]], np.float32)
    dummy = ctypes.c_void_p(16)  # never dereferenced: the call is refused first
    st = lib.mplbm_place_spheres_noalloc(300, 300, 300, sph.ctypes.data, 1, 1, dummy, dummy)
    assert st == _abi.ERR_ARG
    assert "scratch limit" in lib.mplbm_last_error(None).decode()
    many = np.tile(sph, (300000, 1))
    st = lib.mplbm_place_spheres_noalloc(8, 8, 8, many.ctypes.data, len(many), 1, dummy, None)
    assert st == _abi.ERR_ARG and "scratch limit" in lib.mplbm_last_error(None).decode()
    bad = np.array([[4.0, float("nan"), 4.0, 2.0]], np.float32)
    st = lib.mplbm_place_spheres_noalloc(8, 8, 8, bad.ctypes.data, 1, 1, dummy, None)
    assert st == _abi.ERR_ARG and "outside the voxeliser's range" in lib.mplbm_last_error(None).decode()
