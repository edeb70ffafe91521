"""The reference's host<->device calls through the compat layer
(include/iteration_host_compat.h), in the order and with the arguments of
MultiphaseLBM::run_gpu (src/multiphase_lbm.cu:584-606):

    transfer_parameter(&tau, &G, &carn_star, &T_Tc, &rhol_spinodal, &rhog_spinodal,
                       &inject_density, &rho_inject_period, &rho_increment, &rhos,
                       &inject_position, &statPeriod);
    multiphaselbm_initialization(&drho, &dpsi, &dpressure, &dis_fluid,
                                 &dcollision_f, &dstream_f, &dforce);
    for (; step < step_max; step++) multiphaselbm_iteration(step, statPeriod, reverse, inject_type);

64^3 with 40 non-overlapping spheres (the 64^3 workload), G = -5,
injection every 25 steps on y < ly/2.  On each
statistics step the caller's density and pressure arrays hold the oracle's
fields of that step bit for bit (FP32, the reference's precision).  The
caller's arrays are host memory or managed memory (the reference allocates
them with cudaMallocManaged, src/multiphase_lbm.cu:99-101)."""
import ctypes

import numpy as np
import pytest

from oracle import oracle
from tests.helpers import config1_case

pytestmark = pytest.mark.gpu

import multiphase_lbm_amd as mp  # noqa: E402,F401  (fails loudly if libmplbm.so is missing)
from multiphase_lbm_amd import _abi  # noqa: E402

F, B, SZ = ctypes.c_float, ctypes.c_bool, ctypes.c_size_t
PARAMS = dict(tau=1.0, G=-5.0, carn_star=False, T_Tc=0.7, rhol_spinodal=1.285931, rhog_spinodal=0.323507,
              rho_inject_period=25, rho_increment=0.002, rhos=0.7, inject_position=0.5)


def _lib():
    L = ctypes.CDLL(_abi.LIB_PATH)
    L.mplbm_compat_set_domain.argtypes = [ctypes.c_int64] * 3 + [ctypes.c_int32]
    L.transfer_parameter.argtypes = [ctypes.c_void_p] * 12
    L.transfer_parameter.restype = None
    L.multiphaselbm_initialization.argtypes = [ctypes.c_void_p] * 7
    L.multiphaselbm_initialization.restype = None
    L.multiphaselbm_iteration.argtypes = [ctypes.c_long, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint]
    L.multiphaselbm_iteration.restype = None
    L.mplbm_compat_set_abort.argtypes = [ctypes.c_int]
    L.mplbm_compat_status.restype = ctypes.c_int
    return L


class _Managed:
    """n bytes of hipMallocManaged memory with a numpy view."""

    def __init__(self, shape, dtype):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.ptr = ctypes.c_void_p()
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        assert self.hip.hipMallocManaged(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes), 1) == 0
        buf = (ctypes.c_char * nbytes).from_address(self.ptr.value)
        self.a = np.frombuffer(buf, dtype=dtype).reshape(shape)

    def free(self):
        self.hip.hipDeviceSynchronize()
        self.hip.hipFree(self.ptr)


@pytest.mark.parametrize("memory", ["host", "managed"])
def test_reference_call_sequence_matches_oracle(memory):
    shape, fluid, rho0 = config1_case()
    print(f"case compat {memory} 64^3", flush=True)
    st = oracle.State(oracle.Params(*shape, **PARAMS), fluid, rho0, dtype=np.float32)
    p_init = st.p.copy()

    L = _lib()
    L.mplbm_compat_set_abort(0)  # fail the test instead of exiting the process
    managed = []
    if memory == "host":
        drho, dpressure, dis_fluid = rho0.copy(), np.zeros(shape, np.float32), fluid.astype(np.bool_)
    else:
        managed = [_Managed(shape, np.float32), _Managed(shape, np.float32), _Managed(shape, np.bool_)]
        drho, dpressure, dis_fluid = (m.a for m in managed)
        drho[...] = rho0
        dpressure[...] = 0
        dis_fluid[...] = fluid
    try:
        assert L.mplbm_compat_set_domain(*shape, _abi.FP32) == 0
        args = [F(PARAMS["tau"]), F(PARAMS["G"]), B(PARAMS["carn_star"]), F(PARAMS["T_Tc"]),
                F(PARAMS["rhol_spinodal"]), F(PARAMS["rhog_spinodal"]), B(True), SZ(PARAMS["rho_inject_period"]),
                F(PARAMS["rho_increment"]), F(PARAMS["rhos"]), F(PARAMS["inject_position"]), SZ(50)]
        L.transfer_parameter(*[ctypes.byref(a) for a in args])
        assert L.mplbm_compat_status() == 0
        # addresses of the caller's array pointers (Realxyz** ...); psi, both
        # lattices and the force belong to the library
        ptrs = [ctypes.c_void_p(a.ctypes.data) for a in (drho, dpressure, dis_fluid)]
        L.multiphaselbm_initialization(ctypes.byref(ptrs[0]), None, ctypes.byref(ptrs[1]), ctypes.byref(ptrs[2]),
                                       None, None, None)
        assert L.mplbm_compat_status() == 0
        # multiphaselbm_init: rhos on solids, P(rho) on fluid nodes
        assert np.all(drho[~fluid] == np.float32(PARAMS["rhos"]))
        assert np.array_equal(drho[fluid], rho0[fluid])
        assert np.array_equal(dpressure[fluid], p_init[fluid])
        done = 0
        for step in range(0, 101):
            L.multiphaselbm_iteration(step, 50, 1, 0)
            assert L.mplbm_compat_status() == 0, step
            if step % 50 == 0:
                st.run(step + 1 - done)
                done = step + 1
                print(f"  statistics step {step}", flush=True)
                assert np.array_equal(drho[fluid], st.rho[fluid]), step
                assert np.array_equal(dpressure[fluid], st.p[fluid]), step
                assert np.all(dpressure[~fluid] == 0)
    finally:
        L.mplbm_compat_release()
        L.mplbm_compat_set_abort(1)
        for m in managed:
            m.free()
