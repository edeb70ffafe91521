"""Committed golden vectors (tests/golden/make_golden.py): the oracle still
reproduces them bit for bit, and the GPU library matches them."""
import os

import numpy as np
import pytest

from oracle import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.startswith("step_") and f.endswith(".npz"))


def _load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


def _params(prm):
    return oracle.Params(int(prm[0]), int(prm[1]), int(prm[2]), tau=prm[3], G=prm[4], carn_star=bool(prm[5]),
                         T_Tc=prm[6], rhol_spinodal=prm[7], rhog_spinodal=prm[8], rho_inject_period=int(prm[9]),
                         rho_increment=prm[10], rhos=prm[11], inject_position=prm[12])


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    g = _load(name)
    P = _params(g["prm"])
    st = oracle.State(P, g["fluid"], g["rho0"], dtype=g["rho0"].dtype).run(int(g["steps"]),
                                                                          inject_type=int(g["inject_type"]))
    fl = g["fluid"]
    assert np.array_equal(st.fA[fl], g["f"]) and np.array_equal(st.rho[fl], g["rho"])
    assert np.array_equal(st.u[fl], g["u"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("kernel", ["fused", "twopass", "compact"])
def test_gpu_matches_golden(name, kernel):
    import multiphase_lbm_amd as mp
    g = _load(name)
    prm = g["prm"]
    shape = (int(prm[0]), int(prm[1]), int(prm[2]))
    storage = "fp64" if g["rho0"].dtype == np.float64 else "fp32"
    print(f"golden {name} {kernel}", flush=True)
    with mp.Context(*shape, storage=storage, kernel=kernel, tau=prm[3], G=prm[4], carn_star=bool(prm[5]),
                    T_Tc=prm[6], rhol_spinodal=prm[7], rhog_spinodal=prm[8], rho_inject_period=int(prm[9]),
                    rho_increment=prm[10], rhos=prm[11], inject_position=prm[12]) as ctx:
        ctx.set_geometry(g["fluid"])
        ctx.init_fields(g["rho0"])
        ctx.step(0, int(g["steps"]), inject_type=int(g["inject_type"]), write_macro=True)
        f = ctx.read_populations()
        rho, p, u = ctx.read_fields()
    fl = g["fluid"]
    assert np.array_equal(f[fl], g["f"].astype(np.float64))
    assert np.array_equal(rho[fl], g["rho"].astype(np.float64))
    assert np.array_equal(p[fl], g["p"].astype(np.float64))
    assert np.array_equal(u[fl], g["u"].astype(np.float64))
