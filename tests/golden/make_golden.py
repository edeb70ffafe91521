"""Writes tests/golden/step_*.npz: oracle outputs on small seeded lattices.

Each fixture holds the inputs (mask, initial density, parameters) and the
expected outputs after `steps` steps (newest populations of fluid nodes,
density, pressure, shifted velocity) in the oracle's precision.  The
reference itself ships no golden vectors; these pin regressions of the
oracle (tests/test_golden.py) and give the GPU tests fixed targets.

Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402
from tests.helpers import noisy_rho, random_spheres, sphere_mask  # noqa: E402

CASES = {
    "step_sc_16_f32": dict(shape=(16, 16, 16), nsph=3, seed=1, steps=10, dtype="float32",
                           params=dict(tau=1.0, G=-5.0, rhos=0.7, rhol_spinodal=1.285931)),
    "step_sc_16_f64": dict(shape=(16, 16, 16), nsph=3, seed=1, steps=10, dtype="float64",
                           params=dict(tau=1.0, G=-5.0, rhos=0.7, rhol_spinodal=1.285931)),
    "step_inj_12x10x14_f32": dict(shape=(12, 10, 14), nsph=2, seed=4, steps=7, dtype="float32",
                                  params=dict(tau=0.9, G=-4.5, rhos=1.0, rhol_spinodal=1.0, rho_inject_period=3,
                                              rho_increment=0.004, inject_position=0.6), inject_type=1),
}


def build(name, c):
    shape = c["shape"]
    dt = np.dtype(c["dtype"])
    fluid = sphere_mask(shape, random_spheres(shape, c["nsph"], 2.0, 3.5, c["seed"]))
    rho = noisy_rho(shape, 0.7, 0.05, c["seed"] + 1).astype(dt)
    P = oracle.Params(*shape, **c["params"])
    st = oracle.State(P, fluid, rho, dtype=dt).run(c["steps"], inject_type=c.get("inject_type", 0))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), fluid=fluid, rho0=rho, prm=P.vector(),
                        steps=c["steps"], inject_type=c.get("inject_type", 0),
                        f=st.fA[fluid], rho=st.rho[fluid], p=st.p[fluid], u=st.u[fluid])


if __name__ == "__main__":
    for n, c in CASES.items():
        build(n, c)
        print("wrote", n)
