"""Stochastic rounding of the 16-bit deviation storage (mplbm_config.rounding).

With round-to-nearest a steady population is rounded the same way every
step, so its error is a bias and the total mass drifts linearly.  The
stochastic form dithers by up to one unit in the last place before rounding
(unbiased), deterministically per (step, node, population): runs repeat bit
for bit, and the mass drift of a two-phase run shrinks.
"""
import numpy as np
import pytest

from tests.helpers import noisy_rho, random_spheres, sphere_mask

pytestmark = pytest.mark.gpu

import multiphase_lbm_amd as mp  # noqa: E402

SC = dict(tau=1.0, G=-5.0, rhos=0.7, rhol_spinodal=1.285931, rhog_spinodal=0.323507)


def _case(shape, seed):
    fluid = sphere_mask(shape, random_spheres(shape, 6, 2.0, 4.0, seed))
    return fluid, noisy_rho(shape, 0.7, 0.01, seed + 1).astype(np.float32)


def _run(shape, fluid, rho, storage, steps, **kw):
    with mp.Context(*shape, storage=storage, kernel="compact", **SC, **kw) as ctx:
        ctx.set_geometry(fluid)
        ctx.init_fields(rho)
        m0 = ctx.reduce()["mass"]
        ctx.step(0, steps, write_macro=True)
        st = ctx.reduce()
        f = ctx.read_populations()
    return f, st["mass"] / m0 - 1.0, st["nonfinite"]


@pytest.mark.parametrize("storage", ["fp16s", "bf16s"])
def test_stochastic_rounding_is_reproducible_and_drifts_less(storage):
    shape = (32, 32, 32)
    fluid, rho = _case(shape, 12)
    steps = 1500
    print(f"case rounding {storage}", flush=True)
    f_n, drift_n, bad_n = _run(shape, fluid, rho, storage, steps)
    f_s, drift_s, bad_s = _run(shape, fluid, rho, storage, steps, rounding="stochastic")
    f_s2, drift_s2, _ = _run(shape, fluid, rho, storage, steps, rounding="stochastic")
    assert bad_n == 0 and bad_s == 0
    assert np.array_equal(f_s[fluid], f_s2[fluid]) and drift_s == drift_s2
    assert not np.array_equal(f_s[fluid], f_n[fluid])
    print(f"  mass drift nearest {drift_n:.3e} stochastic {drift_s:.3e}", flush=True)
    assert abs(drift_s) < 0.5 * abs(drift_n)


def test_nearest_is_the_default():
    shape = (12, 10, 32)
    fluid, rho = _case(shape, 3)
    a, _, _ = _run(shape, fluid, rho, "fp16s", 20)
    b, _, _ = _run(shape, fluid, rho, "fp16s", 20, rounding="nearest")
    assert np.array_equal(a[fluid], b[fluid])


def test_stochastic_rounding_refusals():
    with pytest.raises(mp.MplbmError, match="stochastic rounding"):
        mp.Context(8, 8, 8, storage="fp32", kernel="compact", rounding="stochastic")
    with pytest.raises(mp.MplbmError, match="stochastic rounding"):
        mp.Context(8, 8, 8, storage="fp16s", kernel="fused", rounding="stochastic")
    from multiphase_lbm_amd.ensemble import Ensemble
    with pytest.raises(mp.MplbmError, match="round-to-nearest"):
        Ensemble([dict(lx=16, ly=16, lz=16, rounding=mp.ROUND_STOCHASTIC)], storage="fp16s")
