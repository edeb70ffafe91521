"""Exact checkpoint/restart (mplbm_save_checkpoint / mplbm_load_checkpoint).

The reference restarts only from a 5-digit density VTK
(src/multiphase_lbm.cu:112-132), which changes the trajectory.  A run saved
at step k and restored into a fresh context must continue bit for bit like
the uninterrupted run, for every storage form (16-bit deviations included)
and every kernel; files of another domain, storage or geometry are refused.
"""
import numpy as np
import pytest

from tests.helpers import noisy_rho, random_spheres, sphere_mask

pytestmark = pytest.mark.gpu

import multiphase_lbm_amd as mp  # noqa: E402

SC = dict(tau=1.0, G=-5.0, rhos=0.7, rhol_spinodal=1.285931, rhog_spinodal=0.323507)


def _case(shape, seed):
    fluid = sphere_mask(shape, random_spheres(shape, 4, 2.0, 3.5, seed))
    return fluid, noisy_rho(shape, 0.7, 0.01, seed + 1)


@pytest.mark.parametrize("storage,kernel,bb", [("fp32", "compact", 1), ("fp16s", "compact", 1),
                                               ("bf16s", "compact", 1), ("fp64", "compact", 1),
                                               ("fp32", "fused", 0), ("fp32", "fused", 1),
                                               ("fp16s", "twopass", 0)])
def test_restart_continues_bit_identically(tmp_path, storage, kernel, bb):
    shape = (14, 12, 64)
    fluid, rho = _case(shape, 3)
    rho = rho.astype(np.float64 if storage == "fp64" else np.float32)
    path = tmp_path / "run.ckpt"
    print(f"case checkpoint {storage} {kernel} bb{bb}", flush=True)
    kw = dict(storage=storage, kernel=kernel, bounce_back=bb, **SC)
    with mp.Context(*shape, **kw) as a:
        a.set_geometry(fluid)
        a.init_fields(rho)
        a.step(0, 5)
        a.save_checkpoint(path, 5)
        f5 = a.read_populations()
        # a context whose freed device memory holds NaN fields: a field that
        # the load leaves unwritten in the next context shows up as NaN
        with mp.Context(*shape, **kw) as junk:
            junk.set_geometry(fluid)
            junk.init_fields(np.full_like(rho, np.nan))
        # context a stays alive, so b cannot inherit its fields
        with mp.Context(*shape, **kw) as b:
            b.set_geometry(fluid)
            assert b.load_checkpoint(path) == 5
            r0, p0, u0 = b.read_fields()
            assert np.isfinite(r0).all() and np.isfinite(p0).all() and np.isfinite(u0).all()
            rhos = SC["rhos"] if storage == "fp64" else float(np.float32(SC["rhos"]))
            assert np.all(r0[~fluid] == rhos)
            assert np.allclose(r0[fluid], f5[fluid].sum(-1), rtol=1e-6)
            b.step(5, 6, write_macro=True)
            fb = b.read_populations()
            rb, _, _ = b.read_fields()
        a.step(5, 6, write_macro=True)
        fa = a.read_populations()
        ra, _, _ = a.read_fields()
    assert np.array_equal(fa[fluid], fb[fluid])
    assert np.array_equal(ra[fluid], rb[fluid])


def test_checkpoint_file_size_and_refusals(tmp_path):
    shape = (10, 9, 8)
    fluid, rho = _case(shape, 5)
    path = tmp_path / "a.ckpt"
    with mp.Context(*shape, storage="fp16s", kernel="compact", **SC) as a:
        a.set_geometry(fluid)
        a.init_fields(rho.astype(np.float32))
        a.save_checkpoint(path, 0)
    assert path.stat().st_size == 88 + int(fluid.sum()) * 19 * 2
    with mp.Context(*shape, storage="fp32", kernel="compact", **SC) as b:  # other storage
        b.set_geometry(fluid)
        with pytest.raises(mp.MplbmError, match="storage"):
            b.load_checkpoint(path)
    other = fluid.copy()
    other.flat[np.flatnonzero(other)[0]] = False
    with mp.Context(*shape, storage="fp16s", kernel="compact", **SC) as c:  # other geometry
        c.set_geometry(other)
        with pytest.raises(mp.MplbmError, match="geometry"):
            c.load_checkpoint(path)
    with mp.Context(9, 9, 8, storage="fp16s", kernel="compact", **SC) as d:  # other domain
        d.set_geometry(np.ones((9, 9, 8), bool))
        with pytest.raises(mp.MplbmError, match="domain"):
            d.load_checkpoint(path)
    bad = tmp_path / "bad.ckpt"
    bad.write_bytes(b"not a checkpoint")
    with mp.Context(*shape, storage="fp16s", kernel="compact", **SC) as e:
        e.set_geometry(fluid)
        with pytest.raises(mp.MplbmError, match="not a checkpoint"):
            e.load_checkpoint(bad)
        trunc = tmp_path / "t.ckpt"
        trunc.write_bytes(path.read_bytes()[:-40])
        with pytest.raises(mp.MplbmError, match="truncated"):
            e.load_checkpoint(trunc)
