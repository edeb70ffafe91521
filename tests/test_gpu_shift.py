"""Node-shifted 16-bit storage (mplbm_config.shift = MPLBM_SHIFT_NODE).

Each node stores its populations as deviations from its own density, and
that density (rounded to the 16-bit type) beside them; a pull decodes every
value with the shift of the node that wrote it.  The stored deviations stay
small in both phases of a two-phase run, so the error against FP64 must be
well below the global-shift form's on the same case, runs must restart
bit for bit from a checkpoint, and the forms the shift does not cover must
be refused.
"""
import numpy as np
import pytest

from tests.helpers import noisy_rho, random_spheres, sphere_mask

pytestmark = pytest.mark.gpu

import multiphase_lbm_amd as mp  # noqa: E402

SC = dict(tau=1.0, G=-5.0, rhos=0.7, rhol_spinodal=1.285931, rhog_spinodal=0.323507)


def _case(shape, seed):
    fluid = sphere_mask(shape, random_spheres(shape, 8, 2.0, 4.5, seed))
    return fluid, noisy_rho(shape, 0.7, 0.01, seed + 1)


def _run(shape, fluid, rho, storage, steps, **kw):
    with mp.Context(*shape, storage=storage, kernel="compact", **kw) as ctx:
        ctx.set_geometry(fluid)
        ctx.init_fields(rho.astype(np.float64 if storage == "fp64" else np.float32))
        m0 = ctx.reduce()["mass"]
        ctx.step(0, steps, write_macro=True)
        r, _, _ = ctx.read_fields()
        st = ctx.reduce()
    return r, st["mass"] / m0 - 1.0, st["nonfinite"]


def _err(r, ref, fluid):
    return float(np.linalg.norm(r[fluid] - ref[fluid]) / np.linalg.norm(ref[fluid]))


@pytest.mark.parametrize("storage", ["fp16s", "bf16s"])
def test_node_shift_tracks_fp64_closer_than_global_shift(storage):
    shape = (40, 36, 64)
    fluid, rho = _case(shape, 21)
    steps = 400
    print(f"case node shift {storage} shape={shape} steps={steps}", flush=True)
    ref, _, _ = _run(shape, fluid, rho, "fp64", steps, **SC)
    rg, drift_g, bad_g = _run(shape, fluid, rho, storage, steps, **SC)
    rn, drift_n, bad_n = _run(shape, fluid, rho, storage, steps, shift="node", **SC)
    eg, en = _err(rg, ref, fluid), _err(rn, ref, fluid)
    print(f"  density rel-L2 global {eg:.3e} node {en:.3e}; mass drift global {drift_g:.3e} node {drift_n:.3e}",
          flush=True)
    assert bad_g == 0 and bad_n == 0
    assert (ref[fluid] > 1.5).any() and (rn[fluid] > 1.5).any()  # both separated into two phases
    assert en < 0.5 * eg
    assert abs(drift_n) < abs(drift_g)


@pytest.mark.parametrize("axis", [0, 1])
def test_node_shift_with_pressure_faces(axis):
    # faces on x (face kernel on the end segments only) or y (face kernel
    # everywhere): anti-bounce-back from the node's own populations, against FP32
    shape = (24, 18, 32)
    fluid, rho = _case(shape, 5)
    ft, fr = [0] * 6, [0.0] * 6
    ft[2 * axis] = ft[2 * axis + 1] = 1
    fr[2 * axis], fr[2 * axis + 1] = 1.04, 0.96
    params = dict(SC, G=-3.0, face_type=tuple(ft), face_rho=tuple(fr))
    print(f"case node shift faces axis {axis}", flush=True)
    r32, _, _ = _run(shape, fluid, rho, "fp32", 300, **params)
    rn, _, bad = _run(shape, fluid, rho, "fp16s", 300, shift="node", **params)
    rg, _, _ = _run(shape, fluid, rho, "fp16s", 300, **params)
    assert bad == 0
    assert _err(rn, r32, fluid) < 0.5 * _err(rg, r32, fluid)


@pytest.mark.parametrize("storage", ["fp16s", "bf16s"])
def test_node_shift_restart_is_bit_identical(tmp_path, storage):
    shape = (14, 12, 64)
    fluid, rho = _case(shape, 3)
    rho = rho.astype(np.float32)
    path = tmp_path / "ns.ckpt"
    kw = dict(storage=storage, kernel="compact", shift="node", **SC)
    print(f"case node shift checkpoint {storage}", flush=True)
    with mp.Context(*shape, **kw) as a:
        a.set_geometry(fluid)
        a.init_fields(rho)
        a.step(0, 7)
        a.save_checkpoint(path, 7)
        a.step(7, 6, write_macro=True)
        fa = a.read_populations()
        ra, _, _ = a.read_fields()
    assert path.stat().st_size == 88 + int(fluid.sum()) * 20 * 2
    with mp.Context(*shape, **kw) as b:
        b.set_geometry(fluid)
        assert b.load_checkpoint(path) == 7
        b.step(7, 6, write_macro=True)
        fb = b.read_populations()
        rb, _, _ = b.read_fields()
    assert np.array_equal(fa[fluid], fb[fluid])
    assert np.array_equal(ra[fluid], rb[fluid])
    with mp.Context(*shape, storage=storage, kernel="compact", **SC) as c:  # global shift refuses the file
        c.set_geometry(fluid)
        with pytest.raises(mp.MplbmError, match="shift"):
            c.load_checkpoint(path)


@pytest.mark.parametrize("tiles", [(8, 32, 5, 0), (32, 32, 32, 1024), (3, 5, 2, 256)])
def test_node_shift_results_do_not_depend_on_the_tiling(tiles):
    # workgroups of 256, 512 and 1024 threads, partial tiles and short
    # segments give the default tiling's populations bit for bit
    shape = (12, 40, 64)
    fluid, rho = _case(shape, 13)
    ty, tz, seg, nt = tiles
    print(f"case node shift tiling {tiles}", flush=True)
    out = []
    for kw in ({}, dict(tile_y=ty, tile_z=tz, seg_x=seg, block_threads=nt)):
        with mp.Context(*shape, storage="fp16s", kernel="compact", shift="node", **SC, **kw) as ctx:
            ctx.set_geometry(fluid)
            ctx.init_fields(rho.astype(np.float32))
            ctx.step(0, 25)
            out.append(ctx.read_populations())
    assert np.array_equal(out[0][fluid], out[1][fluid])


def test_node_shift_population_round_trip():
    # written populations are stored against their node's density: reading
    # them back returns them to 16-bit precision of the deviation
    shape = (10, 12, 32)
    fluid, rho = _case(shape, 9)
    with mp.Context(*shape, storage="fp16s", kernel="compact", shift="node", **SC) as ctx:
        ctx.set_geometry(fluid)
        ctx.init_fields(rho.astype(np.float32))
        ctx.step(0, 50)
        f = ctx.read_populations()
        ctx.write_populations(f)
        g = ctx.read_populations()
        st = ctx.reduce()
    # D3Q19 weights in the reference's direction order (include/lbm_model.cuh:35-39)
    w = np.array([1 / 3 if i == 0 else (1 / 18 if i in (1, 2, 3, 10, 11, 12) else 1 / 36) for i in range(19)])
    rho_n = f[fluid].sum(axis=1, keepdims=True)
    dev = np.abs(f[fluid] / w - rho_n)
    # the deviation from the node's 16-bit density is rounded to 11 bits
    tol = 2.0 ** -10 * (dev + 2.0 ** -10 * rho_n) + 1e-6
    assert np.all(np.abs(g[fluid] - f[fluid]) / w <= tol)
    assert st["nonfinite"] == 0


def test_node_shift_with_stochastic_rounding():
    # dithered deviations from the node's density: reproducible, different
    # from round-to-nearest, and the mass drift stays at the node-shift level
    # (round-to-nearest's is already ~1e-5 here, 30x below the global shift's)
    shape = (40, 36, 64)
    fluid, rho = _case(shape, 21)
    print("case node shift stochastic", flush=True)
    rn, drift_n, _ = _run(shape, fluid, rho, "fp16s", 400, shift="node", **SC)
    rs, drift_s, bad = _run(shape, fluid, rho, "fp16s", 400, shift="node", rounding="stochastic", **SC)
    rs2, drift_s2, _ = _run(shape, fluid, rho, "fp16s", 400, shift="node", rounding="stochastic", **SC)
    print(f"  mass drift nearest {drift_n:.3e} stochastic {drift_s:.3e}", flush=True)
    assert bad == 0
    assert np.array_equal(rs[fluid], rs2[fluid]) and drift_s == drift_s2
    assert not np.array_equal(rs[fluid], rn[fluid])
    assert abs(drift_s) < 5e-5 and abs(drift_n) < 5e-5


def test_node_shift_refusals():
    with pytest.raises(mp.MplbmError, match="node shifts"):
        mp.Context(8, 8, 8, storage="fp32", kernel="compact", shift="node")
    with pytest.raises(mp.MplbmError, match="node shifts"):
        mp.Context(8, 8, 8, storage="fp16s", kernel="fused", shift="node")
    from multiphase_lbm_amd.ensemble import Ensemble
    with pytest.raises(mp.MplbmError, match="global shift"):
        Ensemble([dict(lx=16, ly=16, lz=16, shift=mp.SHIFT_NODE)], storage="fp16s")
