"""GPU parity: libmplbm.so (through the C ABI) against the oracle.

Tolerances (from the specification of this path): after 100 steps per-node
density and velocity agree within 1e-5 relative for the FP32 form (the
reference's precision) and 1e-12 relative for FP64.  The FP32 and FP64 forms
evaluate the same IEEE operation sequence as the oracle (own exp, no fused
multiply-adds), so these tests also assert bit-identical populations where
that is expected; the tolerance assertions stay as the stated bar.
"""
import numpy as np
import pytest

from oracle import oracle
from tests.helpers import DENSITY_L2_BAR, config1_case, noisy_rho, random_spheres, sphere_mask

pytestmark = pytest.mark.gpu

import multiphase_lbm_amd as mp  # noqa: E402  (fails loudly if libmplbm.so is missing)


def _case(shape, nsph, rmin, rmax, seed, rho0=0.7, amp=0.01):
    sph = random_spheres(shape, nsph, rmin, rmax, seed)
    fluid = sphere_mask(shape, sph)
    rho = noisy_rho(shape, rho0, amp, seed + 1)
    return fluid, rho


def _run_pair(shape, fluid, rho, steps, storage="fp32", kernel="fused", step0=0, reverse=1, inject_type=0,
              **params):
    dtype = np.float64 if storage == "fp64" else np.float32
    P = oracle.Params(*shape, **{k: v for k, v in params.items() if k not in LIB_ONLY})
    st = oracle.State(P, fluid, rho.astype(dtype), dtype=dtype)
    st.step = step0
    st.run(steps, reverse=reverse, inject_type=inject_type)
    with mp.Context(*shape, storage=storage, kernel=kernel, **_lib_params(params)) as ctx:
        ctx.set_geometry(fluid)
        ctx.init_fields(rho.astype(dtype))
        ctx.step(step0, steps, reverse=reverse, inject_type=inject_type, write_macro=True)
        g_rho, g_p, g_u = ctx.read_fields(np.float64)
        g_f = ctx.read_populations()
        stats = ctx.reduce()
    return st, (g_rho, g_p, g_u, g_f, stats)


LIB_ONLY = ("tile_y", "tile_z", "seg_x", "rho_ref", "dev_scale", "bounce_back", "plane_pad", "block_threads")


def _lib_params(params):
    return dict(params)


def _relerr(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.maximum(np.abs(b), 1e-30)
    return float(np.max(np.abs(a - b) / scale)) if a.size else 0.0


def _abserr(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)))) if np.size(a) else 0.0


SC = dict(tau=1.0, G=-5.0, rhos=0.7, rhol_spinodal=1.285931, rhog_spinodal=0.323507)


@pytest.mark.parametrize("kernel", ["fused", "twopass", "compact"])
def test_fp32_small_bit_exact(kernel):
    shape = (12, 10, 14)
    fluid, rho = _case(shape, 3, 1.5, 3.0, seed=3)
    print(f"case fp32 {kernel} shape={shape}", flush=True)
    st, (g_rho, g_p, g_u, g_f, _) = _run_pair(shape, fluid, rho, 20, "fp32", kernel, **SC)
    assert _relerr(g_rho[fluid], st.rho[fluid]) <= 1e-5
    assert _abserr(g_u[fluid], st.u[fluid]) <= 1e-5 * max(1e-3, float(np.abs(st.u[fluid]).max()))
    assert np.array_equal(g_f[fluid], st.fA[fluid].astype(np.float64)), "populations not bit-identical"
    assert np.array_equal(g_rho[fluid], st.rho[fluid].astype(np.float64))
    assert np.array_equal(g_p[fluid], st.p[fluid].astype(np.float64))


@pytest.mark.parametrize("kernel,bb", [("fused", 0), ("fused", 1), ("twopass", 0), ("compact", 1)])
def test_fp32_config1_100_steps(kernel, bb):
    """64^3, 40 non-overlapping spheres r in [3,6] (random sequential
    addition), G=-5, tau=1, rho0=0.7 +/- 0.01, 100 steps."""
    shape, fluid, rho = config1_case()
    print(f"case fp32 config1 {kernel} bb {bb}", flush=True)
    st, (g_rho, g_p, g_u, g_f, stats) = _run_pair(shape, fluid, rho, 100, "fp32", kernel, bounce_back=bb, **SC)
    assert _relerr(g_rho[fluid], st.rho[fluid]) <= 1e-5
    umax = float(np.abs(st.u[fluid]).max())
    assert _abserr(g_u[fluid], st.u[fluid]) <= 1e-5 * umax
    assert np.array_equal(g_f[fluid], st.fA[fluid].astype(np.float64)), "populations not bit-identical"
    assert stats["nonfinite"] == 0


@pytest.mark.parametrize("kernel", ["fused", "compact"])
def test_fp64_config1_100_steps(kernel):
    shape, fluid, rho = config1_case()
    print(f"case fp64 config1 {kernel}", flush=True)
    st, (g_rho, g_p, g_u, g_f, stats) = _run_pair(shape, fluid, rho, 100, "fp64", kernel, **SC)
    assert _relerr(g_rho[fluid], st.rho[fluid]) <= 1e-12
    umax = float(np.abs(st.u[fluid]).max())
    assert _abserr(g_u[fluid], st.u[fluid]) <= 1e-12 * umax
    assert np.array_equal(g_f[fluid], st.fA[fluid]), "populations not bit-identical"
    # mass conservation (no injection): FP64 sum of populations
    m0 = float(np.sum(rho.astype(np.float64)[fluid]))
    assert abs(stats["mass"] - m0) <= 1e-12 * m0


@pytest.mark.parametrize("bb", [0, 1])
@pytest.mark.parametrize("tiles", [(6, 8, 3), (5, 16, 2), (32, 5, 7), (0, 0, 1)])
def test_fused_tilings_agree(tiles, bb):
    """Ring tiles in both axes, partial tiles and short segments, with
    write-side or pull-side bounce-back, give the oracle's result."""
    ty, tz, seg = tiles
    shape = (9, 23, 19)
    fluid, rho = _case(shape, 4, 1.5, 3.5, seed=7)
    print(f"case tiling {tiles} bb {bb}", flush=True)
    st, (g_rho, g_p, g_u, g_f, _) = _run_pair(shape, fluid, rho, 12, "fp32", "fused", tile_y=ty, tile_z=tz,
                                              seg_x=seg, bounce_back=bb, **SC)
    assert np.array_equal(g_f[fluid], st.fA[fluid].astype(np.float64))


@pytest.mark.parametrize("tiles", [(1, 1, 1, 0), (3, 5, 2, 0), (8, 32, 3, 0), (0, 0, 0, 0), (2, 19, 9, 512),
                                   (23, 7, 4, 0), (16, 32, 5, 0), (23, 19, 9, 1024), (16, 16, 4, 256)])
def test_compact_tilings_agree(tiles):
    """Compact kernel: single-node tiles, partial tiles in y and z, tiles
    wider than the domain, short and whole-axis segments, workgroups of 256,
    512 and 1024 threads; a 16 x 16 tile with 256 threads, whose psi region
    (324 positions, mostly fluid here) holds more pulls than threads, so the
    extra ring entries take the second pass over the list."""
    ty, tz, seg, nt = tiles
    shape = (9, 23, 19)
    fluid, rho = _case(shape, 4, 1.5, 3.5, seed=7)
    print(f"case compact tiling {tiles}", flush=True)
    st, (g_rho, g_p, g_u, g_f, _) = _run_pair(shape, fluid, rho, 12, "fp32", "compact", tile_y=ty, tile_z=tz,
                                              seg_x=seg, block_threads=nt, **SC)
    assert np.array_equal(g_f[fluid], st.fA[fluid].astype(np.float64))
    assert np.array_equal(g_rho[fluid], st.rho[fluid].astype(np.float64))


@pytest.mark.parametrize("shape,tiles", [((7, 10, 64), (4, 32, 3, 0)), ((6, 9, 64), (5, 64, 4, 0)),
                                         ((5, 11, 96), (3, 16, 2, 0)), ((5, 7, 96), (4, 48, 5, 256)),
                                         ((4, 13, 64), (2, 20, 8, 0)), ((1, 5, 64), (8, 32, 1, 0))])
def test_compact_tiles_across_z_blocks(shape, tiles):
    """lz a multiple of 32 (several 32-wide rank blocks per row): tiles on
    block edges (32, 64 wide), inside blocks (16, 20) and across them (48),
    partial tiles, one plane."""
    ty, tz, seg, nt = tiles
    fluid, rho = _case(shape, 5, 1.5, 3.5, seed=sum(shape))
    print(f"case compact z blocks {shape} {tiles}", flush=True)
    st, (g_rho, g_p, g_u, g_f, _) = _run_pair(shape, fluid, rho, 10, "fp32", "compact", tile_y=ty, tile_z=tz,
                                              seg_x=seg, block_threads=nt, **SC)
    assert np.array_equal(g_f[fluid], st.fA[fluid].astype(np.float64))
    assert np.array_equal(g_rho[fluid], st.rho[fluid].astype(np.float64))


def test_compact_z_faces_and_restart():
    """Pressure faces on z (the face columns are rank-block edges) and a
    population write/read round trip in the middle of the run."""
    shape = (6, 10, 64)
    fluid, rho = _case(shape, 4, 1.5, 3.0, seed=31)
    params = dict(SC, G=-3.0, face_type=(0, 0, 0, 0, 1, 1), face_rho=(0, 0, 0, 0, 1.04, 0.96))
    print(f"case compact z faces {shape}", flush=True)
    st, (g_rho, g_p, g_u, g_f, _) = _run_pair(shape, fluid, rho, 8, "fp32", "compact", **params)
    assert np.isfinite(g_f[fluid]).all()
    assert np.array_equal(g_f[fluid], st.fA[fluid].astype(np.float64))
    P = oracle.Params(*shape, **params)
    ref = oracle.State(P, fluid, rho.astype(np.float32), dtype=np.float32)
    ref.run(4)
    with mp.Context(*shape, storage="fp32", kernel="compact", **params) as ctx:
        ctx.set_geometry(fluid)
        ctx.init_fields(rho.astype(np.float32))
        ctx.step(0, 2)
        f = ctx.read_populations()
        ctx.write_populations(f)
        ctx.step(2, 2, write_macro=True)
        g = ctx.read_populations()
    assert np.array_equal(g[fluid], ref.fA[fluid].astype(np.float64))


@pytest.mark.parametrize("shape", [(5, 2, 3), (2, 3, 40), (1, 6, 5), (3, 1, 1)])
def test_compact_tiny_domains(shape):
    """Axes shorter than the tile and its rings (rows and columns of the LDS
    regions alias the same node), a single plane (both ghost planes are the
    plane itself)."""
    rng = np.random.default_rng(sum(shape))
    fluid = rng.random(shape) < 0.7
    fluid.flat[0] = True
    rho = noisy_rho(shape, 0.7, 0.01, 5)
    print(f"case compact tiny {shape}", flush=True)
    st, (g_rho, g_p, g_u, g_f, _) = _run_pair(shape, fluid, rho, 6, "fp32", "compact", **SC)
    assert np.array_equal(g_f[fluid], st.fA[fluid].astype(np.float64))


@pytest.mark.parametrize("kernel", ["fused", "compact"])
@pytest.mark.parametrize("inject_type", [0, 1, 2])
def test_injection(inject_type, kernel):
    shape = (16, 20, 12)
    fluid, rho = _case(shape, 3, 2.0, 3.0, seed=11, rho0=1.0, amp=0.3)
    print(f"case injection type {inject_type} {kernel}", flush=True)
    params = dict(SC, rho_inject_period=5, rho_increment=0.005, inject_position=0.5, rhol_spinodal=1.0)
    st, (g_rho, g_p, g_u, g_f, _) = _run_pair(shape, fluid, rho, 11, "fp32", kernel, step0=0, reverse=-1,
                                              inject_type=inject_type, **params)
    assert np.array_equal(g_f[fluid], st.fA[fluid].astype(np.float64))
    assert np.array_equal(g_rho[fluid], st.rho[fluid].astype(np.float64))


@pytest.mark.parametrize("kernel", ["fused", "twopass", "compact"])
def test_carnahan_starling(kernel):
    shape = (14, 12, 16)
    fluid, rho = _case(shape, 3, 2.0, 3.0, seed=5, rho0=0.2, amp=0.05)
    print(f"case CS {kernel}", flush=True)
    params = dict(tau=1.0, G=-1.0, carn_star=True, T_Tc=0.7, rhos=0.35, rhol_spinodal=0.2725,
                  rhog_spinodal=0.0484)
    st, (g_rho, g_p, g_u, g_f, _) = _run_pair(shape, fluid, rho, 30, "fp32", kernel, **params)
    assert _relerr(g_rho[fluid], st.rho[fluid]) <= 1e-5
    assert np.array_equal(g_f[fluid], st.fA[fluid].astype(np.float64))


@pytest.mark.parametrize("kernel,seg", [("fused", 0), ("twopass", 0), ("compact", 0), ("compact", 16),
                                        ("compact", 7)])
def test_pressure_faces_and_body_force(kernel, seg):
    """x faces only: the compact kernel runs its face variant on end
    segments [0, 7) and [14, 16) and the plain variant on the planes between
    (seg 0: the default length; 16: one middle segment; 7: one of 7)."""
    shape = (16, 12, 10)
    fluid, rho = _case(shape, 2, 2.0, 3.0, seed=9, rho0=1.0, amp=0.05)
    print(f"case pressure faces {kernel} seg {seg}", flush=True)
    params = dict(SC, G=-3.0, face_type=(1, 1, 0, 0, 0, 0), face_rho=(1.05, 0.95, 0, 0, 0, 0),
                  body_force=(1e-5, 0, 2e-6))
    if kernel == "compact":
        params["seg_x"] = seg
    st, (g_rho, g_p, g_u, g_f, _) = _run_pair(shape, fluid, rho, 15, "fp32", kernel, **params)
    assert np.isfinite(st.fA[fluid]).all()
    assert np.array_equal(g_f[fluid], st.fA[fluid].astype(np.float64))
    assert np.array_equal(g_u[fluid], st.u[fluid].astype(np.float64))


@pytest.mark.parametrize("lx,seg", [(15, 7), (29, 7), (16, 7), (17, 8), (7, 3), (841, 48)])
def test_pressure_faces_short_last_segment(lx, seg):
    """x faces on lengths that exercise the launch plans: end segments in the
    face variant of 6 and 3 planes (lx 15), 8 and 5 (29), 7 and 2 (16), 7 and
    3 (17), 9 and 4 (841, where no end of 2 .. 8 planes leaves a last end of
    two or more) with the planes between in middle segments of seg planes,
    a short last one included; lx 7 has no such plan and runs every segment
    in the face variant."""
    shape = (lx, 12, 10)
    fluid, rho = _case(shape, 2, 2.0, 3.0, seed=lx, rho0=1.0, amp=0.05)
    print(f"case pressure faces lx {lx} seg {seg}", flush=True)
    params = dict(SC, G=-3.0, face_type=(1, 1, 0, 0, 0, 0), face_rho=(1.05, 0.95, 0, 0, 0, 0), seg_x=seg)
    st, (g_rho, g_p, g_u, g_f, _) = _run_pair(shape, fluid, rho, 12, "fp32", "compact", **params)
    assert np.isfinite(st.fA[fluid]).all()
    assert np.array_equal(g_f[fluid], st.fA[fluid].astype(np.float64))
    assert np.array_equal(g_rho[fluid], st.rho[fluid].astype(np.float64))


@pytest.mark.parametrize("kernel", ["fused", "compact"])
def test_checkpoint_roundtrip_continues_identically(kernel):
    shape = (12, 12, 12)
    fluid, rho = _case(shape, 2, 2.0, 3.0, seed=13)
    print(f"case checkpoint {kernel}", flush=True)
    with mp.Context(*shape, storage="fp32", kernel=kernel, **SC) as a, \
            mp.Context(*shape, storage="fp32", kernel=kernel, **SC) as b:
        for c in (a, b):
            c.set_geometry(fluid)
            c.init_fields(rho)
        a.step(0, 7)
        b.step(0, 3)
        f = b.read_populations()
        b.write_populations(f)
        b.step(3, 4)
        assert np.array_equal(a.read_populations()[fluid], b.read_populations()[fluid])


@pytest.mark.parametrize("storage", [k for k in DENSITY_L2_BAR if k != "fp32"])
def test_deviation_storage_tracks_fp64(storage):
    shape = (32, 32, 32)
    fluid, rho = _case(shape, 8, 2.0, 4.0, seed=21)
    print(f"case {storage}", flush=True)
    st, (g_rho, g_p, g_u, g_f, stats) = _run_pair(shape, fluid, rho, 100, "fp64", "fused", **SC)
    with mp.Context(*shape, storage=storage, **SC) as ctx:
        ctx.set_geometry(fluid)
        ctx.init_fields(rho)
        ctx.step(0, 100, write_macro=True)
        h_rho, _, _ = ctx.read_fields()
    err = np.linalg.norm(h_rho[fluid] - st.rho[fluid]) / np.linalg.norm(st.rho[fluid])
    print(f"{storage} density rel-L2 after 100 steps: {err:.3e}", flush=True)
    assert err <= DENSITY_L2_BAR[storage]


def test_fp32_fast_math_error_vs_oracle():
    """FAST arithmetic (FP32 with fused multiply-adds, hardware exp and
    reciprocal) against the reference-arithmetic oracle on the config-1 case
    after 100 steps.  The spinodal start (rho0 = 0.7 inside the G = -5
    spinodal) amplifies rounding differences, so FAST does not meet the 1e-5
    bar of the FP32 form (measured ~2.4e-4); the reference-arithmetic FP32
    form is the one that does (bit-identical).  Asserted: FAST stays within
    1e-3, the bar of the 16-bit storage forms."""
    shape, fluid, rho = config1_case()
    print("case fp32 fast config1", flush=True)
    st, (g_rho, g_p, g_u, g_f, stats) = _run_pair(shape, fluid, rho, 100, "fp32fast", "fused", **SC)
    rel_rho = _relerr(g_rho[fluid], st.rho[fluid])
    umax = float(np.abs(st.u[fluid]).max())
    rel_u = _abserr(g_u[fluid], st.u[fluid]) / umax
    print(f"fp32 fast vs oracle after 100 steps: max rel density {rel_rho:.3e}, max |du|/max|u| {rel_u:.3e}")
    assert rel_rho <= 1e-3


@pytest.mark.parametrize("kernel", ["compact", "fused"])
def test_drainage_faces_two_phase(kernel):
    """The drainage set-up at small size: pressure inlet 2.0 / outlet 0.15 on
    x, G = -5, liquid blobs in vapour (bench.py --drainage)."""
    from bench import DRAIN, drainage_fields
    shape = (40, 24, 20)
    fluid, rho = drainage_fields(shape, 5, 0.4)
    fluid = fluid.astype(bool)
    print(f"case drainage {kernel}", flush=True)
    st, (g_rho, g_p, g_u, g_f, stats) = _run_pair(shape, fluid, rho.astype(np.float32), 30, "fp32", kernel, **DRAIN)
    assert np.isfinite(st.fA[fluid]).all()
    assert np.array_equal(g_f[fluid], st.fA[fluid].astype(np.float64))
    assert np.array_equal(g_rho[fluid], st.rho[fluid].astype(np.float64))


@pytest.mark.parametrize("kernel", ["fused", "compact"])
def test_pressure_faces_on_y(kernel):
    """Faces on y (every segment runs the face variant)."""
    shape = (10, 14, 12)
    fluid, rho = _case(shape, 2, 2.0, 3.0, seed=19, rho0=1.0, amp=0.05)
    print(f"case pressure faces y {kernel}", flush=True)
    params = dict(SC, G=-3.0, face_type=(0, 0, 1, 1, 0, 0), face_rho=(0, 0, 1.04, 0.96, 0, 0))
    st, (g_rho, g_p, g_u, g_f, _) = _run_pair(shape, fluid, rho, 12, "fp32", kernel, **params)
    assert np.isfinite(st.fA[fluid]).all()
    assert np.array_equal(g_f[fluid], st.fA[fluid].astype(np.float64))


@pytest.mark.parametrize("kernel", ["compact", "fused", "twopass"])
@pytest.mark.parametrize("geometry", ["all_solid", "one_fluid_node", "all_fluid"])
def test_degenerate_geometries(geometry, kernel):
    """Edge cases of the geometry: no fluid node at all (work lists and
    rank ranges are empty, every step is a no-op), a single fluid node
    enclosed by solids (every direction bounces back), and a fully fluid
    periodic box (no bounce-back, every rank range full: 17 x 9 x 40 spans
    two z blocks and ragged tiles).  Bit-identical to the oracle where there
    are fluid nodes; mass exactly zero without them."""
    shape = (7, 9, 40) if geometry != "all_fluid" else (17, 9, 40)
    fluid = np.zeros(shape, bool) if geometry != "all_fluid" else np.ones(shape, bool)
    if geometry == "one_fluid_node":
        fluid[3, 4, 33] = True
    rho = noisy_rho(shape, 0.7, 0.01, 11)
    print(f"case degenerate {geometry} {kernel} shape={shape}", flush=True)
    st, (g_rho, g_p, g_u, g_f, stats) = _run_pair(shape, fluid, rho, 12, "fp32", kernel, **SC)
    assert stats["nonfinite"] == 0
    if not fluid.any():
        assert stats["mass"] == 0.0
        return
    assert np.array_equal(g_f[fluid], st.fA[fluid].astype(np.float64))
    assert np.array_equal(g_rho[fluid], st.rho[fluid].astype(np.float64))
    assert abs(stats["mass"] / float(st.fA[fluid].astype(np.float64).sum()) - 1) < 1e-6
