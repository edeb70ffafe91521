"""GPU cluster statistics (mplbm_calc_cluster) against the host harness.

The harness's cluster_stats restates calc_cluster (src/multiphase_lbm.cu:
225-562) with an iterative depth-first search in the reference's scan order
and is pinned by hand-built lattices in tests/test_harness.py.  Here the
device union-find labelling must give the same LiqClusters / GasClusters
arrays, counts, boundary ratios, extrema, cluster order, wet coordination
number and single-cluster radius on the same rho / p fields (read back from
the device).  Counts, labels and float results are compared exactly; the
phase averages are FP64 sums in another order, so they are compared to 1e-6
relative after their cast to float.
"""
import numpy as np
import pytest

from multiphase_lbm_amd import harness

pytestmark = pytest.mark.gpu

import multiphase_lbm_amd as mp  # noqa: E402

RHOL, RHOG = 1.285931, 0.323507
EXACT = ("min_rho", "max_rho", "min_p", "max_p", "saturation", "liq_cluster_count", "gas_cluster_count",
         "cluster_order", "wet_coord_num", "ratio_liquid_solid_boundary", "ratio_liquid_vapor_boundary",
         "cluster_radius", "fluid_count", "liquid_count", "vapor_count")
AVERAGES = ("avg_liq_press", "avg_vap_press", "avg_liq_rho", "avg_vap_rho")


def _packing(shape, nsph, rmin, rmax, seed):
    """Spheres voxelised by place_sphere with their 1-based solid numbers."""
    rng = np.random.default_rng(seed)
    fluid = np.ones(shape, np.uint8)
    sn = np.zeros(shape, np.int64)
    for k in range(nsph):
        c = rng.uniform(0, 1, 3) * np.array(shape)
        harness.place_sphere(fluid, float(c[0]), float(c[1]), float(c[2]), float(rng.uniform(rmin, rmax)), sn, k + 1)
    return fluid, sn, nsph


def _periodic_d2(shape, c):
    d2 = np.zeros(shape, np.float64)
    for a, L in enumerate(shape):
        v = np.abs(np.arange(L) - c[a])
        v = np.minimum(v, L - v)
        sh = [1, 1, 1]
        sh[a] = L
        d2 = d2 + (v ** 2).reshape(sh)
    return d2


def _blobs(shape, seed, nblobs=12, ndrops=25):
    """Vapour background, liquid blobs (some across the periodic faces) with
    intermediate shells, and single-node liquid and vapour drops (clusters
    below 7 nodes)."""
    rng = np.random.default_rng(seed)
    rho = np.full(shape, 0.15, np.float32)
    for _ in range(nblobs):
        c = rng.uniform(0, 1, 3) * np.array(shape)
        r = rng.uniform(1.2, 4.5)
        d2 = _periodic_d2(shape, c)
        rho[(d2 >= r * r) & (d2 < (r + 1.0) ** 2) & (rho < 1.0)] = 0.7
        rho[d2 < r * r] = rng.uniform(1.6, 2.0)
    idx = rng.integers(0, rho.size, ndrops)
    rho.reshape(-1)[idx[: ndrops // 2]] = 1.9
    rho.reshape(-1)[idx[ndrops // 2:]] = 0.05
    return rho


def _compare(ctx, fluid, sn, nsolid):
    rho, p, _ = ctx.read_fields(np.float32)
    got, gl, gg = ctx.calc_cluster(RHOL, RHOG, want_labels=True)
    # thresholds as the reference's floats (the driver passes them so too)
    ref, rl, rg = harness.cluster_stats(fluid, rho, p, sn, float(np.float32(RHOL)), float(np.float32(RHOG)), nsolid,
                                        want_labels=True)
    assert np.array_equal(gl.astype(np.int64), rl), "LiqClusters labels differ"
    assert np.array_equal(gg.astype(np.int64), rg), "GasClusters labels differ"
    for k in EXACT:
        assert got[k] == ref[k], (k, got[k], ref[k])
    for k in AVERAGES:
        assert got[k] == pytest.approx(ref[k], rel=1e-6, abs=1e-12), (k, got[k], ref[k])
    return got


@pytest.mark.parametrize("storage,kernel", [("fp32", "compact"), ("fp32", "fused"), ("fp64", "compact"),
                                            ("fp16s", "compact")])
def test_blobs_match_harness(storage, kernel):
    shape = (26, 22, 18)
    print(f"case clusters blobs {storage} {kernel} shape={shape}", flush=True)
    fluid, sn, nsolid = _packing(shape, 6, 2.0, 4.0, seed=5)
    rho = _blobs(shape, seed=7)
    with mp.Context(*shape, storage=storage, kernel=kernel, G=-5.0, rhol_spinodal=RHOL, rhog_spinodal=RHOG) as ctx:
        ctx.set_geometry(fluid)
        ctx.set_solid_numbers(sn, nsolid)
        ctx.init_fields(rho.astype(np.float64) if storage == "fp64" else rho)
        got = _compare(ctx, fluid, sn, nsolid)
    assert got["liq_cluster_count"] > 1 and got["gas_cluster_count"] >= 1
    assert got["cluster_order"] > 0


def test_single_cluster_across_faces_radius():
    # one liquid slab crossing the x and z faces: unwrapping and radius
    shape = (20, 16, 24)
    print(f"case clusters single shape={shape}", flush=True)
    fluid, sn, nsolid = _packing(shape, 3, 1.5, 2.5, seed=11)
    rho = np.full(shape, 0.15, np.float32)
    x = np.arange(shape[0])[:, None, None]
    y = np.arange(shape[1])[None, :, None]
    z = np.arange(shape[2])[None, None, :]
    slab = ((x < 4) | (x >= shape[0] - 3)) & (y >= 3) & (y < 11) & ((z < 5) | (z >= shape[2] - 6))
    rho[slab] = 1.9
    with mp.Context(*shape, storage="fp32", kernel="compact", G=-5.0) as ctx:
        ctx.set_geometry(fluid)
        ctx.set_solid_numbers(sn, nsolid)
        ctx.init_fields(rho)
        got = _compare(ctx, fluid, sn, nsolid)
    assert got["liq_cluster_count"] == 1
    assert got["cluster_radius"] > 0


def test_after_two_phase_steps():
    # spinodal decomposition in a packing, macro fields of the last step
    shape = (40, 36, 32)
    print(f"case clusters dynamic shape={shape}", flush=True)
    fluid, sn, nsolid = _packing(shape, 10, 2.5, 5.0, seed=2)
    rng = np.random.default_rng(4)
    rho = (0.7 * (1 + 0.02 * (rng.random(shape) - 0.5))).astype(np.float32)
    with mp.Context(*shape, storage="fp32", kernel="compact", G=-5.0, tau=1.0, rhos=0.7) as ctx:
        ctx.set_geometry(fluid)
        ctx.set_solid_numbers(sn, nsolid)
        ctx.init_fields(rho)
        ctx.step(0, 400, write_macro=True)
        got = _compare(ctx, fluid, sn, nsolid)
    assert got["liquid_count"] > 0 and got["vapor_count"] > 0


def test_interfaces_after_short_run():
    # early spinodal decomposition: many interface nodes near the thresholds
    shape = (96, 96, 64)
    print(f"case clusters many shape={shape}", flush=True)
    fluid, sn, nsolid = _packing(shape, 40, 3.0, 7.0, seed=8)
    rng = np.random.default_rng(6)
    rho = (0.7 * (1 + 0.02 * (rng.random(shape) - 0.5))).astype(np.float32)
    with mp.Context(*shape, storage="fp32", kernel="compact", G=-5.0, tau=1.0, rhos=0.7) as ctx:
        ctx.set_geometry(fluid)
        ctx.set_solid_numbers(sn, nsolid)
        ctx.init_fields(rho)
        ctx.step(0, 200, write_macro=True)
        got = _compare(ctx, fluid, sn, nsolid)
    assert got["liquid_count"] > 0 and got["vapor_count"] > 0


def test_large_domain_and_no_solid_numbers():
    # union-find on a larger lattice with long periodic clusters; without
    # solid numbers the cluster order and wet coordination number are 0
    shape = (96, 80, 64)
    print(f"case clusters large shape={shape}", flush=True)
    fluid = np.ones(shape, np.uint8)
    rng = np.random.default_rng(9)
    rho = np.where(rng.random(shape) < 0.35, 1.9, 0.15).astype(np.float32)
    with mp.Context(*shape, storage="fp32", kernel="compact") as ctx:
        ctx.set_geometry(fluid)
        ctx.init_fields(rho)
        got = _compare(ctx, fluid, None, 0)
    assert got["cluster_order"] == 0 and got["wet_coord_num"] == 0


def test_no_liquid():
    shape = (8, 9, 10)
    fluid = np.ones(shape, np.uint8)
    rho = np.full(shape, 0.7, np.float32)
    with mp.Context(*shape, storage="fp32", kernel="compact") as ctx:
        ctx.set_geometry(fluid)
        ctx.init_fields(rho)
        got = _compare(ctx, fluid, None, 0)
    assert got["liq_cluster_count"] == 0 and got["gas_cluster_count"] == 0 and got["saturation"] == 0
