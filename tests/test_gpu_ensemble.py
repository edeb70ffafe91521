"""Batched ensemble (include/mplbm_ensemble.h): every sample of one
persistent launch is bit-identical to the oracle run of that sample alone,
with and without the captured hipGraph, for each workgroup size."""
import numpy as np
import pytest

from oracle import oracle
from tests.helpers import noisy_rho, random_spheres, sphere_mask

pytestmark = pytest.mark.gpu

from multiphase_lbm_amd.ensemble import Ensemble  # noqa: E402

SAMPLES = [
    dict(lx=20, ly=18, lz=24, G=-5.0, tau=1.0, rhos=0.7),
    dict(lx=12, ly=30, lz=70, G=-4.6, tau=0.9, rhos=0.8),   # lz > 64: ring tiles in z
    dict(lx=9, ly=11, lz=13, G=-5.4, tau=1.2, rhos=0.6),
]


def _inputs(k, s):
    shape = (s["lx"], s["ly"], s["lz"])
    fluid = sphere_mask(shape, random_spheres(shape, 3, 1.5, 3.5, 30 + k))
    rho = noisy_rho(shape, 0.7, 0.02, 40 + k)
    return shape, fluid, rho


@pytest.mark.parametrize("kernel", ["compact", "fused"])
@pytest.mark.parametrize("block", [256, 512, 1024])
@pytest.mark.parametrize("use_graph", [True, False])
def test_ensemble_samples_match_oracle(block, use_graph, kernel):
    print(f"ensemble {kernel} block={block} graph={use_graph}", flush=True)
    steps = 9
    with Ensemble(SAMPLES, block_threads=block, kernel=kernel) as ens:
        data = []
        for k, s in enumerate(SAMPLES):
            shape, fluid, rho = _inputs(k, s)
            ens.set_geometry(k, fluid)
            ens.init_fields(k, rho)
            data.append((shape, fluid, rho))
        ens.step(0, steps, write_macro=True, use_graph=use_graph)
        for k, (shape, fluid, rho) in enumerate(data):
            s = SAMPLES[k]
            P = oracle.Params(*shape, G=s["G"], tau=s["tau"], rhos=s["rhos"])
            st = oracle.State(P, fluid, rho).run(steps)
            f = ens.read_populations(k)
            r, p, u = ens.read_fields(k)
            assert np.array_equal(f[fluid], st.fA[fluid].astype(np.float64)), f"sample {k}"
            assert np.array_equal(r[fluid], st.rho[fluid]), f"sample {k}"
        assert ens.summary()["work_items"] > 0


def test_ensemble_per_sample_tiles_match_oracle():
    """Tiles and segment lengths set per sample (the autotuner's knobs) leave
    every sample bit-identical to its oracle run; a tile larger than the
    workgroup is refused and the previous tiles stay."""
    from multiphase_lbm_amd.ensemble import MplbmError
    steps = 7
    with Ensemble(SAMPLES, block_threads=256) as ens:
        data = []
        for k, s in enumerate(SAMPLES):
            shape, fluid, rho = _inputs(k, s)
            ens.set_geometry(k, fluid)
            ens.init_fields(k, rho)
            data.append((shape, fluid, rho))
        ens.set_tiling(0, 4, 16, 3)
        ens.set_tiling(1, 2, 64, 5)
        ens.set_tiling(2, 0, 0, 2)   # fitted tile, 2-plane segments
        assert ens.tiling(0) == (4, 16, 3) and ens.tiling(1)[2] == 5 and ens.tiling(2)[2] == 2
        with pytest.raises(MplbmError, match="exceeds"):
            ens.set_tiling(0, 16, 32, 8)   # 512 nodes > 256 threads
        assert ens.tiling(0) == (4, 16, 3)
        print("ensemble per-sample tiles", flush=True)
        ens.step(0, steps, write_macro=True)
        for k, (shape, fluid, rho) in enumerate(data):
            s = SAMPLES[k]
            P = oracle.Params(*shape, G=s["G"], tau=s["tau"], rhos=s["rhos"])
            st = oracle.State(P, fluid, rho).run(steps)
            assert np.array_equal(ens.read_populations(k)[fluid], st.fA[fluid].astype(np.float64)), f"sample {k}"
