"""Synthetic inputs for the tests: seeded sphere packings and density fields."""
import numpy as np


def random_spheres(shape, n, rmin, rmax, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 1, size=(n, 3)) * np.array(shape)
    r = rng.uniform(rmin, rmax, size=n)
    return np.concatenate([c, r[:, None]], axis=1)


def sphere_mask(shape, spheres):
    """is_fluid [x][y][z] (True = fluid) for periodic spheres (min-image test
    at voxel centres)."""
    fluid = np.ones(shape, bool)
    grids = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    for cx, cy, cz, r in spheres:
        d2 = np.zeros(shape)
        for g, c, L in zip(grids, (cx, cy, cz), shape):
            d = np.abs(g - c)
            d = np.minimum(d, L - d)
            d2 += d * d
        fluid &= d2 >= r * r
    return fluid


def noisy_rho(shape, rho0, amp, seed, dtype=np.float32):
    rng = np.random.default_rng(seed)
    return (rho0 + rng.uniform(-amp, amp, size=shape)).astype(dtype)


# Density relative-L2 bar against FP64 after 100 steps for the 16-bit
# storage forms: 1e-3 (the specification's bar).  BF16 stored against the
# global shift (round-to-nearest or stochastic) is the one stated exception,
# held to 1e-2: 8 mantissa bits for deviations up to |rho - rho_ref| ~ 1 in a
# two-phase field; measured 1.2e-3 to 1.7e-3 on the small cases here.  BF16
# with node shifts meets 1e-3.
DENSITY_L2_BAR = {"fp16s": 1e-3, "fp16sr": 1e-3, "fp16sn": 1e-3, "fp16srn": 1e-3,
                  "bf16s": 1e-2, "bf16sr": 1e-2, "bf16sn": 1e-3, "bf16srn": 1e-3, "fp32": 1e-3}
