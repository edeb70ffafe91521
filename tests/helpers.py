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
