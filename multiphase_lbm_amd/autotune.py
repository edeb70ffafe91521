"""Autotuner of the marching kernels' tile shape.

Times each candidate (tile in y, tile in z, planes per block) with HIP events
on the context stream and returns the fastest.  For FUSED a tile counts
threads including the one-cell ring; for COMPACT it counts interior nodes
(the workgroup is 256 threads up to 256 interior nodes, else 512).  Tiling
never changes results (each candidate computes the same IEEE operations per
node), so the tuner may run on the live state; `autotune` re-initialises the
fields from `rho` afterwards when it is given.
"""
from __future__ import annotations

import itertools


def candidates(lx, ly, lz):
    """Tile shapes for the domain: rings of one cell (interior = threads - 2)
    or whole axes that wrap inside the tile (lz <= 64)."""
    out = []
    tzs = [lz] if lz <= 64 else [18, 34, 66, 130]
    for tz, ty in itertools.product(tzs, [3, 4, 5, 6, 7, 8, 11, 15, 22, 30]):
        t = ty * tz
        if t > 1024 or t < 128 or ty > ly:
            continue
        for seg in (32, 64, 128):
            if seg <= lx:
                out.append((ty, tz, seg))
    return out


def compact_candidates(lx, ly, lz):
    """Interior tiles of the compact kernel: z spans of 32 or 64 (whole
    32-node rank words) or the whole axis when it is shorter."""
    out = []
    tzs = [lz] if lz <= 32 else [32, 64]
    for tz, ty in itertools.product(tzs, [4, 8, 12, 16, 24, 32]):
        if ty > ly or ty * tz > 2048:
            continue
        for seg in (16, 32, 48, 64):
            if seg <= lx:
                out.append((ty, tz, seg))
    return out


def autotune(ctx, steps=6, rho=None, cands=None, verbose=False):
    from . import KERNEL_COMPACT
    lx, ly, lz = ctx.shape
    best = None
    results = []
    if cands is None:
        compact = ctx.info()["kernel"] == KERNEL_COMPACT
        cands = compact_candidates(lx, ly, lz) if compact else candidates(lx, ly, lz)
    for ty, tz, seg in cands:
        try:
            ctx.set_tiling(ty, tz, seg)
        except Exception:
            continue
        ctx.step(0, 1)
        ms = ctx.step_timed(0, steps) / steps
        results.append(((ty, tz, seg), ms))
        if verbose:
            print(f"# tile y={ty} z={tz} seg={seg}: {ms:.3f} ms/step", flush=True)
        if best is None or ms < best[1]:
            best = ((ty, tz, seg), ms)
    if best is not None:
        ctx.set_tiling(*best[0])
    if rho is not None:
        ctx.init_fields(rho)
    return best, results
