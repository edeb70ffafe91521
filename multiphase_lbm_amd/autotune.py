"""Autotuner of the fused kernel's tile shape.

Times each candidate (threads in y, threads in z, planes per block) with HIP
events on the context stream and returns the fastest.  Tiling never changes
results (each candidate computes the same IEEE operations per node), so the
tuner may run on the live state; `autotune` re-initialises the fields from
`rho` afterwards when it is given.
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


def autotune(ctx, steps=6, rho=None, cands=None, verbose=False):
    lx, ly, lz = ctx.shape
    best = None
    results = []
    for ty, tz, seg in cands or candidates(lx, ly, lz):
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
