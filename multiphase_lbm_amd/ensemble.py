"""Python view of the batched ensemble ABI (include/mplbm_ensemble.h): many
periodic packings of uneven sizes advanced by one persistent launch per
step, replayed from a hipGraph."""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _abi
from . import KERNEL_NAMES, MplbmError, STORAGE_NAMES, default_config

VP = ctypes.c_void_p
SIGNATURES = {
    "mplbm_ensemble_create": (ctypes.c_int, [ctypes.POINTER(_abi.Config), ctypes.c_int32, ctypes.c_int32,
                                             ctypes.POINTER(VP)]),
    "mplbm_ensemble_destroy": (None, [VP]),
    "mplbm_ensemble_last_error": (ctypes.c_char_p, [VP]),
    "mplbm_ensemble_set_tiling": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                 ctypes.c_int32]),
    "mplbm_ensemble_set_geometry": (ctypes.c_int, [VP, ctypes.c_int32, VP]),
    "mplbm_ensemble_init_fields": (ctypes.c_int, [VP, ctypes.c_int32, VP]),
    "mplbm_ensemble_step": (ctypes.c_int, [VP, ctypes.c_int64, ctypes.c_int32, ctypes.c_uint32, ctypes.c_int32]),
    "mplbm_ensemble_step_timed": (ctypes.c_int, [VP, ctypes.c_int64, ctypes.c_int32, ctypes.c_uint32,
                                                 ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]),
    "mplbm_ensemble_read_fields": (ctypes.c_int, [VP, ctypes.c_int32, VP, VP, VP]),
    "mplbm_ensemble_read_populations_f64": (ctypes.c_int, [VP, ctypes.c_int32, VP]),
    "mplbm_ensemble_reduce": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.POINTER(_abi.Stats64)]),
    "mplbm_ensemble_info": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.POINTER(_abi.Info)]),
    "mplbm_ensemble_summary": (ctypes.c_int, [VP, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                              ctypes.POINTER(ctypes.c_int32)]),
}

_bound = False


def _lib():
    global _bound
    lib = _abi.load()
    if not _bound:
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return lib


def _p(a):
    return VP(a.ctypes.data)


class Ensemble:
    """samples: list of dicts with keys lx, ly, lz and any mplbm_config
    physics field (G, tau, rhos, ...).  storage/math are shared; kernel is
    "compact" (fluid nodes only, the default) or "fused" (full lattice)."""

    def __init__(self, samples, storage="fp32", math="reference", block_threads=512, kernel="compact"):
        self.lib = _lib()
        n = len(samples)
        arr = (_abi.Config * n)()
        self.shapes = []
        for k, s in enumerate(samples):
            c = default_config()
            c.storage = STORAGE_NAMES[storage]
            c.math = {"reference": 0, "fast": 1}[math]
            c.kernel = KERNEL_NAMES[kernel]
            for key, v in s.items():
                setattr(c, key, v)
            arr[k] = c
            self.shapes.append((int(s["lx"]), int(s["ly"]), int(s["lz"])))
        h = VP()
        st = self.lib.mplbm_ensemble_create(arr, n, int(block_threads), ctypes.byref(h))
        if st != _abi.OK:
            raise MplbmError(st, self.lib.mplbm_ensemble_last_error(None).decode())
        self.h = h
        self.n = n

    def _check(self, st):
        if st != _abi.OK:
            raise MplbmError(st, self.lib.mplbm_ensemble_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.mplbm_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_tiling(self, k, tile_y=0, tile_z=0, seg_x=0):
        """Compact tile of sample k (k = -1: every sample); 0 keeps the
        fitted value of that field."""
        self._check(self.lib.mplbm_ensemble_set_tiling(self.h, int(k), int(tile_y), int(tile_z), int(seg_x)))

    def tiling(self, k):
        """(tile_y, tile_z, seg_x) in use for sample k."""
        i = _abi.Info()
        self._check(self.lib.mplbm_ensemble_info(self.h, int(k), ctypes.byref(i)))
        return i.tile_y, i.tile_z, i.seg_x

    def set_geometry(self, k, is_fluid):
        m = np.ascontiguousarray(np.asarray(is_fluid).astype(np.uint8))
        assert m.shape == self.shapes[k]
        self._check(self.lib.mplbm_ensemble_set_geometry(self.h, k, _p(m)))

    def init_fields(self, k, rho):
        r = np.ascontiguousarray(rho, np.float32)
        assert r.shape == self.shapes[k]
        self._check(self.lib.mplbm_ensemble_init_fields(self.h, k, _p(r)))

    def step(self, step=0, nsteps=1, write_macro=False, use_graph=True):
        self._check(self.lib.mplbm_ensemble_step(self.h, int(step), int(nsteps), 1 if write_macro else 0,
                                                 1 if use_graph else 0))

    def step_timed(self, step=0, nsteps=1, write_macro=False, use_graph=True) -> float:
        ms = ctypes.c_double()
        self._check(self.lib.mplbm_ensemble_step_timed(self.h, int(step), int(nsteps), 1 if write_macro else 0,
                                                       1 if use_graph else 0, ctypes.byref(ms)))
        return ms.value

    def read_fields(self, k):
        shape = self.shapes[k]
        rho = np.empty(shape, np.float32)
        p = np.empty(shape, np.float32)
        u = np.empty(shape + (3,), np.float32)
        self._check(self.lib.mplbm_ensemble_read_fields(self.h, k, _p(rho), _p(p), _p(u)))
        return rho, p, u

    def read_populations(self, k):
        f = np.empty(self.shapes[k] + (19,), np.float64)
        self._check(self.lib.mplbm_ensemble_read_populations_f64(self.h, k, _p(f)))
        return f

    def reduce(self, k):
        s = _abi.Stats64()
        self._check(self.lib.mplbm_ensemble_reduce(self.h, k, ctypes.byref(s)))
        return s.as_dict()

    def summary(self):
        wi, fn, bt = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        self._check(self.lib.mplbm_ensemble_summary(self.h, ctypes.byref(wi), ctypes.byref(fn), ctypes.byref(bt)))
        return {"work_items": wi.value, "fluid_nodes": fn.value, "block_threads": bt.value}


def autotune_block(make_ensemble, steps=4):
    """Time the ensemble at workgroup sizes 256, 512, 1024 (tiles are refit
    per sample for each) and return (best_block, {block: ms_per_step})."""
    times = {}
    for b in (256, 512, 1024):
        try:
            ens = make_ensemble(b)
        except MplbmError:
            continue
        with ens:
            ens.step(0, 2)
            times[b] = ens.step_timed(2, steps) / steps
    best = min(times, key=times.get)
    return best, times


# sample-size classes of the tile search: largest axis up to 80, up to 128, above
SIZE_CLASSES = ((0, 80), (81, 128), (129, 1 << 30))


def size_class(shape):
    n = max(shape)
    return next(k for k, (lo, hi) in enumerate(SIZE_CLASSES) if lo <= n <= hi)


def tile_candidates(block):
    """Compact tiles (tile_y, tile_z) of at most `block` interior nodes; with
    the workgroup size they fix the interior nodes per thread (area / block)."""
    tiles = ((4, 32), (8, 32), (16, 32), (32, 32), (8, 16), (16, 16), (2, 64), (4, 64), (8, 64), (16, 64))
    # fewer than half as many nodes as threads leaves most of the workgroup
    # idle at every plane (measured slower for every class)
    return [(ty, tz) for ty, tz in tiles if block // 2 <= ty * tz <= block]


def autotune(make_ensemble, shapes, blocks=(256, 512), segs=(16, 32, 48, 64), steps=4, log=None):
    """Ensemble autotuner: for each workgroup size, a coordinate search over
    the sample-size classes (largest first) that picks the compact tile
    (tile_y, tile_z) with 32-plane segments, then the segment length, of
    the samples of that class, keeping each choice that lowers the step
    time.  make_ensemble(block) returns a set-up Ensemble.  Returns
    (block, {class: (tile_y, tile_z, seg_x)}, ms_per_step, trace) with the
    trace listing every timed configuration."""
    classes = sorted({size_class(s) for s in shapes}, reverse=True)
    members = {c: [k for k, s in enumerate(shapes) if size_class(s) == c] for c in classes}
    trace = []
    best_all = None

    def timed(ens):
        ens.step(0, 2)
        return ens.step_timed(2, steps) / steps

    def apply(ens, cls, tile):
        for k in members[cls]:
            ens.set_tiling(k, *tile)

    for b in blocks:
        try:
            ens = make_ensemble(b)
        except MplbmError:
            continue
        with ens:
            cur = timed(ens)
            picks = {c: (0, 0, 0) for c in classes}
            trace.append((b, dict(picks), cur))
            if log:
                log(f"block {b}: fitted tiles {cur:.4f} ms")
            for cls in classes:
                for stage in ("tile", "seg"):
                    best, applied = picks[cls], picks[cls]
                    if stage == "tile":
                        cands = [(ty, tz, 32) for ty, tz in tile_candidates(b)]
                    else:
                        cands = [(best[0], best[1], sg) for sg in segs]
                    for cand in cands:
                        if cand == best:
                            continue
                        t0 = time.perf_counter()
                        try:
                            apply(ens, cls, cand)
                        except MplbmError:
                            apply(ens, cls, applied)
                            continue
                        applied = cand
                        t1 = time.perf_counter()
                        t = timed(ens)
                        trace.append((b, {**picks, cls: cand}, t))
                        if log:
                            log(f"block {b} class {cls} tile {cand}: {t:.4f} ms (set-up {t1 - t0:.1f} s, "
                                f"timing {time.perf_counter() - t1:.1f} s)")
                        if t < cur:
                            cur, best = t, cand
                    if applied != best:
                        apply(ens, cls, best)
                    picks[cls] = best
            if best_all is None or cur < best_all[2]:
                best_all = (b, dict(picks), cur)
    return best_all[0], best_all[1], best_all[2], trace
