#!/usr/bin/env python3
"""Throughput of the D3Q19 Shan-Chen step on one MI355X.

Workload (the packing the headline metric is quoted on): 512^3 domain,
overlapping spheres with radii uniform in [6, 18] placed by a seeded
generator until the voxel porosity is 0.45, Shan-Chen G = -5, tau = 1,
rho0 = 0.7 +/- 0.01 (srand(1) initialisation of the reference).  A step is
one stream-collide launch over the whole lattice (by default the compact
kernel, which stores and visits fluid nodes only).

Prints one JSON line: value = fluid-node MLUPS of the FP32 form (whole job,
inputs resident in HBM), the roofline object of the step kernel (the compact
stream-collide kernel unless --kernel says otherwise; algorithmic bytes
2*19*4 = 152 per fluid update / HIP-event-timed average launch duration on
the context stream; traffic from the rocprofv3 counter runs recorded in
profiles/traffic_512.txt), the FP16- and BF16-storage forms beside it with
their error against the FP64 path, and the CPU baseline (the oracle,
OpenMP, on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
Q = 19


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--porosity", type=float, default=0.45)
    ap.add_argument("--rmin", type=float, default=6.0)
    ap.add_argument("--rmax", type=float, default=18.0)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--storage", default="fp32,fp32fast,fp16s,bf16s,fp16sr,fp16sn",
                    help="fp32, fp32fast, fp16s, bf16s, fp64; fp16sr/bf16sr: 16-bit with stochastic rounding; "
                         "fp16sn/bf16sn: 16-bit with node shifts; fp16srn/bf16srn: both")
    ap.add_argument("--kernel", default="compact", help="compact (fluid nodes only), fused or twopass")
    ap.add_argument("--tile-y", type=int, default=0)
    ap.add_argument("--tile-z", type=int, default=0)
    ap.add_argument("--seg-x", type=int, default=0)
    ap.add_argument("--block-threads", type=int, default=0, help="compact kernel workgroup size (0, 256, 512)")
    ap.add_argument("--bounce-back", type=int, default=1, help="0: write-side (ghost) stores, 1: pull-side reads")
    ap.add_argument("--cpu-steps", type=int, default=10, help="timed oracle steps per CPU baseline run (0: skip)")
    ap.add_argument("--cpu-warmup", type=int, default=2, help="untimed oracle steps before the CPU baseline runs")
    ap.add_argument("--cpu-repeats", type=int, default=3, help="CPU baseline runs; the median is reported")
    ap.add_argument("--error-steps", type=int, default=100,
                    help="steps of the error check of each storage form against FP64 (0: skip)")
    ap.add_argument("--profile-only", action="store_true",
                    help="run only the timed loop of the first storage form (with --ensemble: of each form, fitted "
                         "tiles, no autotuner, no saturation run), for rocprofv3")
    ap.add_argument("--precision", action="store_true", help="error study of FP32/FP16S/BF16S against FP64")
    ap.add_argument("--precision-steps", type=int, default=1000)
    ap.add_argument("--slab-steps", type=int, default=3000)
    ap.add_argument("--perm-steps", type=int, default=3000)
    ap.add_argument("--sweep", action="store_true", help="time every autotuner tile candidate (FP32) and exit")
    ap.add_argument("--autotune", action="store_true", help="pick the tile with the autotuner before timing")
    ap.add_argument("--ensemble", action="store_true",
                    help="batched ensemble: 64 packings n^3, n~U{48..160}, one persistent launch per step")
    ap.add_argument("--ensemble-samples", type=int, default=64)
    ap.add_argument("--ensemble-steps", type=int, default=5000, help="steps before the saturation read-out")
    ap.add_argument("--drainage", action="store_true",
                    help="768x512x512 packing between a liquid inlet and a vapour outlet: FP32 vs FP16 throughput, "
                         "and FP64/FP32/FP16 saturation on a 1/8 sub-volume")
    ap.add_argument("--drainage-steps", type=int, default=20000)
    ap.add_argument("--clusters", action="store_true",
                    help="calc_cluster statistics of the 512^3 two-phase run: device (mplbm_calc_cluster) against "
                         "the host path (field read + host cluster_stats), same results")
    ap.add_argument("--cluster-steps", type=int, default=2000)
    ap.add_argument("--voxeliser", action="store_true",
                    help="voxelise the spheres of the 512^3 packing on the device (mplbm_place_spheres) and on the "
                         "host (place_sphere per sphere), same mask and solid numbers")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain device-to-device copy kernel of 2 GiB (rocprofv3 --pmc calibration)")
    return ap.parse_args()


def workload(a):
    from multiphase_lbm_amd import harness as H
    shape = (a.size, a.size, a.size)
    t = time.time()
    fluid, sph = H.generate_packing(shape, a.seed, 10 ** 8, a.rmin, a.rmax, a.porosity, nonoverlap=False)
    rho = H.rho_init(shape, 0.7, 0.02 / 0.7)
    print(f"# workload {shape}: {len(sph)} spheres, porosity {fluid.mean():.4f}, built in {time.time() - t:.1f}s",
          file=sys.stderr, flush=True)
    return shape, fluid, rho, sph


PHYS = dict(tau=1.0, G=-5.0, rhos=0.7, rhol_spinodal=1.285931, rhog_spinodal=0.323507)
# storage names of bench.py and the stored form they time
BASE_STORAGE = {"fp32fast": "fp32", "fp16sr": "fp16s", "bf16sr": "bf16s", "fp16sn": "fp16s", "bf16sn": "bf16s",
                "fp16srn": "fp16s", "bf16srn": "bf16s"}


def bytes_per_update(mp, storage):
    """Algorithmic bytes of one fluid-node update: 19 populations read and
    written, plus the node's 16-bit shift read and written in the node-shifted
    forms."""
    b = 2 * Q * mp.STORAGE_BYTES[mp.STORAGE_NAMES[BASE_STORAGE.get(storage, storage)]]
    return b + (4 if storage in ("fp16sn", "bf16sn", "fp16srn", "bf16srn") else 0)


def run_gpu(a, shape, fluid, rho, storage):
    import multiphase_lbm_amd as mp
    ctx = mp.Context(*shape, storage=storage, kernel=a.kernel, tile_y=a.tile_y, tile_z=a.tile_z, seg_x=a.seg_x,
                     bounce_back=a.bounce_back, block_threads=a.block_threads, **PHYS)
    try:
        ctx.set_geometry(fluid)
        ctx.init_fields(rho)
        if a.autotune:
            from multiphase_lbm_amd.autotune import autotune
            best, _ = autotune(ctx, steps=5, rho=rho)
            print(f"# autotune {storage}: tile y,z,seg = {best[0]} ({best[1]:.3f} ms/step)", file=sys.stderr,
                  flush=True)
        info = ctx.info()
        ctx.step(0, a.warmup)
        ctx.sync()
        t0 = time.perf_counter()
        ev_ms = ctx.step_timed(a.warmup, a.steps)  # HIP events on the context stream around the K launches
        ctx.sync()
        wall = time.perf_counter() - t0
        stats = ctx.reduce()
    finally:
        ctx.close()
    nf = info["fluid_nodes"]
    launches = a.steps * (2 if a.kernel == "twopass" else 1)
    per_launch_s = ev_ms / 1e3 / launches
    bpu = bytes_per_update(mp, storage)
    alg_bytes = bpu * nf  # per step
    return {
        "mlups": nf * a.steps / wall / 1e6,
        "mlups_event": nf * a.steps / (ev_ms / 1e3) / 1e6,
        "ms_per_step": wall * 1e3 / a.steps,
        "kernel_ms": per_launch_s * 1e3,
        "achieved_gbs": alg_bytes / (ev_ms / 1e3 / a.steps) / 1e9,
        "bytes_per_update": bpu,
        "fluid_nodes": nf,
        "tile": [info["tile_y"], info["tile_z"], info["seg_x"]],
        "tag": f"{a.kernel}:{info['tile_y']}x{info['tile_z']}x{info['seg_x']}:bb{a.bounce_back}",
        "mass": stats["mass"],
        "nonfinite": stats["nonfinite"],
    }


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(a, shape, fluid, rho):
    """The oracle (C, OpenMP, the reference's float arithmetic in its 3-sweep
    AoS structure: stream, boundary, collide) on the same 512^3 workload:
    cpu_warmup untimed steps, then cpu_repeats runs of cpu_steps timed steps;
    the median run is reported."""
    from oracle import oracle
    P = oracle.Params(*shape, **PHYS)
    st = oracle.State(P, fluid, rho, dtype=np.float32)
    st.run(a.cpu_warmup)
    nf = int(fluid.sum())
    runs = []
    for _ in range(a.cpu_repeats):
        t = time.perf_counter()
        st.run(a.cpu_steps)
        runs.append(time.perf_counter() - t)
    del st
    dt = sorted(runs)[len(runs) // 2]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(nf * a.cpu_steps / dt / 1e6, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"reference algorithm on CPU (restated): the oracle, {threads} OpenMP threads, "
                      f"{a.cpu_warmup} warm-up steps then {a.cpu_steps} timed steps of the same "
                      f"{shape[0]}^3 packing, median of {a.cpu_repeats} runs "
                      f"({', '.join(f'{r:.2f}' for r in runs)} s)"}


def error_check(a, shape, fluid, rho, storages):
    """Error of each storage form against the FP64 path (bit-identical to the
    FP64 oracle on every parity case) from the same start: density relative
    L2 over fluid nodes after error_steps steps of the G = -5 run on the
    512^3 packing, total mass drift (FP64 sums of the stored populations),
    and the Darcy permeability (G = 0, tau = 0.8, g = 1e-6 along x, 600
    steps) on the same packing generator at 64^3 with radii / 4."""
    import multiphase_lbm_amd as mp
    from multiphase_lbm_amd import harness as H
    fl = fluid.astype(bool)
    m0 = float(rho.astype(np.float64)[fl].sum())
    steps = a.error_steps

    def run(storage, rho0):
        with mp.Context(*shape, storage=storage, kernel=a.kernel, **PHYS) as ctx:
            ctx.set_geometry(fluid)
            ctx.init_fields(rho0)
            ctx.step(0, steps, write_macro=True)
            r, _, _ = ctx.read_fields(np.float64)
            return r, ctx.reduce()["mass"]

    ref, _ = run("fp64", rho.astype(np.float64))
    pshape = (64, 64, 64)
    pfl, _ = H.generate_packing(pshape, a.seed, 10 ** 6, a.rmin / 4, a.rmax / 4, a.porosity)
    tau, g = 0.8, 1e-6

    def darcy(storage):
        with mp.Context(*pshape, storage=storage, kernel=a.kernel, tau=tau, G=0.0, rhos=1.0,
                        body_force=(g, 0.0, 0.0)) as ctx:
            ctx.set_geometry(pfl)
            ctx.init_fields(np.ones(pshape, np.float64 if storage == "fp64" else np.float32))
            ctx.step(0, 600, write_macro=True)
            _, _, u = ctx.read_fields(np.float64)
        ux = np.where(pfl.astype(bool), u[..., 0] - (tau - 0.5) * g, 0.0)
        return (tau - 0.5) / 3 * float(ux.mean()) / g

    k64 = darcy("fp64")
    out = {}
    for s in storages:
        r, mass = run(s, rho)
        k = darcy(s)
        out[s] = {"density_rel_l2": float(np.linalg.norm(r[fl] - ref[fl]) / np.linalg.norm(ref[fl])),
                  "mass_rel_drift": mass / m0 - 1.0, "permeability_rel_err": abs(k / k64 - 1.0),
                  "against": f"FP64 path, {steps} steps of the G=-5 run on this packing; permeability: 600 steps "
                             f"at 64^3 (radii/4, porosity {a.porosity}), k_fp64 = {k64:.6f}"}
        print(f"# error {s}: {out[s]}", file=sys.stderr, flush=True)
    return out


TRAFFIC_WORKLOAD = (512, 0.45, 6.0, 18.0, 3)  # size, porosity, rmin, rmax, seed of profiles/traffic_512.txt


def traffic_from_profiles(a, storage, tag):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, each scaled by the copy-kernel calibration) for this
    workload and kernel configuration, from the `traffic` lines of
    profiles/traffic_512.txt; None when not measured (another packing or
    size, or a configuration without a line)."""
    if (a.size, a.porosity, a.rmin, a.rmax, a.seed) != TRAFFIC_WORKLOAD:
        return None
    path = os.path.join(ROOT, "profiles", "traffic_512.txt")
    try:
        with open(path) as f:
            for line in f:
                w = line.split()
                if len(w) == 4 and w[0] == "traffic" and w[1] == storage and w[2] == tag:
                    return int(w[3])
    except OSError:
        pass
    return None


def precision_study(a, shape, fluid, rho):
    """Error of the 16-bit storage forms (and FP32) against FP64 on the same
    packing.  The FP64 GPU path evaluates the oracle's FP64 operation
    sequence (bit-identical in tests/test_gpu_parity.py), so its fields stand
    for the FP64 oracle at sizes the CPU cannot run.  Prints one JSON line."""
    import multiphase_lbm_amd as mp
    out = {"workload": f"{shape[0]}^3 packing, porosity {fluid.mean():.3f}"}
    fl = fluid.astype(bool)

    def run(storage, params, steps, rho0, early=0, **kw):
        """steps steps from rho0; with early > 0 also the density after the
        first `early` steps (returned last, else None)."""
        with mp.Context(*shape, storage=storage, kernel=a.kernel, **params, **kw) as ctx:
            ctx.set_geometry(fluid)
            ctx.init_fields(rho0)
            t = time.perf_counter()
            r_early = None
            if early:
                ctx.step(0, early, write_macro=True)
                r_early, _, _ = ctx.read_fields()
            ctx.step(early, steps - early, write_macro=True)
            r, p, u = ctx.read_fields()
            dt = time.perf_counter() - t
            st = ctx.reduce()
            info = ctx.info()
        return r, u, st, dt, info, r_early

    # (1) two-phase Shan-Chen run, G = -5
    steps = a.precision_steps
    early = min(100, steps)  # before spinodal decomposition amplifies rounding differences
    ref_r, _, ref_st, _, _, ref_early = run("fp64", PHYS, steps, rho.astype(np.float64), early)
    sat_thr = 0.5 * (1.932442 + 0.156413)
    two = {"steps": steps, "fp64": {"saturation": float((ref_r[fl] > sat_thr).mean()),
                                    "mass_rel_drift": ref_st["mass"] / float(rho.astype(np.float64)[fl].sum()) - 1}}
    variants = [("fp32", {}), ("fp16s", {}), ("fp16s", {"rounding": "stochastic"}), ("fp16s", {"shift": "node"}),
                ("fp16s", {"rounding": "stochastic", "shift": "node"}),
                ("bf16s", {}), ("bf16s", {"rounding": "stochastic"}), ("bf16s", {"shift": "node"}),
                ("bf16s", {"rounding": "stochastic", "shift": "node"})]

    def vkey(storage, kw):
        return (storage + ("_stochastic" if kw.get("rounding") == "stochastic" else "")
                + ("_node" if kw.get("shift") == "node" else ""))

    for storage, kw in variants:
        r, _, st, dt, info, r_early = run(storage, PHYS, steps, rho, early, **kw)
        key = vkey(storage, kw)
        two[key] = {
            "density_rel_l2": float(np.linalg.norm(r[fl] - ref_r[fl]) / np.linalg.norm(ref_r[fl])),
            f"density_rel_l2_step{early}": float(np.linalg.norm(r_early[fl] - ref_early[fl]) /
                                                  np.linalg.norm(ref_early[fl])),
            "density_max_abs": float(np.abs(r[fl] - ref_r[fl]).max()),
            "saturation": float((r[fl] > sat_thr).mean()),
            "mass_rel_drift_fp64_sum": st["mass"] / float(rho.astype(np.float64)[fl].sum()) - 1,
            "rho_ref": info["rho_ref"], "scale": info["dev_scale"], "nonfinite": st["nonfinite"]}
        print(f"# two-phase {key}: {two[key]}", file=sys.stderr, flush=True)
    out["two_phase_G-5"] = two

    # (2) flat interface: liquid slab in vapour, 128x16x16 periodic
    sshape = (128, 16, 16)
    xs = np.arange(128)
    slab = np.where((xs >= 32) & (xs < 96), 1.9, 0.16)[:, None, None] * np.ones(sshape)
    sfl = np.ones(sshape, np.uint8)

    def run_slab(storage, **kw):
        with mp.Context(*sshape, storage=storage, kernel=a.kernel, tau=1.0, G=-5.0, rhos=0.7, **kw) as ctx:
            ctx.set_geometry(sfl)
            ctx.init_fields(slab.astype(np.float64 if storage == "fp64" else np.float32))
            ctx.step(0, a.slab_steps, write_macro=True)
            r, _, _ = ctx.read_fields()
        prof = r.mean(axis=(1, 2))
        mid = 0.5 * (prof.max() + prof.min())
        # first up-crossing of the midpoint density, linear interpolation
        k = int(np.argmax((prof[:-1] < mid) & (prof[1:] >= mid)))
        pos = k + (mid - prof[k]) / (prof[k + 1] - prof[k])
        return {"rho_l": float(prof[64]), "rho_v": float(prof[0]), "interface_x": float(pos)}

    sl = {"steps": a.slab_steps, "fp64": run_slab("fp64")}
    for storage, kw in variants:
        key = vkey(storage, kw)
        v = run_slab(storage, **kw)
        v["rho_l_rel_err"] = abs(v["rho_l"] - sl["fp64"]["rho_l"]) / sl["fp64"]["rho_l"]
        v["rho_v_rel_err"] = abs(v["rho_v"] - sl["fp64"]["rho_v"]) / sl["fp64"]["rho_v"]
        v["interface_shift"] = v["interface_x"] - sl["fp64"]["interface_x"]
        sl[key] = v
    out["flat_interface"] = sl

    # (3) Darcy permeability: single phase G = 0, tau = 0.8, g = 1e-6 along x
    tau, g = 0.8, 1e-6
    nu = (tau - 0.5) / 3
    perm = {"steps": a.perm_steps, "tau": tau, "g": g}
    ones = np.ones(shape)
    for storage, kw in [("fp64", {})] + variants:
        r, u, st, dt, info, _ = run(storage, dict(tau=tau, G=0.0, rhos=1.0, body_force=(g, 0.0, 0.0)), a.perm_steps,
                                    ones.astype(np.float64 if storage == "fp64" else np.float32), **kw)
        storage = vkey(storage, kw)
        # physical velocity (m + F/2)/rho = shifted u - (tau - 1/2) g; superficial average over all cells
        ux = np.where(fl, u[..., 0] - (tau - 0.5) * g, 0.0)
        k = nu * float(ux.mean()) / g
        perm[storage] = {"k_lattice_units": k, "seconds": round(dt, 2)}
        if storage != "fp64":
            perm[storage]["rel_err"] = abs(k - perm["fp64"]["k_lattice_units"]) / perm["fp64"]["k_lattice_units"]
        print(f"# permeability {storage}: {perm[storage]}", file=sys.stderr, flush=True)
    out["permeability"] = perm
    print(json.dumps(out), flush=True)


def copy_calibration():
    """A plain copy kernel of known size: FETCH_SIZE/WRITE_SIZE of this run
    against the 2 GiB read + 2 GiB written give the counter scale."""
    import torch
    n = 1 << 29  # 2 GiB of float32
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}), flush=True)


def ensemble_bench(a):
    """Batched ensemble: 64 seeded packings, cubic sizes n ~ U{48..160},
    porosity ~ U[0.35, 0.55], radii ~ U[2, 8], G ~ U[-5.5, -4.5] (seeds
    0..63), all advanced by one persistent launch per step (hipGraph replay),
    FP32 and FP16 storage; the autotuner picks the workgroup size and, per
    sample-size class, the tile and segment length.  Prints one JSON line
    with the aggregate MLUPS, the autotuner's picks and per-sample
    saturation."""
    from multiphase_lbm_amd import harness as H
    from multiphase_lbm_amd.ensemble import SIZE_CLASSES, Ensemble, autotune, size_class
    rng = np.random.default_rng(2024)
    samples, fluids, rhos = [], [], []
    t = time.time()
    for k in range(a.ensemble_samples):
        n = int(rng.integers(48, 161))
        por = float(rng.uniform(0.35, 0.55))
        G = float(rng.uniform(-5.5, -4.5))
        fl, _ = H.generate_packing((n, n, n), k, 10 ** 7, 2.0, 8.0, por)
        samples.append(dict(lx=n, ly=n, lz=n, G=G, tau=1.0, rhos=0.7, rhol_spinodal=1.285931,
                            rhog_spinodal=0.323507))
        fluids.append(fl)
        rhos.append(H.rho_init((n, n, n), 0.7, 0.02 / 0.7))
    nf = int(sum(int(f.sum()) for f in fluids))
    cells = int(sum(f.size for f in fluids))
    print(f"# ensemble: {len(samples)} samples, {cells} cells, {nf} fluid nodes, built in {time.time() - t:.1f}s",
          file=sys.stderr, flush=True)
    out = {"workload": f"{len(samples)} packings n^3, n~U{{48..160}}, porosity~U[0.35,0.55], r~U[2,8], "
                       f"G~U[-5.5,-4.5], seeds 0..{len(samples) - 1}",
           "cells": cells, "fluid_nodes": nf}
    for storage in a.storage.split(","):
        if storage in ("fp16sr", "bf16sr", "fp16sn", "bf16sn", "fp16srn", "bf16srn"):
            continue  # the ensemble form stores with round-to-nearest and the global shift only
        def make(block):
            e = Ensemble(samples, storage=BASE_STORAGE.get(storage, storage), block_threads=block, kernel=a.kernel,
                         math="fast" if storage == "fp32fast" else "reference")
            for k in range(len(samples)):
                e.set_geometry(k, fluids[k])
                e.init_fields(k, rhos[k])
            return e
        shapes = [(q["lx"], q["ly"], q["lz"]) for q in samples]
        if a.profile_only:
            # fitted tiles at one workgroup size, timed steps only (for rocprofv3)
            block = a.block_threads or (256 if storage in ("fp16s", "bf16s") else 512)
            with make(block) as ens:
                ens.step(0, a.warmup)
                ms = ens.step_timed(a.warmup, a.steps) / a.steps
            print(f"# ensemble {storage}: block {block}, fitted tiles, {ms:.3f} ms/step, {nf / ms / 1e3:.1f} MLUPS",
                  file=sys.stderr, flush=True)
            continue
        t_tune = time.time()
        best, picks, tuned_ms, trace = autotune(
            make, shapes, log=lambda m: print(f"# autotune {storage}: {m}", file=sys.stderr, flush=True))
        t_tune = time.time() - t_tune
        default_ms = {b: ms for b, p, ms in trace if all(v == (0, 0, 0) for v in p.values())}
        with make(best) as ens:
            for k, shp in enumerate(shapes):
                if picks[size_class(shp)] != (0, 0, 0):
                    ens.set_tiling(k, *picks[size_class(shp)])
            tiles = {}
            for cls in sorted(picks):
                k = next(i for i, shp in enumerate(shapes) if size_class(shp) == cls)
                ty, tz, sg = ens.tiling(k)
                tiles[f"n{SIZE_CLASSES[cls][0]}-{min(SIZE_CLASSES[cls][1], 160)}"] = {
                    "tile_y_z_seg": [ty, tz, sg], "interior_nodes_per_thread": round(ty * tz / best, 3)}
            ens.step(0, a.warmup)
            ms = ens.step_timed(a.warmup, a.steps) / a.steps
            done = a.warmup + a.steps
            rest = max(1, a.ensemble_steps - done)
            ens.step(done, rest, write_macro=True)
            sat = []
            for k in range(len(samples)):
                r, _, _ = ens.read_fields(k)
                fl = fluids[k].astype(bool)
                sat.append(round(float((r[fl] > 0.5 * (1.932442 + 0.156413)).mean()), 4))
            summ = ens.summary()
        out[storage] = {"block_threads": best, "tiles": tiles,
                        "untuned_ms_by_block": {str(b): round(v, 4) for b, v in default_ms.items()},
                        "autotune": {"configurations_timed": len(trace), "seconds": round(t_tune, 1)},
                        "ms_per_step": round(ms, 4), "mlups": round(nf / ms / 1e3, 1),
                        "work_items": summ["work_items"], "saturation_step": done + rest, "saturation": sat}
        print(f"# ensemble {storage}: block {best}, {ms:.3f} ms/step, {nf / ms / 1e3:.1f} MLUPS", file=sys.stderr,
              flush=True)
    if not a.profile_only:
        print(json.dumps(out), flush=True)


def drainage_fields(shape, seed, porosity):
    """Packing (porosity, r ~ U[8, 20]) and the initial density: liquid
    rho = 1.9 in the first lx/16 planes (next to the inlet) and in seeded
    liquid spheres up to 3/4 of the length, vapour 0.15 elsewhere (the last
    quarter, next to the outlet, starts as vapour)."""
    from multiphase_lbm_amd import harness as H
    fluid, _ = H.generate_packing(shape, seed, 10 ** 7, 8.0, 20.0, porosity)
    blobs, _ = H.generate_packing(shape, seed + 1000, 10 ** 7, 6.0, 14.0, 0.85)
    rho = np.full(shape, 0.15, np.float64)
    rho[blobs == 0] = 1.9
    rho[(3 * shape[0]) // 4:] = 0.15
    rho[: max(1, shape[0] // 16)] = 1.9
    return fluid, rho


DRAIN = dict(tau=1.0, G=-5.0, rhos=0.7, rhol_spinodal=1.285931, rhog_spinodal=0.323507,
             face_type=(1, 1, 0, 0, 0, 0), face_rho=(2.0, 0.15, 0, 0, 0, 0))


def drainage_bench(a):
    """Config of the large drainage-style run: 768x512x512, porosity 0.4,
    r ~ U[8, 20], pressure inlet rho = 2.0 at x = 0 and outlet rho = 0.15 at
    x = lx-1, y/z periodic.  (1) FP32 and FP16 storage throughput on the full
    domain; (2) FP64, FP32, FP16 from the same start on a 1/8 sub-volume
    (384x256x256) for `--drainage-steps` steps: saturation and density error
    against FP64 (FP16 with round-to-nearest and with stochastic rounding).
    One JSON line."""
    import multiphase_lbm_amd as mp
    thr = 0.5 * (1.932442 + 0.156413)
    out = {"workload": "768x512x512 packing, porosity 0.4, r~U[8,20], inlet rho 2.0 / outlet rho 0.15 on x, "
                       "liquid 1.9 / vapour 0.15 start"}
    shape = (768, 512, 512)
    t = time.time()
    fluid, rho = drainage_fields(shape, 5, 0.4)
    print(f"# drainage domain built in {time.time() - t:.1f}s, porosity {fluid.mean():.4f}", file=sys.stderr,
          flush=True)
    for storage in a.storage.split(","):
        with mp.Context(*shape, storage=storage, kernel=a.kernel, **DRAIN) as ctx:
            ctx.set_geometry(fluid)
            ctx.init_fields(rho.astype(np.float32))
            nf = ctx.info()["fluid_nodes"]
            ctx.step(0, a.warmup)
            ms = ctx.step_timed(a.warmup, a.steps) / a.steps
            st = ctx.reduce()
        bpu = bytes_per_update(mp, storage)
        out[storage] = {"ms_per_step": round(ms, 4), "mlups": round(nf / ms / 1e3, 1),
                        "roofline_frac": round(nf * bpu / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                        "fluid_nodes": nf, "nonfinite": st["nonfinite"]}
        print(f"# drainage {storage}: {out[storage]}", file=sys.stderr, flush=True)
    del fluid, rho
    sub = (384, 256, 256)
    sfluid, srho = drainage_fields(sub, 6, 0.4)
    sfl = sfluid.astype(bool)
    res = {}
    for storage in ("fp64", "fp32", "fp16s", "fp16sr", "fp16sn", "fp16srn"):
        t = time.time()
        with mp.Context(*sub, storage=storage, kernel=a.kernel, **DRAIN) as ctx:
            ctx.set_geometry(sfluid)
            ctx.init_fields(srho if storage == "fp64" else srho.astype(np.float32))
            done = 0
            while done < a.drainage_steps:
                k = min(2000, a.drainage_steps - done)
                ctx.step(done, k, write_macro=done + k >= a.drainage_steps)
                ctx.sync()
                done += k
                print(f"# sub-volume {storage}: {done} steps", file=sys.stderr, flush=True)
            r, _, _ = ctx.read_fields(np.float64)
            st = ctx.reduce()
        res[storage] = (r, st, time.time() - t)
    ref = res["fp64"][0]
    sv = {"shape": list(sub), "steps": a.drainage_steps, "fp64": {"saturation": float((ref[sfl] > thr).mean())}}
    for storage in ("fp32", "fp16s", "fp16sr", "fp16sn", "fp16srn"):
        r, st, dt = res[storage]
        sv[storage] = {"saturation": float((r[sfl] > thr).mean()),
                       "saturation_abs_err": abs(float((r[sfl] > thr).mean()) - sv["fp64"]["saturation"]),
                       "density_rel_l2": float(np.linalg.norm(r[sfl] - ref[sfl]) / np.linalg.norm(ref[sfl])),
                       "nonfinite": st["nonfinite"], "seconds": round(dt, 1)}
    out["sub_volume"] = sv
    print(json.dumps(out), flush=True)


def cluster_bench(a, shape, fluid, rho, sph):
    """Statistics step of the driver (calc_cluster, src/multiphase_lbm.cu:225-562) on the workload after
    --cluster-steps two-phase steps: device union-find labelling against the reference's way (read rho and p
    to the host, then the host cluster search)."""
    import multiphase_lbm_amd as mp
    from multiphase_lbm_amd import harness as H
    m2 = np.ones(shape, np.uint8)
    sn = np.zeros(shape, np.int64)
    for k, (x, y, z, r) in enumerate(sph):
        H.place_sphere(m2, float(x), float(y), float(z), float(r), sn, k + 1)
    assert np.array_equal(m2, fluid.astype(np.uint8)), "solid numbering does not reproduce the packing"
    with mp.Context(*shape, storage="fp32", kernel=a.kernel, **PHYS) as ctx:
        ctx.set_geometry(fluid)
        ctx.set_solid_numbers(sn, len(sph))
        ctx.init_fields(rho)
        ctx.step(0, a.cluster_steps, write_macro=True)
        ctx.sync()
        print(f"# clusters: {a.cluster_steps} steps done", file=sys.stderr, flush=True)
        ctx.calc_cluster(PHYS["rhol_spinodal"], PHYS["rhog_spinodal"])  # first call allocates
        times = []
        for _ in range(3):
            t = time.perf_counter()
            dev = ctx.calc_cluster(PHYS["rhol_spinodal"], PHYS["rhog_spinodal"])
            times.append(time.perf_counter() - t)
        t = time.perf_counter()
        r, p, _ = ctx.read_fields(np.float32)
        t_read = time.perf_counter() - t
    print(f"# clusters: device {sorted(times)[1] * 1e3:.2f} ms, field read {t_read:.2f} s; host search next",
          file=sys.stderr, flush=True)
    t = time.perf_counter()
    # the reference compares float densities with float thresholds (real rhol_spinodal)
    rhol32, rhog32 = float(np.float32(PHYS["rhol_spinodal"])), float(np.float32(PHYS["rhog_spinodal"]))
    host = H.cluster_stats(fluid, r, p, sn, rhol32, rhog32, len(sph))
    t_host = time.perf_counter() - t
    exact = [k for k in dev if k not in ("avg_liq_press", "avg_vap_press", "avg_liq_rho", "avg_vap_rho")]
    diff = {k: [dev[k], host[k]] for k in exact if dev[k] != host[k]}
    diff.update({k: [dev[k], host[k]] for k in dev
                 if k not in exact and abs(dev[k] - host[k]) > 1e-6 * max(1e-30, abs(host[k]))})
    same = not diff
    dev_ms = sorted(times)[1] * 1e3
    print(json.dumps({"mode": "clusters", "workload": f"{shape[0]}^3 packing after {a.cluster_steps} steps, G=-5",
                      "device_ms": round(dev_ms, 2), "host_read_fields_s": round(t_read, 3),
                      "host_cluster_stats_s": round(t_host, 3),
                      "speedup_vs_host_path": round((t_read + t_host) / (dev_ms / 1e3), 1),
                      "results_equal": bool(same), "differences": diff, "liq_clusters": dev["liq_cluster_count"],
                      "gas_clusters": dev["gas_cluster_count"], "saturation": dev["saturation"],
                      "cluster_order": dev["cluster_order"], "wet_coord_num": dev["wet_coord_num"]}), flush=True)


def voxeliser_bench(shape, fluid, sph):
    """place_sphere (src/multiphase_lbm.cu:893-927) over the packing's sphere list, as read_positions calls it:
    the host restatement sphere by sphere against one device call (host arrays in and out, transfers
    included)."""
    import multiphase_lbm_amd as mp
    from multiphase_lbm_amd import harness as H
    spheres = np.ascontiguousarray(np.asarray(sph, np.float32))
    m0, s0 = np.ones(shape, np.uint8), np.zeros(shape, np.int64)
    mp.place_spheres(m0, spheres[:1], 1, s0)  # first device call initialises the runtime
    times = []
    for _ in range(3):
        md, sd = np.ones(shape, np.uint8), np.zeros(shape, np.int64)
        t = time.perf_counter()
        mp.place_spheres(md, spheres, 1, sd)
        times.append(time.perf_counter() - t)
    mh, sh = np.ones(shape, np.uint8), np.zeros(shape, np.int64)
    t = time.perf_counter()
    for k, (x, y, z, r) in enumerate(spheres):
        H.place_sphere(mh, float(x), float(y), float(z), float(r), sh, k + 1)
    t_host = time.perf_counter() - t
    same = bool(np.array_equal(md, mh) and np.array_equal(sd, sh) and np.array_equal(md, fluid.astype(np.uint8)))
    dev_s = sorted(times)[1]
    print(json.dumps({"mode": "voxeliser", "workload": f"{shape[0]}^3 packing, {len(spheres)} spheres",
                      "device_s": round(dev_s, 4), "host_s": round(t_host, 3),
                      "speedup": round(t_host / dev_s, 1), "results_equal": same}), flush=True)


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU slab decomposition is "
              "a design note, not built", file=sys.stderr)
        sys.exit(2)
    if a.ensemble:
        ensemble_bench(a)
        return
    if a.drainage:
        drainage_bench(a)
        return
    shape, fluid, rho, sph = workload(a)
    nsph = len(sph)
    if a.clusters:
        cluster_bench(a, shape, fluid, rho, sph)
        return
    if a.voxeliser:
        voxeliser_bench(shape, fluid, sph)
        return
    if a.precision:
        precision_study(a, shape, fluid, rho)
        return
    if a.sweep:
        import multiphase_lbm_amd as mp
        from multiphase_lbm_amd.autotune import autotune
        for storage in a.storage.split(","):
            with mp.Context(*shape, storage=storage, kernel="fused", bounce_back=a.bounce_back, **PHYS) as ctx:
                ctx.set_geometry(fluid)
                ctx.init_fields(rho)
                nf = ctx.info()["fluid_nodes"]
                best, res = autotune(ctx, steps=5, verbose=False)
                for (ty, tz, seg), ms in sorted(res, key=lambda r: r[1]):
                    print(json.dumps({"storage": storage, "tile_y": ty, "tile_z": tz, "seg_x": seg,
                                      "ms": round(ms, 4), "mlups": round(nf / ms / 1e3, 1)}), flush=True)
        return
    storages = a.storage.split(",")[:1] if a.profile_only else a.storage.split(",")
    if not a.profile_only and "fp32" not in storages:
        sys.exit("the bench line is quoted on the FP32 step: --storage must include fp32 "
                 "(or use --profile-only to time one storage form)")
    res = {}
    for s in storages:
        res[s] = run_gpu(a, shape, fluid, rho, s)
        print(f"# {s}: {json.dumps(res[s])}", file=sys.stderr, flush=True)
    if a.profile_only:
        return
    errors = error_check(a, shape, fluid, rho, storages) if a.error_steps > 0 else {}
    cpu = cpu_baseline(a, shape, fluid, rho) if a.cpu_steps > 0 else None
    r = res["fp32"]
    line = {
        "metric": "MLUPS + % HBM roofline, 512^3 packing: FP32 vs FP16-storage at stated error",
        "value": round(r["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"D3Q19 Shan-Chen step, {shape[0]}x{shape[1]}x{shape[2]} sphere packing "
                               f"(porosity {fluid.mean():.3f}, {nsph} overlapping spheres r~U[{a.rmin},{a.rmax}], "
                               f"seed {a.seed}), G=-5, tau=1, rho0=0.7+/-0.01, {a.kernel} stream-collide kernel",
                   "fluid_nodes": r["fluid_nodes"], "tile_y_z_seg": r["tile"],
                   "mlups_unit": "fluid-node updates per second / 1e6"},
        "roofline": {"bound": "hbm", "achieved": round(r["achieved_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r["achieved_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles(a, "fp32", r["tag"]),
                     "bytes_per_update": r["bytes_per_update"], "kernel_ms": round(r["kernel_ms"], 4)},
        "cpu_baseline": cpu,
        "error": errors.get("fp32"),
    }
    for s in storages:
        if s == "fp32":
            continue
        q = res[s]
        line[s] = {"value": round(q["mlups"], 1), "unit": "MLUPS", "vs_fp32": round(q["mlups"] / r["mlups"], 3),
                   "error": errors.get(s),
                   "roofline": {"bound": "hbm", "achieved": round(q["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                                "unit": "GB/s", "frac": round(q["achieved_gbs"] / HBM_PEAK_GBS, 4),
                                "traffic": traffic_from_profiles(a, s, q["tag"]), "bytes_per_update": q["bytes_per_update"],
                                "kernel_ms": round(q["kernel_ms"], 4)}}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
