#!/usr/bin/env python3
"""Benchmark of the batched FP64 SWE patch step on one MI355X.

Workload: 50,000 patches with sizes drawn uniformly from {8,16,32,64}^2
(seeded), h in [0.5, 2] with Gaussian bumps, small random momenta,
Gaussian-hill bathymetry; each step = ghost fill + CFL reduction + update
of every patch with one global dt.  One "step" advances all 68M cells once.

Prints one JSON line: cell-updates/s of the batched path, its HBM roofline
fraction (update kernel, HIP events around every launch of the timed
steps), the plain per-patch-launch path on the same pool, the hipGraph
replay rate, p50/p99 time-to-solution of a served stream of independent
dam-break requests (Poisson arrivals, continuous batching under an HBM
budget), and the CPU oracle timed on the host cores on a sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 cell-updates/s + %HBM roofline, batched vs per-patch; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE = 56  # read h, hu, hv, b; write h, hu, hv (FP64)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_update_uneven50k_pmc.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic():
    """HBM bytes per update launch from the committed PMC summary, and where
    they come from.  The summary records the source hash of the library it
    was measured on (peanoclaw_amd.build.source_hash); a summary of another
    build gives no traffic figure."""
    from peanoclaw_amd.build import source_hash
    vals = {}
    try:
        for line in open(TRAFFIC_FILE):
            for key in ("traffic_bytes_per_launch:", "library_source_sha256:"):
                if line.startswith(key):
                    vals[key] = line.split(":", 1)[1].split()[0]
    except OSError:
        return None, "no PMC summary"
    rel = os.path.relpath(TRAFFIC_FILE, ROOT)
    if "traffic_bytes_per_launch:" not in vals:
        return None, f"{rel} holds no traffic figure"
    if vals.get("library_source_sha256:") != source_hash():
        return None, f"{rel} was measured on another build of the library"
    return (float(vals["traffic_bytes_per_launch:"]),
            f"{rel}: rocprofv3 --pmc FETCH_SIZE x2.0 + WRITE_SIZE x1.0 per launch (scale "
            f"factors from k_copy_probe), library source sha256 {source_hash()[:16]}")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(num_patches, seconds, threads, runs=5):
    """The oracle (FP64, C, OpenMP over patches) on a sample of the same
    workload generator: ghost fill of every size-class grid + global-dt
    step of every patch.  One warm-up step, then `runs` timed runs of an
    equal number of steps sized to `seconds` in total; the median rate."""
    import ctypes

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OrPool, dptr, lib
    from peanoclaw_amd.scenarios import SIZES, uneven_class_counts, uneven_fields

    L = lib()
    counts = uneven_class_counts(num_patches, 2)
    rng = np.random.default_rng(3)
    dx = 1.0 / 64
    layouts, offs, nxs = [], [], []
    chunks = {k: [] for k in ("h", "hu", "hv", "b")}
    top = 0
    slot = 0
    for n in sorted(SIZES, reverse=True):
        c = counts[n]
        if not c:
            continue
        width = int(math.ceil(math.sqrt(c)))
        h, hu, hv, b, rows = uneven_fields(n, c, width, rng, dx)
        slot_of = -np.ones((width, rows), dtype=np.int32)
        for k in range(c):
            slot_of[k // rows, k % rows] = slot + k
            offs.append(top + k * (n + 2) ** 2)
            nxs.append(n)
        for name, arr in (("h", h), ("hu", hu), ("hv", hv), ("b", b)):
            chunks[name].append(arr.ravel())
        layouts.append((n, np.ascontiguousarray(slot_of)))
        top += c * (n + 2) ** 2
        slot += c
    H = [np.concatenate(chunks["h"]), np.zeros(top)]
    HU = [np.concatenate(chunks["hu"]), np.zeros(top)]
    HV = [np.concatenate(chunks["hv"]), np.zeros(top)]
    B = np.concatenate(chunks["b"])
    off = np.array(offs, dtype=np.int64)
    nx = np.array(nxs, dtype=np.int32)
    dxs = np.full(len(nxs), dx)
    pl = OrPool()
    pl.np = len(nxs)
    pl.nx = nx.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    pl.ny = pl.nx
    pl.off = off.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))
    pl.dx = dxs.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    pl.dy = pl.dx
    for k in range(2):
        pl.h[k], pl.hu[k], pl.hv[k] = dptr(H[k]), dptr(HU[k]), dptr(HV[k])
    pl.b = dptr(B)
    pl.cur = 0
    cells = int(sum(n * n * counts[n] for n in SIZES))

    def step():
        c = pl.cur
        for n, so in layouts:
            L.or_grid_fill(so.shape[0], so.shape[1], n,
                           so.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                           off.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                           dptr(H[c]), dptr(HU[c]), dptr(HV[c]))
        L.or_pool_step(ctypes.byref(pl), math.inf, threads)

    t0 = time.perf_counter()
    step()  # warm-up (also sizes the runs)
    t1 = time.perf_counter() - t0
    per_run = max(1, int(round(seconds / runs / max(t1, 1e-6))))
    rates = []
    for _ in range(runs):
        t0 = time.perf_counter()
        for _ in range(per_run):
            step()
        rates.append(cells * per_run / (time.perf_counter() - t0))
    med = float(np.median(rates))
    return {"value": med, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "runs": [round(r) for r in rates],
            "sample": f"{num_patches} patches of the same generator ({cells} cells); median "
                      f"of {runs} runs of {per_run} global-dt steps each (ghost fill + step) "
                      f"after one warm-up step; FP64 oracle in C (gcc -O3 -march=x86-64-v3 "
                      f"-ffp-contract=off, OpenMP over patches)"}


def serve_stream(args, tuned):
    """Served stream: independent dam-break requests (grid log-uniform in
    64^2..2048^2 cells, 32^2-cell patches, t_end = 1) arriving as a Poisson
    stream, continuously batched on one pool with an HBM budget."""
    import numpy as np

    from peanoclaw_amd.pool import PatchPool
    from peanoclaw_amd.serving import Server, run_stream, served_requests

    reqs = served_requests(args.serve_requests, seed=5)
    with PatchPool(0, hbm_budget_bytes=args.serve_budget_mib << 20) as pool:
        if tuned:
            pool.launch_config(persistent_workgroups=tuned["persistent_workgroups"],
                               fused=tuned["fused"],
                               patches_per_task_cap=tuned["patches_per_task_cap"])
        with Server(pool, max_active=64) as srv:
            res, lat, wall, cells = run_stream(srv, reqs, rate=args.serve_rate, seed=5,
                                               steps_per_round=args.serve_steps_per_round)
    sizes = [r.cells_per_side for r in reqs]
    out = {"requests": len(reqs), "arrival_rate_per_s": args.serve_rate,
           "grid_cells_per_side_min": min(sizes), "grid_cells_per_side_max": max(sizes),
           "hbm_budget_mib": args.serve_budget_mib, "max_active": 64,
           "steps_per_round": args.serve_steps_per_round,
           "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
           "mean_s": float(lat.mean()), "max_s": float(lat.max()), "wall_s": wall,
           "cell_updates": int(cells), "throughput_cell_updates_per_s": cells / wall,
           "total_steps": int(sum(r.steps for r in res.values()))}
    log(f"served stream: {len(reqs)} requests, p50 {out['p50_s']:.3f} s, "
        f"p99 {out['p99_s']:.3f} s, {out['throughput_cell_updates_per_s'] / 1e9:.2f} G/s "
        f"over {wall:.1f} s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--patches", type=int, default=50000)
    ap.add_argument("--per-patch-steps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-patches", type=int, default=4000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--config", default="",
                    help="fixed launch config 'fused,persistent_workgroups,cap' (skips autotune)")
    ap.add_argument("--copy-calibration", type=float, default=1.0,
                    help="GiB copied by the plain copy-kernel probe (HBM rate, counter scale)")
    ap.add_argument("--copy-reps", type=int, default=5)
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="requests of the served stream (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-budget-mib", type=int, default=512)
    ap.add_argument("--serve-steps-per-round", type=int, default=64)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: it measures the patch path on a single "
              "MI355X; sharding patches over several GPUs is not part of this benchmark.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from peanoclaw_amd import autotune
    from peanoclaw_amd.pool import PatchPool
    from peanoclaw_amd.scenarios import uneven_patches

    t_setup = time.perf_counter()
    pool = PatchPool(0)
    wl = uneven_patches(pool, num_patches=args.patches)
    cells = wl.cells_per_step
    log(f"setup: {args.patches} patches, {cells} cells/step, "
        f"{pool.bytes_in_use() / 2**30:.2f} GiB pool, {time.perf_counter() - t_setup:.1f} s")
    mass0 = pool.mass()

    tuned = None
    if args.config:
        f, pw, cap = (int(v) for v in args.config.split(","))
        pool.launch_config(persistent_workgroups=pw, fused=f, patches_per_task_cap=cap)
        tuned = {"fused": f, "persistent_workgroups": pw, "patches_per_task_cap": cap}
    elif not args.no_autotune:
        tuned = autotune.tune(pool, steps=3, verbose=True)
        log("autotune picked", {k: v for k, v in tuned.items() if k != "candidates"})

    pool.step(args.warmup)
    pool.synchronize()
    t0 = time.perf_counter()
    pool.step(args.steps)
    pool.synchronize()
    wall = time.perf_counter() - t0
    ev_total_ms, upd_ms = pool.last_timing()
    ms_per_step = 1e3 * wall / args.steps
    value = cells * args.steps / wall
    log(f"batched: {ms_per_step:.3f} ms/step (events {ev_total_ms / args.steps:.3f}), "
        f"update kernel {upd_ms:.3f} ms, {value / 1e9:.2f} G cell-updates/s")

    # hipGraph replay of {ghost fill, CFL reduction, dt, update}
    pool.graph_capture()
    pool.graph_replay(2)
    pool.synchronize()
    t0 = time.perf_counter()
    pool.graph_replay(args.steps)
    pool.synchronize()
    gwall = time.perf_counter() - t0
    graph_value = cells * args.steps / gwall
    log(f"graph replay: {1e3 * gwall / args.steps:.3f} ms/step, {graph_value / 1e9:.2f} G/s")

    # plain per-patch-launch form on the same pool (same results bitwise)
    per_patch_value = None
    if args.per_patch_steps > 0:
        pool.step_per_patch(1)
        pool.synchronize()
        t0 = time.perf_counter()
        pool.step_per_patch(args.per_patch_steps)
        pool.synchronize()
        pwall = time.perf_counter() - t0
        per_patch_value = cells * args.per_patch_steps / pwall
        log(f"per-patch launches: {1e3 * pwall / args.per_patch_steps:.1f} ms/step, "
            f"{per_patch_value / 1e9:.3f} G/s")

    # a plain copy kernel of known size (k_copy_probe): the achievable HBM
    # rate, and the byte scale of FETCH_SIZE / WRITE_SIZE in a profiled run
    copy_bytes = int(args.copy_calibration * 2**30) // 16 * 16
    copy_gbs = pool.copy_probe(copy_bytes, reps=args.copy_reps)
    log(f"copy probe: {copy_bytes} bytes read + {copy_bytes} written per launch, "
        f"best {copy_gbs:.0f} GB/s")
    # the same traffic mix as the update (4 doubles read, 3 written per
    # cell) as one flat stream without neighbours: the achievable rate of
    # the update's bytes on this GPU
    stream_gbs = pool.stream_probe(cells, reps=args.copy_reps)
    log(f"4-read/3-write stream probe over {cells} cells: best {stream_gbs:.0f} GB/s "
        f"(56 B per cell)")
    # the update's own traversal of the pool (tasks, row march, prefetch)
    # without the edge arithmetic: the memory-pattern ceiling of the layout
    march_gbs = pool.march_probe(reps=args.copy_reps)
    log(f"row-march probe (update traversal, no arithmetic): best {march_gbs:.0f} GB/s "
        f"(56 B per cell)")

    mass1 = pool.mass()
    t, dt, est, nsteps = pool.scenario_get(0)

    achieved = cells * BYTES_PER_CELL_UPDATE / (upd_ms * 1e-3) / 1e9
    traffic, traffic_source = read_traffic()
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "traffic_source": traffic_source,
                "kernel": "k_update (row-march f-wave update, all patches)",
                "algorithmic_bytes_per_launch": cells * BYTES_PER_CELL_UPDATE,
                "avg_launch_ms": round(upd_ms, 4),
                "copy_kernel_gbs": round(copy_gbs, 1),
                "frac_of_copy_kernel": round(achieved / copy_gbs, 4),
                "stream_4r3w_gbs": round(stream_gbs, 1),
                "frac_of_stream_4r3w": round(achieved / stream_gbs, 4),
                "march_probe_gbs": round(march_gbs, 1),
                "frac_of_march_probe": round(achieved / march_gbs, 4)}

    serving = None
    if args.serve_requests > 0:
        serving = serve_stream(args, tuned)

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(args.cpu_patches, args.cpu_seconds, threads)
        # the same sample on one core (a shorter run: the rate is steady)
        one = cpu_baseline(args.cpu_patches, max(args.cpu_seconds / 2, 1.0), 1)
        cpu["single_core_value"] = one["value"]
        cpu["single_core_sample"] = one["sample"]
        cpu["host_cpus"] = os.cpu_count()
        cpu["cpu_model"] = cpu_model()
        log(f"cpu baseline: {cpu['value'] / 1e6:.1f} M cell-updates/s on {threads} threads, "
            f"{one['value'] / 1e6:.1f} M on one core ({cpu['cpu_model']})")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "50k uneven SWE patches {8,16,32,64}^2 (seeded), FP64 f-wave, "
                               "global dt; ghost fill + CFL reduction + update per step",
                   "patches": args.patches, "cells_per_step": cells,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "hbm_frac_whole_step": round(value * BYTES_PER_CELL_UPDATE / (HBM_PEAK_GBS * 1e9), 4),
        "graph_replay_value": graph_value,
        "per_patch_launch_value": per_patch_value,
        "batched_vs_per_patch": value / per_patch_value if per_patch_value else None,
        "autotune": None if tuned is None else {k: v for k, v in tuned.items()
                                                if k not in ("candidates", "stage2")},
        "mass_rel_drift": abs(mass1 - mass0) / mass0,
        "p50_solve_latency_s": serving["p50_s"] if serving else None,
        "p99_solve_latency_s": serving["p99_s"] if serving else None,
        "serving": serving,
        "sim_time": t, "steps_done": nsteps,
    }
    print(json.dumps(out), flush=True)
    pool.close()


if __name__ == "__main__":
    main()
