"""Autotuner of the batched step's launch shape.

Knobs (all keep the results bitwise identical, they only change the
schedule): whether the next step's CFL speeds are computed inside the
update kernel (exactly for every edge, or pruned by a cheap bound with an
exact fallback) or by a separate speed pass, how many same-size patches
share one wavefront task (the segments of the row march), and whether the
update kernel runs one task per workgroup or as a persistent grid of N
workgroups pulling tasks from the device work list (large patches first).
A second stage tunes, for the chosen schedule, the pruning slack (how close
to the last CFL limit an edge's bound must come for its exact speed to be
computed; more slack computes more speeds, less makes the exact fallback
pass more likely) and the splitting of patches into row segments (automatic,
off, or forced to 2 or 4).  Each candidate is timed over a few steps on the
real pool; the fastest is applied and returned.
"""
import itertools
import sys


def candidates(num_cus=256):
    fused = [2, 1, 0]
    persistent = [0, 8 * num_cus, 16 * num_cus]
    caps = [0, 2]
    return list(itertools.product(fused, persistent, caps))


def tune(pool, steps=3, num_cus=256, verbose=False, rounds=2):
    """Time each (fused, persistent_workgroups, patches_per_task_cap)
    candidate and keep the fastest.  The first candidate is run once untimed
    (clocks and caches settle), then every candidate is timed in `rounds`
    passes over the list and scored by its best pass, so no candidate is
    favoured by its place in the order.  Advances the pool by
    steps * (2 + 2 * rounds * len(candidates)) steps."""
    cands = candidates(num_cus)
    f0, pw0, cap0 = cands[0]
    pool.launch_config(persistent_workgroups=pw0, fused=f0, patches_per_task_cap=cap0)
    pool.step(2 * steps)
    best_of = {}
    for _ in range(rounds):
        for fused, pw, cap in cands:
            pool.launch_config(persistent_workgroups=pw, fused=fused, patches_per_task_cap=cap)
            pool.step(steps)  # rebuild the work list, warm caches
            pool.step(steps)
            total_ms, upd_ms = pool.last_timing()
            key = (fused, pw, cap)
            prev = best_of.get(key)
            if prev is None or total_ms < prev[0]:
                best_of[key] = (total_ms, upd_ms)
    results = []
    best = None
    for (fused, pw, cap) in cands:
        total_ms, upd_ms = best_of[(fused, pw, cap)]
        results.append({"fused": fused, "persistent_workgroups": pw,
                        "patches_per_task_cap": cap, "ms_per_step": total_ms / steps,
                        "update_ms": upd_ms})
        if verbose:
            print(f"autotune: fused={fused} persistent={pw:5d} cap={cap} -> "
                  f"{total_ms / steps:.3f} ms/step (update {upd_ms:.3f} ms)", file=sys.stderr,
                  flush=True)
        if best is None or total_ms < best[0]:
            best = (total_ms, fused, pw, cap)
    pool.launch_config(persistent_workgroups=best[2], fused=best[1], patches_per_task_cap=best[3])
    # second stage, on the chosen schedule: slack (pruned mode only), then
    # row segments
    slack, segments = 1.03125, 0
    stage2 = []

    def time_config():
        pool.step(steps)
        pool.step(steps)
        return pool.last_timing()

    def pick(options, apply, name):
        nonlocal best
        choice, best_ms = None, None
        for _ in range(rounds):
            for v in options:
                apply(v)
                total_ms, upd_ms = time_config()
                stage2.append({name: v, "ms_per_step": total_ms / steps, "update_ms": upd_ms})
                if verbose:
                    print(f"autotune: {name}={v} -> {total_ms / steps:.3f} ms/step "
                          f"(update {upd_ms:.3f} ms)", file=sys.stderr, flush=True)
                if best_ms is None or total_ms < best_ms:
                    choice, best_ms = v, total_ms
        apply(choice)
        return choice

    if best[1] == 2:
        slack = pick([1.0078125, 1.03125, 1.125], pool.prune_slack, "prune_slack")
    segments = pick([0, 1, 2, 4], pool.row_segments, "row_segments")
    return {"fused": best[1], "persistent_workgroups": best[2], "patches_per_task_cap": best[3],
            "prune_slack": slack, "row_segments": segments, "candidates": results,
            "stage2": stage2}


__all__ = ["tune", "candidates"]
