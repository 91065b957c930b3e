"""Autotuner of the batched step's launch shape.

Knobs (all keep the results bitwise identical, they only change the
schedule): how many same-size patches share one wavefront task (the
segments of the row march), and whether the update kernel runs one task
per workgroup or as a persistent grid of N workgroups pulling tasks from
the device work list (large patches first).  Each candidate is timed over a
few steps on the real pool; the fastest is applied and returned.
"""
import itertools
import time


def candidates(num_cus=256):
    persistent = [0, 8 * num_cus, 16 * num_cus, 20 * num_cus]
    caps = [0, 2, 1]
    return list(itertools.product(persistent, caps))


def tune(pool, steps=3, num_cus=256, verbose=False):
    """Time each (persistent_workgroups, patches_per_task_cap) candidate and
    keep the fastest.  Advances the pool by (#candidates + 1) * steps steps."""
    best = None
    results = []
    for pw, cap in candidates(num_cus):
        pool.launch_config(persistent_workgroups=pw, patches_per_task_cap=cap)
        pool.step(1)  # rebuild the work list, warm caches
        pool.step(steps)
        total_ms, upd_ms = pool.last_timing()
        results.append({"persistent_workgroups": pw, "patches_per_task_cap": cap,
                        "ms_per_step": total_ms / steps, "update_ms": upd_ms})
        if verbose:
            print(f"autotune: persistent={pw:5d} cap={cap} -> {total_ms / steps:.3f} ms/step "
                  f"(update {upd_ms:.3f} ms)", flush=True)
        if best is None or total_ms < best[0]:
            best = (total_ms, pw, cap)
    pool.launch_config(persistent_workgroups=best[1], patches_per_task_cap=best[2])
    return {"persistent_workgroups": best[1], "patches_per_task_cap": best[2],
            "candidates": results}


__all__ = ["tune", "candidates"]
