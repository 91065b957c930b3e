"""Run the batched adaptive-grid step on one of the two AMR workloads and
print its time per step:

    python -m peanoclaw_amd.run_amr c1 200     # 2-level dam break, 27 x 27 x 16^2
    python -m peanoclaw_amd.run_amr c3 200     # 3-level wet/dry basin, 9 x 9 x 27^2

c1: coarse patches refined by factor 3 where the BreakingDam criterion
(|grad h| > 0.05) asks for it on the initial data (dam centre jittered with
seed 1); c3: the wet/dry basin refined to 3 levels along the front.
"""
import sys
import time

import numpy as np

from .amr import (AmrGrid, BatchedAmr, DeviceBackend, TwoLevelGrid, dam_break_fields,
                  front_refinement, refined_set_from_flags, wet_dry_fields)
from .pool import PatchPool


def build(pool, case):
    if case == "c1":
        rng = np.random.default_rng(1)
        centre = tuple(0.5 + rng.uniform(-0.05, 0.05, 2))
        fn = dam_break_fields(centre=centre, radius=0.25)
        refined = refined_set_from_flags(pool, 27, 16, fn)
        return TwoLevelGrid(DeviceBackend(pool), 27, 27, 16, refined, fn)
    if case == "c3":
        fn = wet_dry_fields()
        refine = front_refinement(9, 9, 27, fn, 3)
        return AmrGrid(DeviceBackend(pool), 9, 9, 27, refine, fn, max_level=3)
    raise ValueError("case must be c1 or c3")


def main(argv):
    case = argv[1] if len(argv) > 1 else "c1"
    steps = int(argv[2]) if len(argv) > 2 else 200
    with PatchPool(0) as pool:
        grid = build(pool, case)
        amr = BatchedAmr(grid)
        amr.step(2)
        pool.synchronize()
        t0 = time.perf_counter()
        amr.step(steps)
        pool.synchronize()
        el = time.perf_counter() - t0
        cells = grid.cells
        print(f"{case}: {len(grid.leaves)} leaves, {cells} leaf cells, {steps} steps: "
              f"{el / steps * 1e3:.3f} ms/step, {cells * steps / el / 1e9:.3f} G leaf "
              f"cell-updates/s", flush=True)
        amr.close()


if __name__ == "__main__":
    main(sys.argv)
