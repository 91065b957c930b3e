"""Workloads of the SWE patch path, built directly into a PatchPool.

* ``dam_break_grid``   — uniform px x py grid of n x n patches on [0,1]^2 with
  the radial BreakingDam initial state (native/scenarios/BreakingDam.cpp:
  53-83, 165-176), flat bathymetry and reflective walls; the 9x9 x 16^2
  case is the geometry of testscenarios/test9x9PatchesOf16x16Cells.py.
* ``uneven_patches``   — N patches whose sizes are drawn uniformly from
  {8,16,32,64}^2; each size class is tiled row-major into a ceil(sqrt(N_c))
  wide grid of same-size patches with dx = dy = 1/64 (holes and the border
  are reflective walls), h in [0.5, 2] from uniform noise plus 3 Gaussian
  bumps, hu, hv ~ U[-0.1, 0.1], b = sum of 5 Gaussian hills (height <= 0.3)
  evaluated at every cell centre including ghosts.

All generators are seeded numpy PCG64 streams, so the same arguments give
the same bits.
"""
import math

import numpy as np

from ._abi import GHOST_STRIP_DTYPE
from .pool import make_description, tile_stride

SIZES = (8, 16, 32, 64)


class GridLayout:
    """Slots of one equal-size patch grid: slot_of[i, j] (-1 = hole); its
    patches have n x ny cells (ny = n unless given)."""

    def __init__(self, n, slot_of, ny=None):
        self.n = n
        self.ny = n if ny is None else ny
        self.slot_of = slot_of


class Workload:
    def __init__(self, pool, layouts, scenario, cells_per_step, slots):
        self.pool = pool
        self.layouts = layouts
        self.scenario = scenario
        self.cells_per_step = cells_per_step
        self.slots = slots

    def build_ghost_table(self, fused_sides=True):
        """Device neighbour table of the grid(s): fused ghost sources (the
        fast path), or explicit ghost strips gathered into the tiles."""
        if fused_sides:
            self.pool.set_ghost_strips(np.zeros(0, dtype=GHOST_STRIP_DTYPE))
            for L in self.layouts:
                self.pool.set_grid_sides(L.slot_of)
            return None
        self.pool.clear_sides()
        tables = [self.pool.grid_strips(L.slot_of) for L in self.layouts]
        strips = np.concatenate(tables) if tables else np.zeros(0, dtype=GHOST_STRIP_DTYPE)
        self.pool.set_ghost_strips(strips)
        return strips


def _register_grid(pool, px, py, n, dx, scenario, origin=(0.0, 0.0), count=None):
    count = px * py if count is None else count
    slot_of = -np.ones((px, py), dtype=np.int32)
    k = 0
    for i in range(px):
        for j in range(py):
            if k >= count:
                break
            d = make_description((origin[0] + i * n * dx, origin[1] + j * n * dx), n * dx, n,
                                 scenario=scenario)
            slot_of[i, j] = pool.register(d)
            k += 1
    return slot_of


def dam_break_grid(pool, px=9, py=9, n=16, scenario=0, centre=(0.5, 0.5), radius=0.25,
                   h_in=2.0, h_out=1.0):
    """Config: [0,1]^2, px x py patches of n^2 cells, radial dam break."""
    if px != py:
        raise ValueError("the dam break grid is square")
    dx = 1.0 / (px * n)
    slot_of = _register_grid(pool, px, py, n, dx, scenario)
    pool.scenario_set(scenario, 0.0, 0.0, np.inf)
    for i in range(px):
        for j in range(py):
            s = int(slot_of[i, j])
            d = pool.description(s)
            sub = d.size[0] / n
            X = d.position[0] + (np.arange(n)[:, None] + 0.5) * sub
            Y = d.position[1] + (np.arange(n)[None, :] + 0.5) * sub
            r = np.sqrt((X - centre[0]) * (X - centre[0]) + (Y - centre[1]) * (Y - centre[1]))
            unew = np.zeros((3, n, n))
            unew[0] = h_in * (r <= radius) + h_out * (r > radius)
            pool.upload(s, unew=unew, b=np.zeros((n + 2, n + 2)))
    wl = Workload(pool, [GridLayout(n, slot_of)], scenario, px * py * n * n,
                  [int(s) for s in slot_of.ravel()])
    wl.build_ghost_table()
    return wl


def wet_dry_basin(X, Y, rng_params):
    """Bathymetry of the wet/dry basin: a slope from -1 (x = 0) to +0.5
    (x = 1) plus seeded Gaussian bumps of height <= 0.15."""
    b = -1.0 + 1.5 * X
    for cx, cy, s, a in rng_params:
        b = b + a * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * s * s))
    return b


def wet_dry_grid(pool, px=9, py=9, n=16, seed=4, scenario=0):
    """Wet/dry stress case on [0,1]^2: sloping basin (b from -1 to +0.5,
    seeded bumps), still water at surface level 0 where b < 0, dry land
    (h = 0) elsewhere, and a dam-break column of depth 1.5 over 10% of the
    domain (x in [0.05, 0.35], y in [0.3, 0.6333]); reflective walls."""
    rng = np.random.default_rng(seed)
    params = [(rng.uniform(0.2, 0.9), rng.uniform(0.1, 0.9), rng.uniform(0.03, 0.1),
               rng.uniform(-0.15, 0.15)) for _ in range(6)]
    dx = 1.0 / (px * n)
    slot_of = _register_grid(pool, px, py, n, dx, scenario)
    pool.scenario_set(scenario, 0.0, 0.0, np.inf)
    for i in range(px):
        for j in range(py):
            s = int(slot_of[i, j])
            d = pool.description(s)
            sub = d.size[0] / n
            ii = np.arange(-1, n + 1)
            X = d.position[0] + (ii[:, None] + 0.5) * sub
            Y = d.position[1] + (ii[None, :] + 0.5) * sub
            X, Y = np.broadcast_arrays(X, Y)
            b = wet_dry_basin(X, Y, params)
            h = np.maximum(0.0, 0.0 - b)
            col = (X >= 0.05) & (X <= 0.35) & (Y >= 0.3) & (Y <= 0.3 + 0.1 / 0.3)
            h = np.where(col, 1.5, h)
            unew = np.zeros((3, n, n))
            unew[0] = h[1:-1, 1:-1]
            pool.upload(s, unew=unew, b=np.ascontiguousarray(b))
    wl = Workload(pool, [GridLayout(n, slot_of)], scenario, px * py * n * n,
                  [int(s) for s in slot_of.ravel()])
    wl.build_ghost_table()
    return wl


def uneven_class_counts(num_patches, seed=2):
    rng = np.random.default_rng(seed)
    sizes = np.asarray(SIZES)[rng.integers(0, len(SIZES), num_patches)]
    return {n: int((sizes == n).sum()) for n in SIZES}


def uneven_fields(n, count, width, rng, dx):
    """Fields of one size class tiled into a `width`-wide grid: arrays of
    shape (count, n+2, n+2) including the ghost ring."""
    L = width * n * dx
    rows = -(-count // width)
    Ly = rows * n * dx
    # Gaussian bumps and hills of this class domain
    bc = rng.uniform(0, 1, (3, 2)) * (L, Ly)
    bs = rng.uniform(0.05, 0.2, 3) * max(L, Ly)
    ba = rng.uniform(0.2, 0.5, 3)
    hc = rng.uniform(0, 1, (5, 2)) * (L, Ly)
    hs = rng.uniform(0.05, 0.15, 5) * max(L, Ly)
    ha = rng.uniform(0.0, 0.3, 5)
    k = np.arange(count)
    pi, pj = k // rows, k % rows
    ii = np.arange(-1, n + 1)
    X = (pi[:, None, None] * n + ii[None, :, None] + 0.5) * dx
    Y = (pj[:, None, None] * n + ii[None, None, :] + 0.5) * dx
    X, Y = np.broadcast_arrays(X, Y)
    bump = np.zeros(X.shape)
    for c, s, a in zip(bc, bs, ba):
        bump += a * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2) / (2 * s * s))
    noise = rng.uniform(0.0, 1.0, X.shape)
    h = np.clip(0.8 + 0.2 * noise + bump, 0.5, 2.0)
    hu = rng.uniform(-0.1, 0.1, X.shape)
    hv = rng.uniform(-0.1, 0.1, X.shape)
    b = np.zeros(X.shape)
    for c, s, a in zip(hc, hs, ha):
        b += a * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2) / (2 * s * s))
    return h, hu, hv, b, rows


def uneven_patches(pool, num_patches=50000, seed=2, field_seed=3, scenario=0, dx=1.0 / 64):
    """Config: N uneven patches (see module docstring), one scenario."""
    counts = uneven_class_counts(num_patches, seed)
    rng = np.random.default_rng(field_seed)
    layouts, slots = [], []
    cells = 0
    pool.reserve(sum(c * ((n + 2) * tile_stride(n) + 16) for n, c in counts.items()))
    pool.scenario_set(scenario, 0.0, 0.0, np.inf)
    state = pool.current_state(scenario)
    # large classes first: their tiles come first in the pool
    for n in sorted(SIZES, reverse=True):
        c = counts[n]
        if c == 0:
            continue
        width = int(math.ceil(math.sqrt(c)))
        h, hu, hv, b, rows = uneven_fields(n, c, width, rng, dx)
        slot_of = _register_grid(pool, width, rows, n, dx, scenario, count=c)
        # ghost-ring values of h, hu, hv are rewritten by the ghost fill
        order = [int(slot_of[i // rows, i % rows]) for i in range(c)]
        offs = np.array([pool.offset(s) for s in order])
        ts = tile_stride(n)

        def rows(a):
            # tiles with rows padded to the pool's stride (padding zero)
            out = np.zeros(a.shape[:-1] + (ts,))
            out[..., :n + 2] = a
            return out

        if np.all(np.diff(offs) == (n + 2) * ts):
            pool.copy_fields(True, state, int(offs[0]), rows(h).ravel(), rows(hu).ravel(),
                             rows(hv).ravel(), rows(b).ravel())
        else:
            for t, s in enumerate(order):
                pool.copy_fields(True, state, int(offs[t]), rows(h[t]).ravel(),
                                 rows(hu[t]).ravel(), rows(hv[t]).ravel(), rows(b[t]).ravel())
        layouts.append(GridLayout(n, slot_of))
        slots.extend(order)
        cells += c * n * n
    wl = Workload(pool, layouts, scenario, cells, slots)
    wl.build_ghost_table()
    return wl


__all__ = ["dam_break_grid", "wet_dry_grid", "uneven_patches", "uneven_class_counts",
           "uneven_fields",
           "Workload", "GridLayout", "SIZES"]
