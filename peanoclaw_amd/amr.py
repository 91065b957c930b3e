"""Two-level adaptive patch grid (refinement factor 3) with global time steps.

The grid is px x py coarse patches of n x n cells on [0,1]^2; a refined
coarse patch becomes virtual and holds 3 x 3 fine patches of n x n cells.
One step is the sequence of Numerics calls PeanoClaw's traversal makes for
a grid whose leaves all share one clock (maximumTimestepSize = the global
CFL minimum, SolveTimestep.cpp:591-751):

1. every virtual patch is cleared and receives the restriction of its nine
   children, then the SWE post-processing (GridLevelTransfer ->
   Numerics::restrictSolution / postProcessRestriction, SWEKernel.cpp:288-316);
2. every leaf copies uNew to uOld (SolveTimestep.cpp:638);
3. every leaf face gets its ghost layer from the patch across it: a
   same-level leaf or virtual patch by transferGhostlayer
   (GhostlayerCompositorFunctors.cpp:59), a coarse leaf next to a fine patch
   by interpolateSolution(toUOld, not toCurrentTime, useTimeUNewOrUOld)
   (GhostlayerCompositorFunctors.cpp:61-69);
4. domain walls by fillBoundaryLayer in the order left, top, right, bottom
   (SolveTimestep.cpp:87-142);
5. the global step = min over leaves of their CFL step (bounded by the end
   time) and solveTimestep on every leaf with it.

Corner ghost cells are not filled from diagonal neighbours: the update never
reads them.  The driver talks to a backend object; `DeviceBackend` runs
every call on the GPU through the C ABI (each call one launch, the
reference's per-patch call pattern).
"""
import ctypes

import numpy as np

from ._abi import check, iarr, load
from .pool import make_description

LA_EPS = 1e-8
FACES = ((0, 0), (0, 1), (1, 0), (1, 1))  # (dimension, upper): x-, x+, y-, y+
WALL_ORDER = ((0, 0), (1, 1), (0, 1), (1, 0))  # left, top, right, bottom


class AmrPatch:
    def __init__(self, handle, level, pos, size, n, virtual, ij):
        self.handle = handle
        self.level = level
        self.pos = pos
        self.size = size
        self.n = n
        self.virtual = virtual
        self.ij = ij  # index on its level
        self.children = []


class TwoLevelGrid:
    def __init__(self, backend, px, py, n, refined, fn):
        """fn(X, Y) -> (h, b) at cell centres (arrays of equal shape)."""
        self.b = backend
        self.px, self.py, self.n = px, py, n
        self.refined = set(refined)
        size_c = 1.0 / px
        self.coarse = {}
        self.fine = {}
        for i in range(px):
            for j in range(py):
                pos = (i * size_c, j * size_c)
                virt = (i, j) in self.refined
                cp = AmrPatch(backend.new_patch(pos, size_c, n, 1, virt), 1, pos, size_c, n,
                              virt, (i, j))
                self.coarse[(i, j)] = cp
                if virt:
                    sf = size_c / 3.0
                    for a in range(3):
                        for c in range(3):
                            fpos = (pos[0] + a * sf, pos[1] + c * sf)
                            fp = AmrPatch(backend.new_patch(fpos, sf, n, 2, False), 2, fpos, sf,
                                          n, False, (3 * i + a, 3 * j + c))
                            self.fine[fp.ij] = fp
                            cp.children.append(fp)
        self.virtuals = [p for p in self.coarse.values() if p.virtual]
        self.leaves = [p for p in self.coarse.values() if not p.virtual] + list(self.fine.values())
        for p in list(self.coarse.values()) + list(self.fine.values()):
            ii = np.arange(-1, n + 1)
            sub = p.size / n
            X = p.pos[0] + (ii[:, None] + 0.5) * sub
            Y = p.pos[1] + (ii[None, :] + 0.5) * sub
            X, Y = np.broadcast_arrays(X, Y)
            h, bb = fn(X, Y)
            backend.set_state(p.handle, np.ascontiguousarray(h[1:-1, 1:-1], dtype=np.float64),
                              np.ascontiguousarray(bb, dtype=np.float64))
        self.time, self.dt = 0.0, 0.0
        self.steps = 0
        self._plan = self._ghost_plan()

    @property
    def cells(self):
        return len(self.leaves) * self.n * self.n

    def _face(self, p, dim, upper):
        """('transfer'|'interp'|'wall', source patch) for a face of leaf p."""
        if p.level == 1:
            i, j = p.ij
            q = (i + (1 if upper else -1), j) if dim == 0 else (i, j + (1 if upper else -1))
            if not (0 <= q[0] < self.px and 0 <= q[1] < self.py):
                return ("wall", None)
            return ("transfer", self.coarse[q])
        I, J = p.ij
        q = (I + (1 if upper else -1), J) if dim == 0 else (I, J + (1 if upper else -1))
        if not (0 <= q[0] < 3 * self.px and 0 <= q[1] < 3 * self.py):
            return ("wall", None)
        if q in self.fine:
            return ("transfer", self.fine[q])
        return ("interp", self.coarse[(q[0] // 3, q[1] // 3)])

    def _ghost_plan(self):
        n = self.n
        plan = []
        for p in self.leaves:
            for dim, upper in FACES:
                kind, src = self._face(p, dim, upper)
                if kind == "wall":
                    continue
                size = (1, n) if dim == 0 else (n, 1)
                if dim == 0:
                    doff = (n if upper else -1, 0)
                    soff = (0 if upper else n - 1, 0)
                else:
                    doff = (0, n if upper else -1)
                    soff = (0, 0 if upper else n - 1)
                plan.append((kind, src, p, size, soff, doff))
        walls = []
        for p in self.leaves:
            for dim, upper in WALL_ORDER:
                if self._face(p, dim, upper)[0] == "wall":
                    walls.append((p, dim, upper))
        return plan, walls

    def step(self, end_time=np.inf):
        """One global step; returns the applied dt (0 when the end time is
        reached)."""
        B = self.b
        max_dt = end_time - (self.time + self.dt)
        if not (max_dt - 0.0 > LA_EPS):
            return 0.0
        self.fill_ghosts()
        cfl = min(B.cfl(p.handle) for p in self.leaves)
        bound = min(cfl, max_dt)
        dt = 0.0
        for p in self.leaves:
            dt = B.solve(p.handle, bound)
        if dt - 0.0 > LA_EPS:
            self.time, self.dt = self.time + self.dt, dt
        self.steps += 1
        return dt

    def fill_ghosts(self):
        """Restriction into the virtual patches, uNew -> uOld, and the ghost
        layers of every leaf (steps 1-4 above)."""
        B = self.b
        for v in self.virtuals:
            B.clear(v.handle)
            for c in v.children:
                B.restrict(c.handle, v.handle)
            B.post_restrict(v.handle)
        for p in self.leaves:
            B.copy_unew_to_uold(p.handle)
        plan, walls = self._plan
        for kind, src, dst, size, soff, doff in plan:
            if kind == "transfer":
                B.transfer(size, soff, doff, src.handle, dst.handle)
            else:
                B.interpolate(size, doff, src.handle, dst.handle, 1, 0, 1)
        for p, dim, upper in walls:
            B.fill_boundary(p.handle, dim, upper)

    def leaf_states(self):
        """uNew of every leaf, keyed by (level, i, j)."""
        return {(p.level,) + p.ij: self.b.state(p.handle) for p in self.leaves}


class DeviceBackend:
    """Numerics calls on the GPU through the C ABI (one pool)."""

    def __init__(self, pool):
        self.pool = pool
        self.lib = load()

    def new_patch(self, pos, size, n, level, virtual):
        return self.pool.register(make_description(pos, size, n, level=level,
                                                   is_virtual=virtual))

    def set_state(self, s, h, b):
        n = h.shape[0]
        unew = np.zeros((3, n, n))
        unew[0] = h
        self.pool.upload(s, unew=unew, uold=np.zeros((3, n + 2, n + 2)), b=b)

    def clear(self, s):
        n = self.pool.dims(s)[0]
        self.pool.upload(s, unew=np.zeros((3, n, n)), uold=np.zeros((3, n + 2, n + 2)))

    def restrict(self, src, dst):
        check(self.lib.pc_restrict_solution(self.pool.h, src, dst, 0))

    def post_restrict(self, dst):
        check(self.lib.pc_post_process_restriction(self.pool.h, dst, 0))

    def copy_unew_to_uold(self, s):
        check(self.lib.pc_copy_unew_to_uold(self.pool.h, s))

    def transfer(self, size, soff, doff, src, dst):
        check(self.lib.pc_transfer_ghostlayer(self.pool.h, iarr(*size), iarr(*soff),
                                              iarr(*doff), src, dst))

    def interpolate(self, size, doff, src, dst, to_uold, to_current, use_time):
        check(self.lib.pc_interpolate_solution(self.pool.h, iarr(*size), iarr(*doff), src, dst,
                                               to_uold, to_current, use_time))

    def fill_boundary(self, s, dim, upper):
        check(self.lib.pc_fill_boundary_layer(self.pool.h, s, dim, upper))

    def cfl(self, s):
        v = ctypes.c_double()
        check(self.lib.pc_patch_cfl(self.pool.h, s, ctypes.byref(v)))
        return v.value

    def solve(self, s, max_dt):
        dt, est = ctypes.c_double(), ctypes.c_double()
        check(self.lib.pc_solve_timestep(self.pool.h, s, float(max_dt), iarr(0, 0, 0, 0),
                                         ctypes.byref(dt), ctypes.byref(est)))
        return dt.value

    def state(self, s):
        return self.pool.download(s)[0]


def refined_set_from_flags(pool, px, n, fn, threshold=0.05):
    """Coarse patches the BreakingDam criterion refines on the initial data
    (SolveTimestep::leaveCell with the demanded mesh width of
    BreakingDam.cpp:85-139; minimal width one level finer, maximal the
    coarse width): evaluated on the device for a px x px grid of n x n
    cells in a scratch scenario of `pool`, which is released afterwards."""
    size = 1.0 / px
    slots = {}
    for i in range(px):
        for j in range(px):
            s = pool.register(make_description((i * size, j * size), size, n, scenario=0))
            slots[(i, j)] = s
    ii = np.arange(-1, n + 1)
    for (i, j), s in slots.items():
        sub = size / n
        X = i * size + (ii[:, None] + 0.5) * sub
        Y = j * size + (ii[None, :] + 0.5) * sub
        X, Y = np.broadcast_arrays(X, Y)
        h, b = fn(X, Y)
        unew = np.zeros((3, n, n))
        unew[0] = h[1:-1, 1:-1]
        pool.upload(s, unew=unew, b=np.ascontiguousarray(b, dtype=np.float64))
    dx = size / n
    flags = pool.refinement_flags(threshold, dx / 3.0, dx, max(slots.values()) + 1)
    refined = {ij for ij, s in slots.items() if flags[s] == 1}
    for s in slots.values():
        pool.release(s)
    return refined


def dam_break_fields(centre=(0.5, 0.5), radius=0.25, h_in=2.0, h_out=1.0, hills=()):
    """Initial data of the 2-level dam break: radial column, optional
    Gaussian hills (x, y, sigma, height) as bathymetry."""
    def fn(X, Y):
        r = np.sqrt((X - centre[0]) ** 2 + (Y - centre[1]) ** 2)
        h = np.where(r <= radius, h_in, h_out)
        b = np.zeros(X.shape)
        for x, y, s, a in hills:
            b = b + a * np.exp(-((X - x) ** 2 + (Y - y) ** 2) / (s * s))
        return h, b
    return fn


__all__ = ["TwoLevelGrid", "DeviceBackend", "dam_break_fields", "refined_set_from_flags"]
