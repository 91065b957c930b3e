"""Adaptive patch grids (refinement factor 3) with global time steps.

The grid is px x py square coarse patches of n x n cells, size 1/px (the
domain [0, 1] x [0, py/px]); a refined patch becomes virtual and holds
3 x 3 patches of n x n cells one level finer, to any depth.  One step is the sequence of Numerics calls
PeanoClaw's traversal makes for a grid whose leaves all share one clock
(maximumTimestepSize = the global CFL minimum, SolveTimestep.cpp:591-751):

1. every virtual patch is cleared; every leaf, in depth-first tree order,
   is restricted into each virtual patch that contains it, nearest first
   (GridLevelTransfer::restrictToOverlappingVirtualSubgrids,
   GridLevelTransfer.cpp:140-198: only leaves restrict, so a leaf two levels
   down feeds its grandparent directly and restricted data, which lives on
   the [0,1] time basis, is never restricted again); then the SWE
   post-processing of every virtual patch (Numerics::restrictSolution /
   postProcessRestriction, SWEKernel.cpp:288-316);
2. every leaf copies uNew to uOld (SolveTimestep.cpp:638);
3. every leaf face gets its ghost layer from the patch across it: a
   same-level leaf or virtual patch by transferGhostlayer
   (GhostlayerCompositorFunctors.cpp:59), otherwise the coarser leaf covering
   that region by interpolateSolution(toUOld, not toCurrentTime,
   useTimeUNewOrUOld) (GhostlayerCompositorFunctors.cpp:61-69);
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
        self.age = 0  # global steps since creation
        self.k = -1  # index in the packed patch arrays (packed_plan_native)


class AmrGrid:
    """Patch tree with refinement factor 3: level 1 is a px x py grid of
    square patches of n x n cells, size 1/px (the domain [0, 1] x [0, py/px]); a patch at level l that `refine(l,
    (i, j))` selects becomes virtual and holds 3 x 3 patches at level l+1
    (indices (3i+a, 3j+c)), down to `max_level`.  `regrid` refines and
    coarsens during a run."""

    def __init__(self, backend, px, py, n, refine, fn, max_level=2, pure=False):
        self.b = backend
        # pure: transfers between levels read h + b without transforming the
        # source in place (the semantics of the batched device step)
        self.pure = pure
        self.px, self.py, self.n = px, py, n
        self.max_level = max_level
        self.fn = fn
        self.levels = {}
        # square patches of size 1/px: the domain is [0, 1] x [0, py/px]
        todo = [(1, (i, j), (i / px, j / px), 1.0 / px) for i in range(px) for j in range(py)]
        while todo:
            nxt = []
            for lev, ij, pos, size in todo:
                virt = lev < max_level and bool(refine(lev, ij))
                pt = AmrPatch(backend.new_patch(pos, size, n, lev, virt), lev, pos, size, n, virt,
                              ij)
                self.levels.setdefault(lev, {})[ij] = pt
                if virt:
                    sf = size / 3.0
                    for a in range(3):
                        for c in range(3):
                            nxt.append((lev + 1, (3 * ij[0] + a, 3 * ij[1] + c),
                                        (pos[0] + a * sf, pos[1] + c * sf), sf))
            todo = nxt
        for lev in sorted(self.levels):
            for pt in self.levels[lev].values():
                h, bb = self._fields(pt)
                backend.set_state(pt.handle, np.ascontiguousarray(h[1:-1, 1:-1]), bb)
        self.time, self.dt = 0.0, 0.0
        self.steps = 0
        self._rebuild()

    def _fields(self, pt):
        """(h, b) of the initial data at the cell centres of pt, ghost ring
        included."""
        n = self.n
        ii = np.arange(-1, n + 1)
        sub = pt.size / n
        X = pt.pos[0] + (ii[:, None] + 0.5) * sub
        Y = pt.pos[1] + (ii[None, :] + 0.5) * sub
        X, Y = np.broadcast_arrays(X, Y)
        h, bb = self.fn(X, Y)
        return (np.ascontiguousarray(h, dtype=np.float64),
                np.ascontiguousarray(bb, dtype=np.float64))

    def _b_many(self, pts):
        """b of the initial data of several patches (as _fields, one
        evaluation of the field function over all of them; the same
        operations per cell, so the same values)."""
        if not pts:
            return []
        n = self.n
        ii = np.arange(-1, n + 1)
        px = np.array([pt.pos[0] for pt in pts])[:, None, None]
        py = np.array([pt.pos[1] for pt in pts])[:, None, None]
        sub = np.array([pt.size / n for pt in pts])[:, None, None]
        X = px + (ii[None, :, None] + 0.5) * sub
        Y = py + (ii[None, None, :] + 0.5) * sub
        X, Y = np.broadcast_arrays(X, Y)
        _, bb = self.fn(X, Y)
        bb = np.broadcast_to(bb, X.shape)
        return [np.ascontiguousarray(bb[k], dtype=np.float64) for k in range(len(pts))]

    def _rebuild(self, changed=None):
        """Children, leaves, virtual patches, restriction and ghost plans
        from the current tree."""
        levels = self.levels
        order = sorted(levels)
        for lev in order:
            finer = levels.get(lev + 1)
            for (i, j), pt in levels[lev].items():
                if pt.virtual:
                    i3, j3 = 3 * i, 3 * j
                    pt.children = [finer[(i3 + a, j3 + c)] for a in (0, 1, 2) for c in (0, 1, 2)]
                elif pt.children:
                    pt.children = []
        every = [pt for lev in order for pt in levels[lev].values()]
        self.virtuals = [pt for pt in every if pt.virtual]
        self.leaves = [pt for pt in every if not pt.virtual]
        # (leaf, virtual ancestors nearest first) in depth-first tree order;
        # the leaves below one virtual patch share its (read-only) list
        restrictions = self._restrictions = []

        def walk(pt, ancestors):
            nearest = [pt] + ancestors
            for c in pt.children:
                if c.virtual:
                    walk(c, nearest)
                else:
                    restrictions.append((c, nearest))
        top = levels[1]
        for i in range(self.px):
            for j in range(self.py):
                pt = top[(i, j)]
                if pt.virtual:
                    walk(pt, [])
        # the face plan is derived when first used (the batched driver takes
        # the library's packed rows instead): the changes since the last
        # plan accumulate for its incremental update (None: everything)
        self._plan_value = None
        if changed is None or getattr(self, "_plan_changed", None) is None:
            self._plan_changed = None  # a full derivation is pending
        else:
            self._plan_changed |= set(changed)

    @property
    def _plan(self):
        """(face and corner entries, walls) of every leaf (_ghost_plan)."""
        if self._plan_value is None:
            self._plan_value = self._ghost_plan(self._plan_changed)
            self._plan_changed = set()
        return self._plan_value

    def regrid(self, threshold=0.05, minimal_width=None, maximal_width=None):
        """One remeshing pass on the current state with the BreakingDam
        criterion (demanded mesh width minimal_width where the maximal
        forward-difference |grad h| of a leaf exceeds `threshold`, else
        maximal_width; SolveTimestep::leaveCell, SolveTimestep.cpp:753-797),
        defaults: the finest and the coarsest subcell width.
        - A leaf flagged for refinement below max_level becomes virtual and
          gets 3 x 3 new leaves on the leaves' clock, b from the initial
          data, uNew and uOld interpolated from it (Remesh::createCell,
          Remesh.cpp:331-417: interpolateSolution to uNew at the new time,
          then to uOld at the current time, without neighbour times).
        - A virtual patch whose nine children are leaves that all allow
          coarsening and have taken at least 5 steps (the age veto,
          SolveTimestep.cpp:791-797) receives the restriction of all of
          them and becomes a leaf: its [0, 1]-basis values are evaluated at
          the clock's time and time + dt (Patch::switchToLeaf,
          Patch.cpp:793-830); the children are released.
        Decisions are per patch: Peano's vertex-based refinement (refining
        a cell's vertices refines the cells around them; erase vetoes
        shared across vertices) is not reproduced.
        Returns (refined, coarsened) patch counts."""
        B = self.b
        n = self.n
        finest = 1.0 / (self.px * 3 ** (self.max_level - 1) * n)
        coarsest = 1.0 / (self.px * n)
        minw = finest if minimal_width is None else minimal_width
        maxw = coarsest if maximal_width is None else maximal_width
        flags = B.flags([p.handle for p in self.leaves], threshold, minw, maxw)
        flag = {p.handle: f for p, f in zip(self.leaves, flags)}
        refine = [p for p in self.leaves if flag[p.handle] == 1 and p.level < self.max_level]
        coarsen = [v for v in self.virtuals
                   if all(not c.virtual and flag.get(c.handle) == 0 and c.age >= 5
                          for c in v.children)]
        changed = set()
        # the coarsened patches share no patch: their clear / restrict /
        # post-process / switch sequences run as batches, in that order
        if coarsen:
            B.clear_batch([v.handle for v in coarsen])
            B.restrict_batch([(c.handle, v.handle) for v in coarsen for c in v.children])
            for v in coarsen:
                B.post_restrict(v.handle)
            B.switch_to_leaf_batch([v.handle for v in coarsen], self.time, self.dt)
        for v in coarsen:
            changed.add(v.handle)
            changed.update(c.handle for c in v.children)
            for c in v.children:
                del self.levels[c.level][c.ij]
                B.release(c.handle)
            v.virtual = False
            v.age = 0
        # all new patches are registered first, then filled: a device pool
        # rebuilds its work lists once instead of once per new patch
        created = []
        for p in refine:
            sf = p.size / 3.0
            lev = p.level + 1
            for a in range(3):
                for c in range(3):
                    ij = (3 * p.ij[0] + a, 3 * p.ij[1] + c)
                    pos = (p.pos[0] + a * sf, p.pos[1] + c * sf)
                    ch = AmrPatch(B.new_patch(pos, sf, n, lev, False, self.time, self.dt), lev,
                                  pos, sf, n, False, ij)
                    created.append((p, ch))
                    changed.add(ch.handle)
                    self.levels.setdefault(lev, {})[ij] = ch
        # b of the new patches in one upload, then their interpolations
        # (uNew at the new time, uOld at the current time; the nine children
        # of a parent in order) as one batch
        B.set_b_batch([ch.handle for _, ch in created], self._b_many([ch for _, ch in created]))
        B.interpolate_batch([call for p, ch in created
                             for call in (((n, n), (0, 0), p.handle, ch.handle, 0, 0, 0),
                                          ((n, n), (0, 0), p.handle, ch.handle, 1, 1, 0))])
        for p in refine:
            B.set_virtual(p.handle, True)
            p.virtual = True
            changed.add(p.handle)
        if refine or coarsen:
            self._rebuild(changed)
        return len(refine), len(coarsen)

    @property
    def cells(self):
        return len(self.leaves) * self.n * self.n

    def refined_fraction(self):
        """Fraction of the domain covered by leaves finer than level 1."""
        return sum(p.size * p.size for p in self.leaves if p.level > 1)

    def run(self, steps, regrid_every=0, end_time=np.inf, **criterion):
        """`steps` global steps with a regrid after every `regrid_every`
        (0: never); returns the refined fraction after each regrid."""
        fractions = []
        for k in range(1, steps + 1):
            self.step(end_time)
            if regrid_every > 0 and k % regrid_every == 0:
                self.regrid(**criterion)
                fractions.append(self.refined_fraction())
        return fractions

    def _face(self, p, dim, upper):
        """('transfer'|'interp'|'wall', source patch) for a face of leaf p: the
        same-level patch across the face, else the leaf covering that region
        on the nearest coarser level."""
        lev = p.level
        i, j = p.ij
        q = (i + (1 if upper else -1), j) if dim == 0 else (i, j + (1 if upper else -1))
        k = 3 ** (lev - 1)
        if not (0 <= q[0] < self.px * k and 0 <= q[1] < self.py * k):
            return ("wall", None)
        same = self.levels[lev].get(q)
        if same is not None:
            return ("transfer", same)
        while lev > 1:
            lev -= 1
            q = (q[0] // 3, q[1] // 3)
            if q in self.levels[lev]:
                return ("interp", self.levels[lev][q])
        raise RuntimeError("face region not covered")

    def _ghost_plan(self, changed=None):
        """Face and corner entries and walls of every leaf.  With `changed`
        (handles of the patches a regrid created, refined, coarsened or
        released), only the leaves that are new or whose sources changed are
        looked up again; the others keep their entries."""
        n = self.n
        # region geometry per face: (size, source offset, destination offset)
        geom = {(0, 0): ((1, n), (n - 1, 0), (-1, 0)), (0, 1): ((1, n), (0, 0), (n, 0)),
                (1, 0): ((n, 1), (0, n - 1), (0, -1)), (1, 1): ((n, 1), (0, 0), (0, n))}
        face = self._face
        wall_rank = [FACES.index(f) for f in WALL_ORDER]
        steps = [(-1, 0), (1, 0), (0, -1), (0, 1)]  # FACES order
        old = getattr(self, "_face_cache", {}) if changed is not None else {}
        cache = {}
        plan = []
        walls = []
        for p in self.leaves:
            c = old.get(p.handle)
            if c is None or c[0] is not p or p.handle in changed or not c[3].isdisjoint(changed):
                # same-level neighbours by direct lookup, the rest by _face
                lv = self.levels[p.level]
                k = 3 ** (p.level - 1)
                nx_, ny_ = self.px * k, self.py * k
                i, j = p.ij
                res = []
                for (di, dj), (dim, upper) in zip(steps, FACES):
                    qi, qj = i + di, j + dj
                    if not (0 <= qi < nx_ and 0 <= qj < ny_):
                        res.append(("wall", None))
                        continue
                    same = lv.get((qi, qj))
                    res.append(("transfer", same) if same is not None else face(p, dim, upper))
                pe = [(kind, src, p) + geom[f] for f, (kind, src) in zip(FACES, res)
                      if kind != "wall"]
                # corner ghosts from the diagonal neighbour, as the edge
                # functor fills them (GhostlayerCompositorFunctors.cpp:87-141):
                # the patch of this level there (transfer), else the coarser
                # leaf covering it (interpolation); outside the domain the
                # walls fill them
                for di, dj in ((-1, -1), (-1, 1), (1, -1), (1, 1)):
                    qi, qj = i + di, j + dj
                    if not (0 <= qi < nx_ and 0 <= qj < ny_):
                        continue
                    src = self._covering(p.level, (qi, qj))
                    kind = "transfer" if src.level == p.level else "interp"
                    soff = (n - 1 if di == -1 else 0, n - 1 if dj == -1 else 0)
                    doff = (-1 if di == -1 else n, -1 if dj == -1 else n)
                    pe.append((kind, src, p, (1, 1), soff, doff))
                we = [(p, FACES[k][0], FACES[k][1]) for k in wall_rank if res[k][0] == "wall"]
                # the same entries packed as the rows of the device plan
                # (pc_amr_ghost: kind, src, dst, size, src_offset, dst_offset;
                # pc_amr_wall: slot, dimension, upper)
                rg = np.array([(0 if e[0] == "transfer" else 1, e[1].handle, p.handle,
                                e[3][0], e[3][1], e[4][0], e[4][1], e[5][0], e[5][1])
                               for e in pe], dtype=np.int32).reshape(-1, 9)
                rw = np.array([(p.handle, d, u) for _, d, u in we], dtype=np.int32).reshape(-1, 3)
                c = (p, pe, we, {e[1].handle for e in pe}, rg, rw)
            cache[p.handle] = c
            plan.extend(c[1])
            walls.extend(c[2])
        self._face_cache = cache
        return plan, walls

    def packed_plan_native(self):
        """The rows of packed_plan() built by the library (pc_amr_face_plan)
        from the tree alone: the same rules, without deriving the Python
        plan.  Patch handles must be ints."""
        lib = load()
        pts = [pt for lev in sorted(self.levels) for pt in self.levels[lev].values()]
        m, nl = len(pts), len(self.leaves)
        for k, pt in enumerate(pts):
            pt.k = k
        level = np.fromiter((pt.level for pt in pts), np.int32, m)
        ij = np.fromiter((x for pt in pts for x in pt.ij), np.int32, 2 * m)
        handle = np.fromiter((pt.handle for pt in pts), np.int32, m)
        leaves = np.fromiter((p.k for p in self.leaves), np.int32, nl)
        gh = np.empty((8 * nl + 1, 9), np.int32)
        wl = np.empty((4 * nl + 1, 3), np.int32)
        ng, nw = ctypes.c_int(), ctypes.c_int()
        vp = ctypes.c_void_p
        check(lib.pc_amr_face_plan(self.px, self.py, self.n, len(pts),
                                   level.ctypes.data_as(vp), ij.ctypes.data_as(vp),
                                   handle.ctypes.data_as(vp), leaves.ctypes.data_as(vp),
                                   len(leaves), gh.ctypes.data_as(vp), len(gh),
                                   ctypes.byref(ng), wl.ctypes.data_as(vp), len(wl),
                                   ctypes.byref(nw)))
        return gh[:ng.value].copy(), wl[:nw.value].copy()

    def packed_plan(self):
        """The face entries and walls of the current plan as int32 rows in
        the layout of pc_amr_ghost and pc_amr_wall (plan order)."""
        self._plan  # derive the plan (and its per-leaf rows) if stale
        cache = self._face_cache
        rows = [cache[p.handle] for p in self.leaves]
        gh = np.concatenate([c[4] for c in rows]) if rows else np.zeros((0, 9), np.int32)
        wl = np.concatenate([c[5] for c in rows]) if rows else np.zeros((0, 3), np.int32)
        return gh, wl

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
            p.age += 1
        if dt - 0.0 > LA_EPS:
            self.time, self.dt = self.time + self.dt, dt
        self.steps += 1
        return dt

    def fill_ghosts(self):
        """Restriction of the leaves into the virtual patches, uNew -> uOld,
        and the ghost layers of every leaf (steps 1-4 above)."""
        B = self.b
        for v in self.virtuals:
            B.clear(v.handle)
        restrict = B.restrict_pure if self.pure else B.restrict
        for leaf, ancestors in self._restrictions:
            for v in ancestors:
                restrict(leaf.handle, v.handle)
        for v in self.virtuals:
            B.post_restrict(v.handle)
        for p in self.leaves:
            B.copy_unew_to_uold(p.handle)
        plan, walls = self._plan
        interpolate = B.interpolate_pure if self.pure else B.interpolate
        for kind, src, dst, size, soff, doff in plan:
            if kind == "transfer":
                B.transfer(size, soff, doff, src.handle, dst.handle)
            else:
                interpolate(size, doff, src.handle, dst.handle, 1, 0, 1)
        for p, dim, upper in walls:
            B.fill_boundary(p.handle, dim, upper)

    # ---- local time stepping ----

    def _covering(self, lev, q):
        """The patch at `lev` with index q, else the patch of the nearest
        coarser level covering that region (None outside the domain)."""
        k = 3 ** (lev - 1)
        if not (0 <= q[0] < self.px * k and 0 <= q[1] < self.py * k):
            return None
        while lev >= 1:
            if q in self.levels[lev]:
                return self.levels[lev][q]
            lev -= 1
            q = (q[0] // 3, q[1] // 3)
        return None

    def _lts_tables(self):
        """Per leaf: the patches sharing a face or a corner with it (same
        level if present, else the coarser patch covering the region: the
        four patches around each of its vertices, GhostLayerCompositor.h:52-56),
        its ghost-face entries and its walls; per virtual patch the leaves
        below it in depth-first order."""
        plan, walls = self._plan
        faces = {p.handle: [] for p in self.leaves}
        for entry in plan:
            faces[entry[2].handle].append(entry)
        wl = {p.handle: [] for p in self.leaves}
        for p, dim, upper in walls:
            wl[p.handle].append((dim, upper))
        nbrs = {}
        for p in self.leaves:
            seen = []
            for di in (-1, 0, 1):
                for dj in (-1, 0, 1):
                    if di == 0 and dj == 0:
                        continue
                    q = self._covering(p.level, (p.ij[0] + di, p.ij[1] + dj))
                    if q is not None and q is not p and all(q is not x for x in seen):
                        seen.append(q)
            nbrs[p.handle] = seen
        below = {v.handle: [] for v in self.virtuals}
        for leaf, ancestors in self._restrictions:
            for v in ancestors:
                below[v.handle].append(leaf)
        return faces, wl, nbrs, below

    def clock(self, pt):
        """(time, timestep size) of a leaf, or for a virtual patch the end of
        the minimal time interval of the leaves below it as (end, 0)
        (TimeIntervals::updateMinimalFineGridTimeInterval,
        TimeIntervals.cpp:190-214: the intersection of the fine intervals
        ends at their minimal end)."""
        if not pt.virtual:
            return self.b.clock(pt.handle)
        return (min(sum(self.b.clock(x.handle)) for x in self._lts[3][pt.handle]), 0.0)

    def lts_advancing(self, end_time):
        """The leaves that may advance now (TimeIntervals::
        isAllowedToAdvanceInTime, TimeIntervals.cpp:24-78): the step must
        not pass the global end time, and the leaf's current end time must
        not exceed the time constraint (end of the current interval) of any
        patch sharing a face or a corner with it (isBlockedByNeighbors,
        TimeIntervals.cpp:80-84; a virtual patch constrains by the minimal
        end of its leaves), with tarch's 1e-8 tolerance."""
        if getattr(self, "_lts", None) is None or self._lts_key is not self._plan:
            self._lts = self._lts_tables()
            self._lts_key = self._plan
        nbrs = self._lts[2]
        end = {p.handle: sum(self.b.clock(p.handle)) for p in self.leaves}
        vend = {v.handle: min(end[x.handle] for x in self._lts[3][v.handle])
                for v in self.virtuals}
        out = []
        for p in self.leaves:
            e = end[p.handle]
            if not ((end_time - e) - 0.0 > LA_EPS):
                continue
            c = min((vend[q.handle] if q.virtual else end[q.handle] for q in nbrs[p.handle]),
                    default=np.inf)
            if e - c > LA_EPS:
                continue
            out.append(p)
        return out

    def lts_iteration(self, end_time):
        """One grid iteration with local time stepping: every leaf that may
        advance (`lts_advancing`, decided on the clocks at the start of the
        iteration) takes one step of its own CFL size
        (SolveTimestep::enterCell, SolveTimestep.cpp:591-751), phase by
        phase over those leaves: restriction of all leaves into the virtual
        patches; the ghost faces of the advancing leaves from their
        neighbours' current intervals (the time factors of DefaultTransfer /
        DefaultInterpolation; every neighbour is read before any leaf
        changes); uNew -> uOld; walls; solveTimestep with
        maximumTimestepSize = end time - (time + timestep size).  Returns
        the number of leaves that advanced."""
        B = self.b
        adv = self.lts_advancing(end_time)
        if not adv:
            return 0
        faces, walls = self._lts[0], self._lts[1]
        for v in self.virtuals:
            B.clear(v.handle)
        restrict = B.restrict_pure if self.pure else B.restrict
        for leaf, ancestors in self._restrictions:
            for v in ancestors:
                restrict(leaf.handle, v.handle)
        for v in self.virtuals:
            B.post_restrict(v.handle)
        interpolate = B.interpolate_pure if self.pure else B.interpolate
        for p in adv:
            for kind, src, dst, size, soff, doff in faces[p.handle]:
                if kind == "transfer":
                    B.transfer(size, soff, doff, src.handle, dst.handle)
                else:
                    interpolate(size, doff, src.handle, dst.handle, 1, 0, 1)
        for p in adv:
            B.copy_unew_to_uold(p.handle)
        for p in adv:
            for dim, upper in walls[p.handle]:
                B.fill_boundary(p.handle, dim, upper)
        for p in adv:
            t, dt = B.clock(p.handle)
            B.solve(p.handle, end_time - (t + dt))
            p.age += 1
        self.steps += 1
        return len(adv)

    def run_lts(self, end_time, max_iterations=100000):
        """Grid iterations with local time stepping until every leaf has
        reached `end_time` (or `max_iterations`); returns the number of
        iterations and the number of leaf steps."""
        its = leaf_steps = 0
        while its < max_iterations:
            k = self.lts_iteration(end_time)
            its += 1
            leaf_steps += k
            if k == 0:
                break
        return its, leaf_steps

    def leaf_states(self):
        """uNew of every leaf, keyed by (level, i, j)."""
        return {(p.level,) + p.ij: self.b.state(p.handle) for p in self.leaves}


class TwoLevelGrid(AmrGrid):
    """Two levels: the coarse patches in `refined` hold 3 x 3 fine patches."""

    def __init__(self, backend, px, py, n, refined, fn):
        refined = set(refined)
        super().__init__(backend, px, py, n, lambda lev, ij: lev == 1 and ij in refined, fn,
                         max_level=2)
        self.coarse = self.levels[1]
        self.fine = self.levels.get(2, {})


class DeviceBackend:
    """Numerics calls on the GPU through the C ABI (one pool)."""

    def __init__(self, pool):
        self.pool = pool
        self.lib = load()

    def new_patch(self, pos, size, n, level, virtual, time=0.0, dt=0.0):
        return self.pool.register(make_description(pos, size, n, level=level,
                                                   is_virtual=virtual, time=time,
                                                   timestep_size=dt))

    def set_b(self, s, b):
        self.pool.upload(s, b=b)

    def set_virtual(self, s, virtual):
        d = self.pool.description(s)
        d.is_virtual = 1 if virtual else 0
        self.pool.set_description(s, d)

    def release(self, s):
        self.pool.release(s)

    def flags(self, slots, threshold, minimal_width, maximal_width):
        f = self.pool.refinement_flags(threshold, minimal_width, maximal_width,
                                       max(slots) + 1)
        return [int(f[s]) for s in slots]

    def switch_to_leaf(self, s, time, dt):
        self.switch_to_leaf_batch([s], time, dt)

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

    # batched forms: the effect of the per-call methods in the order given
    def interpolate_batch(self, calls):
        """calls: (size, doff, src, dst, to_uold, to_current, use_time)."""
        if not calls:
            return
        size = _ints([c[0] for c in calls])
        doff = _ints([c[1] for c in calls])
        src, dst = _ints([c[2] for c in calls]), _ints([c[3] for c in calls])
        opts = _ints([c[4:7] for c in calls])
        check(self.lib.pc_interpolate_batch(self.pool.h, len(calls), _ip(size), _ip(doff),
                                            _ip(src), _ip(dst), _ip(opts)))

    def restrict_batch(self, pairs):
        """pairs: (src, dst), restricted into the virtual dst in order."""
        if not pairs:
            return
        src, dst = _ints([a for a, _ in pairs]), _ints([b for _, b in pairs])
        check(self.lib.pc_restrict_batch(self.pool.h, len(pairs), _ip(src), _ip(dst), 0))

    def clear_batch(self, handles):
        if handles:
            check(self.lib.pc_patches_clear(self.pool.h, len(handles), _ip(_ints(handles))))

    def set_b_batch(self, handles, bs):
        if handles:
            b = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64).ravel()
                                                     for x in bs]))
            check(self.lib.pc_patches_set_b(self.pool.h, len(handles), _ip(_ints(handles)),
                                            b.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def switch_to_leaf_batch(self, handles, time, dt):
        if handles:
            t = np.full(len(handles), float(time))
            d = np.full(len(handles), float(dt))
            dp = ctypes.POINTER(ctypes.c_double)
            check(self.lib.pc_switch_to_leaf(self.pool.h, len(handles), _ip(_ints(handles)),
                                             t.ctypes.data_as(dp), d.ctypes.data_as(dp)))

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

    def clock(self, s):
        d = self.pool.description(s)
        return d.time, d.timestep_size

    def state(self, s):
        return self.pool.download(s)[0]


def _ints(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int32).ravel())


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


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


def wet_dry_fields(seed=4, column=(0.05, 0.35, 0.3, 0.6333333333333333), column_h=1.5):
    """Wet/dry basin: b a slope from -1 (x = 0) to +0.5 (x = 1) plus six
    seeded bumps (|height| <= 0.15), still water at surface level 0 where
    b < 0, dry land (h = 0) elsewhere, a column of depth 1.5 over
    x0 <= x <= x1, y0 <= y <= y1 (10% of the domain)."""
    rng = np.random.default_rng(seed)
    bumps = [(rng.uniform(0.2, 0.9), rng.uniform(0.1, 0.9), rng.uniform(0.03, 0.1),
              rng.uniform(-0.15, 0.15)) for _ in range(6)]
    x0, x1, y0, y1 = column

    def fn(X, Y):
        b = -1.0 + 1.5 * X
        for cx, cy, s, a in bumps:
            b = b + a * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * s * s))
        h = np.maximum(0.0, 0.0 - b)
        col = (X >= x0) & (X <= x1) & (Y >= y0) & (Y <= y1)
        return np.where(col, column_h, h), b
    return fn


class BatchedAmr:
    """The grid's global steps as one batched device plan (pc_amr_*): the
    same restriction, ghost and wall sequence as AmrGrid.step with
    pure=True, seven launches per step.  `grid` must run on a DeviceBackend;
    afterwards its leaves hold the state and clock of the batched steps."""

    def __init__(self, grid):
        self.grid = grid
        self.pool = grid.b.pool
        self.lib = load()
        self.h = None
        self.steps_before = 0  # steps of earlier plans (before regrids)
        self._create()

    def _create(self):
        from ._abi import AmrGhost, AmrWall
        grid = self.grid
        leaves = iarr_list([p.handle for p in grid.leaves])
        virtuals = iarr_list([v.handle for v in grid.virtuals])
        pairs = [x for leaf, anc in grid._restrictions for v in anc for x in (leaf.handle,
                                                                                v.handle)]
        rest = iarr_list(pairs)
        # the face entries and walls as int32 rows in the layout of
        # pc_amr_ghost and pc_amr_wall, from the library
        assert ctypes.sizeof(AmrGhost) == 9 * 4 and ctypes.sizeof(AmrWall) == 3 * 4
        gh, wl = grid.packed_plan_native()
        ng, nw = len(gh), len(wl)
        gh = np.ascontiguousarray(gh if ng else np.zeros((1, 9), np.int32))
        wl = np.ascontiguousarray(wl if nw else np.zeros((1, 3), np.int32))
        h = ctypes.c_void_p()
        check(self.lib.pc_amr_create(self.pool.h, leaves, len(grid.leaves), virtuals,
                                     len(grid.virtuals), rest, len(pairs) // 2,
                                     gh.ctypes.data_as(ctypes.c_void_p), ng,
                                     wl.ctypes.data_as(ctypes.c_void_p), nw,
                                     ctypes.byref(h)))
        self.h = h

    def step(self, k=1, end_time=np.inf):
        """k global steps; returns the last step's dt (0 once the end time
        is reached)."""
        dt = ctypes.c_double()
        before = self.clock()[2]
        check(self.lib.pc_amr_step(self.h, int(k), float(end_time), ctypes.byref(dt)))
        taken = self.clock()[2] - before
        for p in self.grid.leaves:
            p.age += taken
        return dt.value

    def lts_init(self):
        """Switch the plan to local time stepping (pc_amr_lts_init): the
        face and corner neighbours of every leaf from the grid's tables,
        the leaves' clocks from their descriptors."""
        grid = self.grid
        if getattr(grid, "_lts", None) is None or grid._lts_key is not grid._plan:
            grid._lts = grid._lts_tables()
            grid._lts_key = grid._plan
        nbrs = grid._lts[2]
        ptr, idx = [0], []
        for p in grid.leaves:
            idx += [q.handle for q in nbrs[p.handle]]
            ptr.append(len(idx))
        check(self.lib.pc_amr_lts_init(self.h, iarr_list(ptr), iarr_list(idx)))

    def lts_iterations(self, k, end_time):
        """k grid iterations with local time stepping (the semantics of
        AmrGrid.lts_iteration with pure=True); returns the leaf steps taken."""
        n = ctypes.c_longlong()
        check(self.lib.pc_amr_lts_step(self.h, int(k), float(end_time), ctypes.byref(n)))
        return n.value

    def run_lts(self, end_time, batch=16, max_iterations=1000000):
        """Iterations in batches of `batch` until every leaf has reached
        `end_time`; returns (iterations launched, leaf steps)."""
        its = steps = 0
        while its < max_iterations:
            k = self.lts_iterations(batch, end_time)
            its += batch
            steps += k
            if k == 0:
                break
        return its, steps

    def clock(self):
        """(time, timestep_size, steps) of the leaves."""
        t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        check(self.lib.pc_amr_clock(self.h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)))
        return t.value, dt.value, self.steps_before + n.value

    def regrid(self, **criterion):
        """AmrGrid.regrid on the current state, then a new device plan."""
        t, dt, n = self.clock()
        self.grid.time, self.grid.dt = t, dt
        counts = self.grid.regrid(**criterion)
        if counts != (0, 0):
            self.steps_before = n
            check(self.lib.pc_amr_destroy(self.h))
            self.h = None
            self._create()
        return counts

    def run(self, steps, regrid_every=0, end_time=np.inf, **criterion):
        """`steps` global steps with a regrid after every `regrid_every`
        (0: never); returns the refined leaf-cell fraction after each
        regrid."""
        fractions = []
        done = 0
        while done < steps:
            k = steps - done if regrid_every <= 0 else min(regrid_every, steps - done)
            self.step(k, end_time)
            done += k
            if regrid_every > 0 and done % regrid_every == 0:
                self.regrid(**criterion)
                fractions.append(self.grid.refined_fraction())
        return fractions

    def close(self):
        if self.h:
            check(self.lib.pc_amr_destroy(self.h))
            self.h = None


def iarr_list(v):
    """int32 array of v (at least one element) as an int pointer; the pointer
    keeps the array alive."""
    a = np.ascontiguousarray(np.asarray(v if len(v) else [0], dtype=np.int32))
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def front_refinement(px, py, n, fn, max_level, jump=0.5):
    """refine(level, ij) for the wet/dry basin: a patch refines when its
    cells (initial data) hold both dry (h = 0) and wet cells, or when the
    depth jumps by more than `jump` between neighbouring cells (the column
    edge).  Evaluated once on the host at the cell centres."""
    def refine(lev, ij):
        if lev >= max_level:
            return False
        size = 1.0 / (px * 3 ** (lev - 1))
        c = (np.arange(n) + 0.5) * (size / n)
        X, Y = np.meshgrid(ij[0] * size + c, ij[1] * size + c, indexing="ij")
        h, _ = fn(X, Y)
        dry = h == 0.0
        if dry.any() and not dry.all():
            return True
        return bool(max(np.abs(np.diff(h, axis=0)).max(), np.abs(np.diff(h, axis=1)).max())
                    > jump)
    return refine


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


__all__ = ["AmrGrid", "BatchedAmr", "TwoLevelGrid", "DeviceBackend", "dam_break_fields",
           "front_refinement", "refined_set_from_flags", "wet_dry_fields"]
