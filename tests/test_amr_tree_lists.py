"""The tree lists the AMR drivers derive after every change of the grid
(leaves, virtual patches, the (leaf, virtual ancestors) restriction pairs in
depth-first order) and the packed patch arrays of the native face plan,
against an independent walk of the same tree.  Host only: the backend does
no numerical work."""
import numpy as np
import pytest

from peanoclaw_amd.amr import AmrGrid


class _TreeOnly:
    """A backend that only hands out patch handles and replays given
    refinement flags (no data)."""

    def __init__(self, seed):
        self.next = 0
        self.rng = np.random.default_rng(seed)

    def new_patch(self, *args, **kwargs):
        self.next += 1
        return self.next

    def flags(self, handles, *args):
        return [int(f) for f in self.rng.choice([0, 1, 2], size=len(handles), p=[0.85, 0.1, 0.05])]

    def __getattr__(self, name):
        return lambda *args, **kwargs: None


def _flat_field(X, Y):
    return np.ones_like(X), np.zeros_like(X)


def _expected(grid):
    """Leaves, virtual patches (both by level, then in the level's insertion
    order) and the restriction pairs of a recursive depth-first walk from the
    level-1 patches in (i, j) order, children in (a, c) order."""
    pts = [pt for lev in sorted(grid.levels) for pt in grid.levels[lev].values()]
    leaves = [pt for pt in pts if not pt.virtual]
    virtuals = [pt for pt in pts if pt.virtual]
    pairs = []

    def walk(pt, chain):
        if not pt.virtual:
            if chain:
                pairs.append((pt.handle, [v.handle for v in chain]))
            return
        i, j = pt.ij
        for a in range(3):
            for c in range(3):
                walk(grid.levels[pt.level + 1][(3 * i + a, 3 * j + c)], [pt] + chain)
    for i in range(grid.px):
        for j in range(grid.py):
            walk(grid.levels[1][(i, j)], [])
    return leaves, virtuals, pairs


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_tree_lists_after_regrids(seed):
    rng = np.random.default_rng(100 + seed)
    chosen = {(lev, (int(i), int(j))) for lev in (1, 2)
              for i, j in rng.integers(0, 4 * 3 ** (lev - 1), size=(6, 2))}
    grid = AmrGrid(_TreeOnly(seed), 4, 3, 4, lambda lev, ij: (lev, ij) in chosen, _flat_field,
                   max_level=3)
    refined = coarsened = 0
    for it in range(8):
        leaves, virtuals, pairs = _expected(grid)
        assert grid.leaves == leaves and grid.virtuals == virtuals
        assert [(p.handle, [v.handle for v in anc]) for p, anc in grid._restrictions] == pairs
        for pt in grid.virtuals:
            i, j = pt.ij
            assert [c.ij for c in pt.children] == [(3 * i + a, 3 * j + c)
                                                   for a in range(3) for c in range(3)]
        assert all(pt.children == [] for pt in grid.leaves)
        # the packed arrays the native face plan takes: its rows equal the
        # Python driver's rows of the same tree
        gh_n, wl_n = grid.packed_plan_native()
        gh_p, wl_p = grid.packed_plan()
        assert np.array_equal(gh_n, gh_p) and np.array_equal(wl_n, wl_p)
        for pt in grid.leaves:
            pt.age = 10
        r, c = grid.regrid()
        refined += r
        coarsened += c
    assert refined > 0 and coarsened > 0


def test_deep_corner_tree_face_plan_with_dense_and_hashed_levels():
    """A tree refined at the domain corner down to level 12: the library's
    face plan looks the coarse levels up in dense tables and the fine ones
    (grids of up to 3^11 x 3^11 patches) in its hash map; its rows equal the
    Python driver's."""
    grid = AmrGrid(_TreeOnly(0), 1, 1, 2, lambda lev, ij: ij == (0, 0), _flat_field,
                   max_level=12)
    assert max(grid.levels) == 12 and len(grid.leaves) == 1 + 8 * 11
    gh_n, wl_n = grid.packed_plan_native()
    gh_p, wl_p = grid.packed_plan()
    assert len(gh_n) > 0 and np.array_equal(gh_n, gh_p) and np.array_equal(wl_n, wl_p)
