"""The batched level transfers of a regrid (pc_interpolate_batch,
pc_restrict_batch, pc_patches_clear, pc_patches_set_b, pc_switch_to_leaf)
through the C ABI.

A batch promises exactly the effect of its per-call entry points made one
after another, which are themselves pinned against the oracle
(test_gpu_parity, test_amr_gpu).  So each batch runs on one pool and the
same calls one by one on a second pool built identically; every field of
every patch must then be bitwise equal.  The batches hold calls that share
patches (the nine children of a parent, uNew then uOld; a second
interpolation into a patch already written; several leaves into one virtual
patch, with only-overlapped regions), which the batch must keep in order.
pc_switch_to_leaf is compared with the oracle's Patch::switchToLeaf."""
import ctypes

import numpy as np
import pytest

from oracle_lib import Patch as OPatch, lib as olib

pytestmark = pytest.mark.gpu

N = 9  # cells per side: a child covers 3 x 3 parent cells


def _ip(v):
    a = np.ascontiguousarray(np.asarray(v, dtype=np.int32).ravel())
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _fill(pool, s, rng, n, virtual=False):
    unew = np.stack([rng.uniform(0.5, 2.0, (n, n)), rng.uniform(-0.4, 0.4, (n, n)),
                     rng.uniform(-0.4, 0.4, (n, n))])
    uold = np.stack([rng.uniform(0.5, 2.0, (n + 2, n + 2)),
                     rng.uniform(-0.4, 0.4, (n + 2, n + 2)),
                     rng.uniform(-0.4, 0.4, (n + 2, n + 2))])
    b = rng.uniform(0.0, 0.3, (n + 2, n + 2))
    pool.upload(s, unew=unew, uold=uold, b=b)


def _grid(seed, virtual_parents):
    """Two identical pools: 3 parents side by side (size 1, at x = 0, 1, 2)
    and their 27 children (size 1/3), random data and b, parents on
    [0.2, 0.25], children on [0.21, 0.22]."""
    from peanoclaw_amd.pool import PatchPool, make_description
    pools, layout = [], None
    for _ in range(2):
        rng = np.random.default_rng(seed)
        pool = PatchPool(0)
        parents, children = [], []
        for k in range(3):
            d = make_description((float(k), 0.0), 1.0, N, level=1, time=0.2,
                                 timestep_size=0.05, is_virtual=virtual_parents)
            if virtual_parents:
                # only-overlapped restriction: strips next to the borders
                d.restriction_lower_bounds[0] = k + 0.62
                d.restriction_lower_bounds[1] = 0.62
                d.restriction_upper_bounds[0] = k + 0.3
                d.restriction_upper_bounds[1] = 0.3
            s = pool.register(d)
            _fill(pool, s, rng, N)
            parents.append(s)
            for a in range(3):
                for c in range(3):
                    cs = pool.register(make_description((k + a / 3.0, c / 3.0), 1.0 / 3.0, N,
                                                        level=2, time=0.21, timestep_size=0.01))
                    _fill(pool, cs, rng, N)
                    children.append((s, cs))
        pools.append(pool)
        layout = (parents, children)
    return pools, layout


def _assert_same(pa, pb, slots):
    for s in slots:
        for x, y, name in zip(pa.download(s), pb.download(s), ("unew", "uold", "b")):
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), (s, name)


def test_interpolate_batch_equals_the_calls_in_order():
    (pa, pb), (parents, children) = _grid(7, False)
    try:
        calls = []
        for p, c in children:
            calls.append(((N, N), (0, 0), p, c, 0, 0, 0))   # uNew at the new time
            calls.append(((N, N), (0, 0), p, c, 1, 1, 0))   # uOld at the current time
        # a ghost-layer strip into a child that was written above, from the
        # next parent (shares patches with calls of two parents)
        p1, c0 = parents[1], children[8][1]
        calls.append(((1, N), (N, 0), p1, c0, 1, 0, 1))
        calls.append(((N - 2, 2), (1, 3), parents[0], children[4][1], 0, 1, 1))
        for c in calls:
            size, doff = (ctypes.c_int * 2)(*c[0]), (ctypes.c_int * 2)(*c[1])
            assert pa.lib.pc_interpolate_solution(pa.h, size, doff, *c[2:]) == 0
        size, psize = _ip([c[0] for c in calls])
        doff, pdoff = _ip([c[1] for c in calls])
        src, psrc = _ip([c[2] for c in calls])
        dst, pdst = _ip([c[3] for c in calls])
        opts, popts = _ip([c[4:7] for c in calls])
        assert pb.lib.pc_interpolate_batch(pb.h, len(calls), psize, pdoff, psrc, pdst,
                                           popts) == 0
        _assert_same(pa, pb, parents + [c for _, c in children])
        print(f"{len(calls)} interpolations: batch bitwise equal to the calls in order")
    finally:
        pa.close()
        pb.close()


@pytest.mark.parametrize("only_overlapped", [0, 1])
def test_restrict_batch_equals_the_calls_in_order(only_overlapped):
    (pa, pb), (parents, children) = _grid(11 + only_overlapped, True)
    try:
        for pool in (pa, pb):
            slots, ps = _ip(parents)
            assert pool.lib.pc_patches_clear(pool.h, len(parents), ps) == 0
        pairs = [(c, p) for p, c in children]
        for c, p in pairs:
            assert pa.lib.pc_restrict_solution(pa.h, c, p, only_overlapped) == 0
        src, psrc = _ip([c for c, _ in pairs])
        dst, pdst = _ip([p for _, p in pairs])
        assert pb.lib.pc_restrict_batch(pb.h, len(pairs), psrc, pdst, only_overlapped) == 0
        _assert_same(pa, pb, parents + [c for _, c in children])
        # the clear zeroed uNew's interior and all of uOld before the restrictions
        unew, uold, _ = pb.download(parents[0])
        assert np.count_nonzero(unew) > 0
    finally:
        pa.close()
        pb.close()


def test_clear_and_set_b():
    from peanoclaw_amd.pool import PatchPool, make_description
    rng = np.random.default_rng(3)
    with PatchPool(0) as pool:
        slots = [pool.register(make_description((k, 0.0), 1.0, n)) for k, n in
                 enumerate((4, 9, 16))]
        for s, n in zip(slots, (4, 9, 16)):
            _fill(pool, s, rng, n)
        bs = [rng.uniform(-1, 1, (n + 2, n + 2)) for n in (4, 9, 16)]
        _, ps = _ip(slots)
        flat = np.ascontiguousarray(np.concatenate([b.ravel() for b in bs]))
        assert pool.lib.pc_patches_set_b(pool.h, 3, ps, flat.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double))) == 0
        before = [pool.download(s) for s in slots]
        for s, b in zip(slots, bs):
            assert np.array_equal(pool.download(s)[2], b)
        _, p2 = _ip(slots[:2])
        assert pool.lib.pc_patches_clear(pool.h, 2, p2) == 0
        for k, s in enumerate(slots):
            unew, uold, b = pool.download(s)
            assert np.array_equal(b, bs[k])
            if k < 2:
                assert not unew.any() and not uold.any()
            else:
                assert np.array_equal(unew, before[k][0]) and np.array_equal(uold, before[k][1])
        # a slot twice, or a dead slot: rejected, nothing changed
        _, pdup = _ip([slots[2], slots[2]])
        assert pool.lib.pc_patches_clear(pool.h, 2, pdup) == 1
        _, pbad = _ip([slots[2], 99])
        assert pool.lib.pc_patches_clear(pool.h, 2, pbad) == 1
        assert np.array_equal(pool.download(slots[2])[0], before[2][0])


def test_switch_to_leaf_matches_the_oracle():
    from peanoclaw_amd.pool import PatchPool, make_description
    rng = np.random.default_rng(5)
    times = [(0.41, 0.02), (0.37, 0.0), (3.0, 0.5)]
    with PatchPool(0) as pool:
        slots, refs = [], []
        for k, (t, dt) in enumerate(times):
            s = pool.register(make_description((float(k), 0.0), 1.0, N, is_virtual=True))
            _fill(pool, s, rng, N)
            unew, uold, b = pool.download(s)
            P = OPatch(3, 1, N, 1, (float(k), 0.0), 1.0, 0, 0.0, 0.0)
            P.unew[:] = unew
            P.uold[:] = uold
            P.s.is_virtual, P.s.is_leaf = 1, 0
            P.s.time, P.s.dt = t, dt
            olib().or_switch_to_leaf(P.ref)
            slots.append(s)
            refs.append(P)
        _, ps = _ip(slots)
        tt = np.array([t for t, _ in times])
        dd = np.array([d for _, d in times])
        dp = ctypes.POINTER(ctypes.c_double)
        assert pool.lib.pc_switch_to_leaf(pool.h, len(slots), ps, tt.ctypes.data_as(dp),
                                          dd.ctypes.data_as(dp)) == 0
        for s, P, (t, dt) in zip(slots, refs, times):
            unew, uold, _ = pool.download(s)
            assert np.array_equal(unew.view(np.uint64), np.array(P.unew).view(np.uint64))
            assert np.array_equal(uold.view(np.uint64), np.array(P.uold).view(np.uint64))
            d = pool.description(s)
            assert d.is_virtual == 0 and d.time == t and d.timestep_size == dt


def test_invalid_call_fails_the_batch_with_nothing_changed():
    (pa, pb), (parents, children) = _grid(13, False)
    try:
        before = {s: pb.download(s) for s in parents + [c for _, c in children]}
        calls = [((N, N), (0, 0), p, c, 0, 0, 0) for p, c in children[:5]]
        calls.append(((N + 5, N), (0, 0), parents[0], children[0][1], 0, 0, 0))  # too large
        size, psize = _ip([c[0] for c in calls])
        doff, pdoff = _ip([c[1] for c in calls])
        src, psrc = _ip([c[2] for c in calls])
        dst, pdst = _ip([c[3] for c in calls])
        opts, popts = _ip([c[4:7] for c in calls])
        assert pb.lib.pc_interpolate_batch(pb.h, len(calls), psize, pdoff, psrc, pdst,
                                           popts) == 1
        # restriction into a patch that is not virtual
        src, psrc = _ip([children[0][1]])
        dst, pdst = _ip([parents[0]])
        assert pb.lib.pc_restrict_batch(pb.h, 1, psrc, pdst, 0) == 1
        for s, (u, o, b) in before.items():
            x = pb.download(s)
            assert np.array_equal(x[0], u) and np.array_equal(x[1], o) and np.array_equal(x[2], b)
    finally:
        pa.close()
        pb.close()
