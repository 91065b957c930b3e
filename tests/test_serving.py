"""Served stream (continuous batching) through the C ABI.

* each request's result is bit-identical alone and in a changing mix;
* a served request matches the oracle run of its initial state;
* admission respects the pool's HBM budget and the active limit.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle_lib import OrPool, dptr as odptr, lib as olib

pytestmark = pytest.mark.gpu

LA_EPS = 1e-8


def _requests():
    from peanoclaw_amd.serving import make_request
    rng = np.random.default_rng(11)
    out = []
    for n in (64, 96, 128, 64, 160, 96):
        hills = [(rng.uniform(0, 1), rng.uniform(0, 1), rng.uniform(0.05, 0.15),
                  rng.uniform(0, 0.3)) for _ in range(5)]
        out.append(make_request(n, 32, end_time=float(rng.uniform(0.02, 0.06)),
                                column=rng.uniform(0.2, 0.8, 2), column_radius=0.1,
                                column_height=float(rng.uniform(1.2, 3.0)), depth=1.0,
                                hills=hills))
    return out


def _run_all(server, reqs, stagger):
    """Submit in waves (`stagger` rounds apart) and run to completion."""
    ids = {}
    pending = list(range(len(reqs)))
    res = {}
    rounds = 0
    while len(res) < len(reqs):
        if pending and rounds % stagger == 0:
            for k in pending[:2]:
                ids[server.submit(reqs[k])] = k
            pending = pending[2:]
        server.round(3)
        rounds += 1
        for r in server.poll():
            res[ids[r.id]] = r
        assert rounds < 100000
    return res


def test_served_results_identical_alone_and_in_mix():
    from peanoclaw_amd.pool import PatchPool
    from peanoclaw_amd.serving import Server
    reqs = _requests()
    with PatchPool(0) as pool:
        pool.launch_config(persistent_workgroups=1024, fused=2)
        with Server(pool, max_active=1) as srv:  # one at a time
            alone = _run_all(srv, reqs, 1)
    with PatchPool(0) as pool:
        pool.launch_config(persistent_workgroups=1024, fused=2)
        with Server(pool, max_active=8) as srv:  # mixed, admissions mid-run
            mixed = _run_all(srv, reqs, 4)
    for k in range(len(reqs)):
        a, m = alone[k], mixed[k]
        print(f"request {k}: N={reqs[k].cells_per_side} steps {a.steps}/{m.steps} "
              f"hash {a.state_hash:016x}/{m.state_hash:016x}")
        assert a.steps == m.steps
        assert a.time == m.time
        assert a.state_hash == m.state_hash
        assert a.mass == m.mass
        assert abs(a.time - reqs[k].end_time) <= LA_EPS
    # the mix really overlapped: some request was admitted while others ran
    assert any(mixed[k].admitted_round > 0 for k in mixed)


def test_served_request_matches_oracle():
    from peanoclaw_amd.pool import PatchPool
    from peanoclaw_amd.serving import Server, make_request
    req = make_request(64, 32, end_time=0.03, column=(0.35, 0.6), column_radius=0.12,
                       column_height=2.5, depth=1.0,
                       hills=[(0.3, 0.4, 0.1, 0.25), (0.7, 0.6, 0.08, 0.2)])
    with PatchPool(0) as pool, Server(pool, max_active=4, keep_completed=True) as srv:
        pool.launch_config(persistent_workgroups=512, fused=2)
        rid = srv.submit(req)
        srv.round(0)  # admission only
        slots = srv.slots(rid)
        P, n = 2, 32
        s = n + 2
        H = [np.zeros(P * P * s * s), np.zeros(P * P * s * s)]
        HU = [np.zeros_like(H[0]), np.zeros_like(H[0])]
        HV = [np.zeros_like(H[0]), np.zeros_like(H[0])]
        B = np.zeros_like(H[0])
        for k, sl in enumerate(slots):
            unew, _, b = pool.download(sl)
            for arr, q in ((H[0], 0), (HU[0], 1), (HV[0], 2)):
                arr[k * s * s:(k + 1) * s * s].reshape(s, s)[1:-1, 1:-1] = unew[q]
            B[k * s * s:(k + 1) * s * s] = b.ravel()
        # the oracle: same grid, global dt bounded by the end time
        off = np.array([k * s * s for k in range(P * P)], dtype=np.int64)
        nx = np.full(P * P, n, dtype=np.int32)
        dx = np.full(P * P, 1.0 / 64)
        pl = OrPool()
        pl.np = P * P
        pl.nx = pl.ny = nx.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        pl.off = off.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))
        pl.dx = pl.dy = dx.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        for k in range(2):
            pl.h[k], pl.hu[k], pl.hv[k] = odptr(H[k]), odptr(HU[k]), odptr(HV[k])
        pl.b = odptr(B)
        pl.cur = 0
        so = np.arange(P * P, dtype=np.int32)
        t, dt, steps = 0.0, 0.0, 0
        while True:
            max_dt = req.end_time - (t + dt)
            if not (max_dt - 0.0 > LA_EPS):
                break
            c = pl.cur
            olib().or_grid_fill(P, P, n, so.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                off.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                odptr(H[c]), odptr(HU[c]), odptr(HV[c]))
            d = olib().or_pool_step(ctypes.byref(pl), max_dt, 4)
            if d - 0.0 > LA_EPS:
                t, dt = t + dt, d
            steps += 1
        # the server
        res = None
        for _ in range(10000):
            srv.round(5)
            out = srv.poll()
            if out:
                res = out[0]
                break
        assert res is not None
        assert res.steps == steps
        assert res.time == t + dt
        worst = 0.0
        exact = 0
        for k, sl in enumerate(slots):
            unew, _, _ = pool.download(sl)
            c = pl.cur
            for arr, q in ((H[c], 0), (HU[c], 1), (HV[c], 2)):
                ref = arr[k * s * s:(k + 1) * s * s].reshape(s, s)[1:-1, 1:-1]
                d = np.abs(unew[q] - ref)
                worst = max(worst, float((d / np.maximum(np.abs(ref), 1.0)).max()))
                exact += int((unew[q].view(np.uint64) == ref.view(np.uint64)).sum())
        print(f"served vs oracle: {steps} steps, max rel {worst:.3e}, "
              f"bitwise {exact}/{3 * P * P * n * n}")
        assert worst <= 1e-12
        srv.release(rid)


def test_admission_waits_for_budget():
    from peanoclaw_amd.pool import PatchPool
    from peanoclaw_amd.serving import Server, make_request
    from peanoclaw_amd._abi import PeanoClawAmdError
    tile_cells = 64 * 34 * 34  # N = 256 in 32^2 patches
    budget = int(tile_cells * 56 * 1.5)  # room for one such request, not two
    reqs = [make_request(256, 32, end_time=0.004, column=(0.5, 0.5), column_height=2.0)
            for _ in range(2)]
    with PatchPool(0, hbm_budget_bytes=budget) as pool, Server(pool, max_active=8) as srv:
        pool.launch_config(persistent_workgroups=512, fused=2)
        with pytest.raises(PeanoClawAmdError):
            srv.submit(make_request(512, 32))  # can never fit
        a = srv.submit(reqs[0])
        b = srv.submit(reqs[1])
        active, queued = srv.round(0)
        assert (active, queued) == (1, 1)
        res = {}
        for _ in range(10000):
            srv.round(4)
            for r in srv.poll():
                res[r.id] = r
            if len(res) == 2:
                break
        assert res[b].admitted_round >= res[a].completed_round
        assert res[a].state_hash == res[b].state_hash  # same request, same bits
