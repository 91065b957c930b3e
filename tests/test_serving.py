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


def _run_all(server, reqs, stagger, steps=3):
    """Submit in waves (`stagger` rounds apart) and run to completion,
    `steps` batched steps per round."""
    ids = {}
    pending = list(range(len(reqs)))
    res = {}
    rounds = 0
    while len(res) < len(reqs):
        if pending and rounds % stagger == 0:
            for k in pending[:2]:
                ids[server.submit(reqs[k])] = k
            pending = pending[2:]
        server.round(steps)
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


def _oracle_run_of_request(pool, srv, req, threads=4):
    """Admit `req`, seed the oracle with its initial state, run the oracle
    on the same grid to the end time (global dt bounded by it), then serve
    the request; returns (result, steps, t + dt, worst rel diff, bitwise
    count, cell count)."""
    P = req.cells_per_side // req.patch_cells
    n = req.patch_cells
    rid = srv.submit(req)
    srv.round(0)  # admission only
    slots = srv.slots(rid)
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
    off = np.array([k * s * s for k in range(P * P)], dtype=np.int64)
    nx = np.full(P * P, n, dtype=np.int32)
    dxv = np.array([pool.description(sl).size[0] / n for sl in slots])
    pl = OrPool()
    pl.np = P * P
    pl.nx = pl.ny = nx.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    pl.off = off.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))
    pl.dx = pl.dy = dxv.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
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
        d = olib().or_pool_step(ctypes.byref(pl), max_dt, threads)
        if d - 0.0 > LA_EPS:
            t, dt = t + dt, d
        steps += 1
    res = None
    for _ in range(100000):
        srv.round(64)
        out = srv.poll()
        if out:
            res = out[0]
            break
    assert res is not None
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
    srv.release(rid)
    return res, steps, t + dt, worst, exact, 3 * P * P * n * n


def test_served_request_matches_oracle():
    from peanoclaw_amd.pool import PatchPool
    from peanoclaw_amd.serving import Server, make_request
    req = make_request(64, 32, end_time=0.03, column=(0.35, 0.6), column_radius=0.12,
                       column_height=2.5, depth=1.0,
                       hills=[(0.3, 0.4, 0.1, 0.25), (0.7, 0.6, 0.08, 0.2)])
    with PatchPool(0) as pool, Server(pool, max_active=4, keep_completed=True) as srv:
        pool.launch_config(persistent_workgroups=512, fused=2)
        res, steps, tend, worst, exact, total = _oracle_run_of_request(pool, srv, req)
        print(f"served vs oracle: {steps} steps, max rel {worst:.3e}, bitwise {exact}/{total}")
        assert res.steps == steps
        assert res.time == tend
        assert worst <= 1e-12


def test_served_256_request_to_t1_matches_oracle():
    """A 256^2 request of the served workload's kind (8 x 8 patches of 32^2,
    five Gaussian hills) run to t = 1.0 (about 4,500 steps): bitwise equal
    to the oracle's run of the same initial state (the bar is 1e-10)."""
    from peanoclaw_amd.pool import PatchPool
    from peanoclaw_amd.serving import Server, make_request
    rng = np.random.default_rng(256)
    hills = [(rng.uniform(0, 1), rng.uniform(0, 1), rng.uniform(0.05, 0.15),
              rng.uniform(0.0, 0.3)) for _ in range(5)]
    req = make_request(256, 32, end_time=1.0, column=(0.42, 0.57), column_radius=0.1,
                       column_height=2.6, depth=1.0, hills=hills)
    with PatchPool(0) as pool, Server(pool, max_active=4, keep_completed=True) as srv:
        pool.launch_config(fused=2)
        res, steps, tend, worst, exact, total = _oracle_run_of_request(pool, srv, req,
                                                                       threads=16)
        print(f"256^2 to t=1: {steps} steps, max rel {worst:.3e}, bitwise {exact}/{total}",
              flush=True)
        assert res.steps == steps
        assert res.time == tend
        assert abs(tend - 1.0) <= LA_EPS
        assert worst <= 1e-10


def test_request_initial_state_is_the_breaking_dam():
    """A request with the dam break's centre, radius and heights and no
    hills starts every patch from what the oracle's restatement of
    BreakingDamSWEScenario::initializePatch writes for its geometry."""
    from test_oracle_initial_data import breaking_dam
    from peanoclaw_amd.pool import PatchPool
    from peanoclaw_amd.serving import Server, make_request
    req = make_request(288, 32, end_time=1.0, column=(0.5, 0.5), column_radius=0.25,
                       column_height=2.0, depth=1.0)
    with PatchPool(0) as pool, Server(pool, max_active=1, keep_completed=True) as srv:
        rid = srv.submit(req)
        srv.round(0)
        slots = srv.slots(rid)
        assert len(slots) == 81
        for sl in slots:
            d = pool.description(sl)
            unew, _, b = pool.download(sl)
            ref, _ = breaking_dam((d.position[0], d.position[1]), d.size[0], 32)
            assert np.array_equal(unew, ref), sl
            assert not b.any()


def test_stream_requests_up_to_2048_alone_and_in_mix():
    """Requests of the served-stream workload (served_requests, seed 5) at
    their stated sizes and run to t = 1.0: the first five (64^2 .. 1120^2)
    and the first 2048^2 one.  Served one at a time and as a Poisson
    stream under a 512 MiB budget (admissions while others run), each
    result is the same: steps, end time, state digest and mass."""
    from peanoclaw_amd.pool import PatchPool
    from peanoclaw_amd.serving import Server, run_stream, served_requests
    allr = served_requests(204, seed=5)
    reqs = [allr[k] for k in (0, 1, 2, 3, 4, 203)]
    sizes = [r.cells_per_side for r in reqs]
    assert max(sizes) == 2048 and sum(n >= 1024 for n in sizes) >= 3
    with PatchPool(0) as pool:
        pool.launch_config(fused=2)
        with Server(pool, max_active=1) as srv:
            alone = _run_all(srv, reqs, 1, steps=64)
    print("alone done: " + ", ".join(f"{n}^2 {alone[k].steps} steps"
                                     for k, n in enumerate(sizes)), flush=True)
    with PatchPool(0, hbm_budget_bytes=512 << 20) as pool:
        pool.launch_config(fused=2)
        with Server(pool, max_active=64) as srv:
            res, lat, wall, cells = run_stream(srv, reqs, rate=200.0, seed=5,
                                               steps_per_round=64)
    for k in range(len(reqs)):
        a, m = alone[k], res[k]
        print(f"request {k}: N={sizes[k]} steps {a.steps}/{m.steps} "
              f"hash {a.state_hash:016x}/{m.state_hash:016x} latency {lat[k]:.2f} s")
        assert a.steps == m.steps
        assert a.time == m.time
        assert abs(a.time - 1.0) <= LA_EPS
        assert a.state_hash == m.state_hash
        assert a.mass == m.mass
    assert any(res[k].admitted_round > 0 for k in res)


def test_admission_waits_for_budget():
    from peanoclaw_amd.pool import PatchPool
    from peanoclaw_amd.serving import Server, make_request
    from peanoclaw_amd._abi import PeanoClawAmdError
    from peanoclaw_amd.pool import tile_stride
    tile_cells = 64 * 34 * tile_stride(32)  # N = 256 in 32^2 patches (padded rows)
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
