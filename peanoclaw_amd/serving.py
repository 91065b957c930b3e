"""Served stream of independent dam-break scenarios (continuous batching).

`Server` wraps the library's request server (pc_server_* in the C ABI):
requests queue in arrival order, each round admits those that fit the
pool's HBM budget, advances every admitted scenario by a few shared batched
steps (each with its own CFL dt) and retires the ones that reached their
end time.  `served_requests` draws the requests of the served-stream
workload (seeded); `run_stream` replays them with Poisson arrivals against
the wall clock and reports time-to-solution percentiles and throughput.
"""
import ctypes
import math
import time

import numpy as np

from ._abi import MAX_HILLS, Request, RequestResult, check, load


def make_request(cells_per_side, patch_cells=32, end_time=1.0, column=(0.5, 0.5),
                 column_radius=0.1, column_height=2.0, depth=1.0, hills=()):
    """hills: iterable of (x, y, sigma, height)."""
    r = Request()
    r.cells_per_side = int(cells_per_side)
    r.patch_cells = int(patch_cells)
    r.end_time = float(end_time)
    r.column_x, r.column_y = float(column[0]), float(column[1])
    r.column_radius = float(column_radius)
    r.column_height = float(column_height)
    r.depth = float(depth)
    hills = list(hills)
    if len(hills) > MAX_HILLS:
        raise ValueError("at most %d hills" % MAX_HILLS)
    r.num_hills = len(hills)
    for k, (x, y, s, a) in enumerate(hills):
        r.hill_x[k], r.hill_y[k], r.hill_sigma[k], r.hill_height[k] = x, y, s, a
    return r


def served_requests(count, seed=5, n_min=64, n_max=2048, patch_cells=32, end_time=1.0):
    """Requests of the served-stream workload: grid size log-uniform in
    [n_min, n_max] cells per side rounded to a multiple of the 32-cell patch,
    column centre U[0.2,0.8]^2 with radius 0.1 and height U[1.2,3.0] over
    depth 1, five Gaussian hills (centre U[0,1]^2, sigma U[0.05,0.15],
    height U[0,0.3]), run to t = end_time."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        n = int(min(n_max, max(n_min, patch_cells * round(n / patch_cells))))
        c = rng.uniform(0.2, 0.8, 2)
        h = rng.uniform(1.2, 3.0)
        hills = [(rng.uniform(0, 1), rng.uniform(0, 1), rng.uniform(0.05, 0.15),
                  rng.uniform(0.0, 0.3)) for _ in range(5)]
        out.append(make_request(n, patch_cells, end_time, c, 0.1, h, 1.0, hills))
    return out


class Server:
    def __init__(self, pool, max_active=64, keep_completed=False):
        self.lib = load()
        self.pool = pool
        h = ctypes.c_void_p()
        check(self.lib.pc_server_create(pool.h, int(max_active), 1 if keep_completed else 0,
                                        ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            check(self.lib.pc_server_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, req):
        i = ctypes.c_int()
        check(self.lib.pc_server_submit(self.h, ctypes.byref(req), ctypes.byref(i)))
        return i.value

    def round(self, steps):
        a, q = ctypes.c_int(), ctypes.c_int()
        check(self.lib.pc_server_round(self.h, int(steps), ctypes.byref(a), ctypes.byref(q)))
        return a.value, q.value

    def poll(self):
        out = (RequestResult * 64)()
        res = []
        while True:
            n = ctypes.c_int()
            check(self.lib.pc_server_poll(self.h, out, 64, ctypes.byref(n)))
            res.extend(RequestResult.from_buffer_copy(out[k]) for k in range(n.value))
            if n.value < 64:
                return res

    def slots(self, rid):
        n = ctypes.c_int()
        check(self.lib.pc_server_request_slots(self.h, rid, None, 0, ctypes.byref(n)))
        arr = (ctypes.c_int * max(n.value, 1))()
        check(self.lib.pc_server_request_slots(self.h, rid, arr, n.value, ctypes.byref(n)))
        return list(arr[:n.value])

    def release(self, rid):
        check(self.lib.pc_server_release(self.h, rid))


def run_stream(server, requests, rate=200.0, seed=5, steps_per_round=8, clock=time.perf_counter):
    """Submit `requests` at Poisson arrival times (mean `rate` per second,
    seeded) measured on `clock`, run rounds until all are done.  Returns
    (results by request index, per-request time-to-solution in seconds,
    wall seconds, total cell updates)."""
    rng = np.random.default_rng(seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(requests)))
    ids, finish = {}, {}
    results = {}
    nxt = 0
    t0 = clock()
    while len(results) < len(requests):
        now = clock() - t0
        while nxt < len(requests) and arrivals[nxt] <= now:
            ids[server.submit(requests[nxt])] = nxt
            nxt += 1
        if nxt == len(results):  # nothing submitted is unfinished
            time.sleep(max(0.0, arrivals[nxt] - now))  # idle until the next arrival
            continue
        server.round(steps_per_round)
        t = clock() - t0
        for r in server.poll():
            k = ids[r.id]
            results[k] = r
            finish[k] = t
    wall = clock() - t0
    lat = np.array([finish[k] - arrivals[k] for k in range(len(requests))])
    cells = sum(results[k].cell_updates for k in range(len(requests)))
    return results, lat, wall, cells


__all__ = ["Server", "make_request", "served_requests", "run_stream"]
