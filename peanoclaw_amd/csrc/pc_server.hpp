// Continuous-batching server of independent shallow-water scenarios over one
// patch pool (included at the end of pc_abi.hip; uses the pool internals).
//
// A request is a dam break on [0,1]^2 with its own grid size, bathymetry
// and end time.  Requests wait in arrival order; each round admits every
// waiting request that fits the pool's HBM budget and the active limit
// (first fit in arrival order), registers its N/n x N/n patches in one
// contiguous tile range, writes its initial state on the device, then
// advances all active requests by `steps` shared batched steps (each
// scenario with its own CFL dt; replayed from a captured graph while the
// set of requests is unchanged), and retires the requests that reached
// their end time: digest of the final state, mass, step count, patches
// released.  The reference has no counterpart (it runs one simulation per
// process); the per-step numerics are those of the batched pool step.
#pragma once

#include <deque>

static_assert(PC_MAX_HILLS == kMaxHills, "hill count of the C ABI and the kernels");

namespace {

struct ServerReq {
  pc_request req{};
  int status = 0;  // 0 queued, 1 active, 2 done (patches kept), 3 done (released)
  int scenario = -1;
  int P = 0;
  long long tile_cells = 0;
  long long admitted_round = -1;
  std::vector<int> slots;  // grid order, k = i * P + j
  pc_request_result res{};
};

}  // namespace

struct pc_server {
  pc_pool* pool = nullptr;
  int max_active = 0;
  int keep = 0;
  long long round = 0;
  std::vector<ServerReq> reqs;  // indexed by request id
  std::deque<int> queue;
  std::vector<int> active;
  std::vector<int> done;  // completed, not yet polled
  std::vector<char> scn_used;
  int* d_slots = nullptr;
  size_t d_slots_cap = 0;
  unsigned long long* d_hash = nullptr;
};

namespace {

long long live_tile_cells(pc_pool* p) {
  long long c = 0;
  for (const Slot& s : p->slots)
    if (s.live) c += s.cells;
  return c;
}

int server_validate(const pc_request* r) {
  if (!r) return fail(PC_E_INVALID, "request is NULL");
  const int n = r->patch_cells, N = r->cells_per_side;
  if (n < 2 || n > 64) return fail(PC_E_INVALID, "patch_cells must be in [2, 64]");
  if (N < n || N % n != 0) return fail(PC_E_INVALID, "cells_per_side must be a multiple of patch_cells");
  if (N / n > 4096) return fail(PC_E_INVALID, "too many patches per side");
  if (r->num_hills < 0 || r->num_hills > PC_MAX_HILLS)
    return fail(PC_E_INVALID, "num_hills out of range");
  if (!(r->end_time >= 0.0) || !(r->depth >= 0.0) || !(r->column_height >= 0.0) ||
      !(r->column_radius >= 0.0))
    return fail(PC_E_INVALID, "negative or non-finite request parameter");
  for (int m = 0; m < r->num_hills; ++m)
    if (!(r->hill_sigma[m] > 0.0)) return fail(PC_E_INVALID, "hill_sigma must be positive");
  return PC_OK;
}

int server_scenario(pc_server* s) {
  for (size_t k = 0; k < s->scn_used.size(); ++k)
    if (!s->scn_used[k]) {
      s->scn_used[k] = 1;
      return (int)k;
    }
  s->scn_used.push_back(1);
  return (int)s->scn_used.size() - 1;
}

int server_upload_slots(pc_server* s, const std::vector<int>& slots) {
  pc_pool* p = s->pool;
  CHK(ensure_dev(&s->d_slots, &s->d_slots_cap, std::max<size_t>(slots.size(), 1)));
  HIPCHK(hipMemcpyAsync(s->d_slots, slots.data(), slots.size() * sizeof(int),
                        hipMemcpyHostToDevice, p->stream));
  return PC_OK;
}

// Register the P x P patches of a request in one contiguous tile range.
int server_register(pc_server* s, ServerReq& q) {
  pc_pool* p = s->pool;
  const int n = q.req.patch_cells, P = q.P;
  const long long per = (long long)(n + 2) * tile_stride(n);
  const long long total = per * P * P;
  long long off = 0;
  // per is a multiple of kTileAlign for padded rows: every tile keeps the phase
  CHK(alloc_cells(p, total, &off, tile_phase(n)));
  // the other state buffer starts zeroed like a fresh heap vector; the
  // current one and b are written by k_init_request
  if (off < p->hwm) {
    const int cur = (int)(p->scn.size() > (size_t)q.scenario ? p->scn[q.scenario].stepc & 1 : 0);
    for (int k = 0; k < 3; ++k)
      HIPCHK(hipMemsetAsync(p->f[2 * k + (1 - cur)] + off, 0, (size_t)total * sizeof(double),
                            p->stream));
  }
  p->hwm = std::max(p->hwm, off + total);
  std::vector<int> free_slots;
  for (size_t i = 0; i < p->slots.size() && free_slots.size() < (size_t)P * P; ++i)
    if (!p->slots[i].live) free_slots.push_back((int)i);
  while (free_slots.size() < (size_t)P * P) {
    free_slots.push_back((int)p->slots.size());
    p->slots.emplace_back();
  }
  const double dx = 1.0 / q.req.cells_per_side;
  q.slots.assign((size_t)P * P, -1);
  for (int i = 0; i < P; ++i)
    for (int j = 0; j < P; ++j) {
      const int k = i * P + j;
      Slot sl;
      sl.live = true;
      sl.off = off + per * k;
      sl.cells = per;
      sl.stride = tile_stride(n);
      sl.clock_from_scenario = true;
      pc_cell_description& d = sl.desc;
      d.position[0] = i * n * dx;
      d.position[1] = j * n * dx;
      d.size[0] = d.size[1] = n * dx;
      d.subdivision_factor[0] = d.subdivision_factor[1] = n;
      d.ghostlayer_width = 1;
      d.level = 0;
      d.is_virtual = 0;
      d.scenario = q.scenario;
      d.minimal_neighbor_time_constraint = 1.7976931348623157e308;
      d.restriction_lower_bounds[0] = d.restriction_lower_bounds[1] = 1.7976931348623157e308;
      d.restriction_upper_bounds[0] = d.restriction_upper_bounds[1] = -1.7976931348623157e308;
      p->slots[free_slots[k]] = sl;
      q.slots[k] = free_slots[k];
    }
  p->topo_dirty = true;
  return PC_OK;
}

void server_release_slots(pc_server* s, ServerReq& q) {
  pc_pool* p = s->pool;
  for (int sl : q.slots) {
    if (sl < 0 || sl >= (int)p->slots.size() || !p->slots[sl].live) continue;
    free_cells(p, p->slots[sl].off, p->slots[sl].cells);
    p->slots[sl].live = false;
    if ((size_t)sl < p->h_sides.size()) p->h_sides[sl] = SideDev{};
  }
  p->topo_dirty = true;
  q.slots.clear();
  if (q.scenario >= 0 && (size_t)q.scenario < s->scn_used.size()) s->scn_used[q.scenario] = 0;
}

int server_admit(pc_server* s, int id) {
  pc_pool* p = s->pool;
  ServerReq& q = s->reqs[id];
  CHK(sync_scn_host(p));
  q.scenario = server_scenario(s);
  if ((int)p->scn.size() <= q.scenario) {
    const size_t old = p->scn.size();
    p->scn.resize(q.scenario + 1);
    for (size_t k = old; k < p->scn.size(); ++k) {
      ScnDev z{};
      z.end_time = INFINITY;
      z.minbits = 0x7FEFFFFFFFFFFFFFull;
      p->scn[k] = z;
    }
  }
  int rc = server_register(s, q);
  if (rc != PC_OK) {
    server_release_slots(s, q);
    q.scenario = -1;
    return rc;
  }
  // fused ghost sources of the request's grid (walls at the box)
  rc = pc_set_grid_sides(p, q.P, q.P, q.slots.data());
  if (rc == PC_OK) rc = prepare(p);  // topology, per-slot device arrays
  if (rc != PC_OK) {
    server_release_slots(s, q);
    q.scenario = -1;
    return rc;
  }
  // a fresh clock; estimate 0 = no pruned step before an exact one
  ScnDev& S = p->scn[q.scenario];
  const int stepc = S.stepc, pbeg = S.pbeg, pend = S.pend;
  S = ScnDev{};
  S.stepc = stepc;
  S.pbeg = pbeg;
  S.pend = pend;
  S.end_time = q.req.end_time;
  S.minbits = 0x7FEFFFFFFFFFFFFFull;
  CHK(upload_scn(p));
  // initial state and edge buffers of the new patches
  RequestInit a{};
  a.dx = 1.0 / q.req.cells_per_side;
  a.cx = q.req.column_x;
  a.cy = q.req.column_y;
  a.radius = q.req.column_radius;
  a.col_h = q.req.column_height;
  a.depth = q.req.depth;
  a.nh = q.req.num_hills;
  a.P = q.P;
  a.n = q.req.patch_cells;
  a.cur = stepc & 1;
  for (int m = 0; m < a.nh; ++m) {
    a.hx[m] = q.req.hill_x[m];
    a.hy[m] = q.req.hill_y[m];
    a.hs2[m] = q.req.hill_sigma[m] * q.req.hill_sigma[m];
    a.ha[m] = q.req.hill_height[m];
  }
  CHK(server_upload_slots(s, q.slots));
  const int ns = (int)q.slots.size();
  hipLaunchKernelGGL(k_init_request, dim3(ns), dim3(256), 0, p->stream, s->d_slots, ns,
                     p->d_pd, fields_of(p), a);
  HIPCHK(hipGetLastError());
  if (!p->edges_dirty) {
    hipLaunchKernelGGL(k_edges_init, dim3(ns), dim3(64), 0, p->stream, s->d_slots, ns, p->d_pd,
                       fields_of(p), p->d_scn, edges_of(p), p->d_bcol);
    HIPCHK(hipGetLastError());
  }
  if (fused_mode(p) != 2) p->fused_ready = false;  // exact modes recompute all speeds
  HIPCHK(hipStreamSynchronize(p->stream));
  q.status = 1;
  q.admitted_round = s->round;
  s->active.push_back(id);
  return PC_OK;
}

int server_finish(pc_server* s, int id) {
  pc_pool* p = s->pool;
  ServerReq& q = s->reqs[id];
  const ScnDev& S = p->scn[q.scenario];
  pc_request_result& r = q.res;
  r.id = id;
  r.scenario = q.scenario;
  r.steps = S.steps;
  r.cell_updates = S.steps * (long long)q.req.cells_per_side * q.req.cells_per_side;
  r.time = S.time + S.dt;
  r.admitted_round = q.admitted_round;
  r.completed_round = s->round;
  CHK(pc_scenario_mass(p, q.scenario, &r.mass));
  CHK(server_upload_slots(s, q.slots));
  HIPCHK(hipMemsetAsync(s->d_hash, 0, sizeof(unsigned long long), p->stream));
  const int ns = (int)q.slots.size();
  hipLaunchKernelGGL(k_hash_request, dim3(ns), dim3(256), 0, p->stream, s->d_slots, ns,
                     p->d_pd, fields_of(p), S.stepc & 1, s->d_hash);
  HIPCHK(hipGetLastError());
  unsigned long long hsh = 0;
  HIPCHK(hipMemcpyAsync(&hsh, s->d_hash, sizeof(hsh), hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  r.state_hash = hsh;
  if (s->keep) {
    q.status = 2;
  } else {
    server_release_slots(s, q);
    q.status = 3;
  }
  s->done.push_back(id);
  return PC_OK;
}

}  // namespace

extern "C" {

int pc_server_create(pc_pool* pool, int max_active, int keep_completed, pc_server** out) {
  if (!pool || !out || max_active < 1) return fail(PC_E_INVALID, "invalid argument");
  hipSetDevice(pool->device);
  pc_server* s = new pc_server;
  s->pool = pool;
  s->max_active = max_active;
  s->keep = keep_completed ? 1 : 0;
  if (hipMalloc(&s->d_hash, sizeof(unsigned long long)) != hipSuccess) {
    delete s;
    return fail(PC_E_HIP, "hipMalloc failed");
  }
  *out = s;
  return PC_OK;
}

int pc_server_destroy(pc_server* s) {
  if (!s) return fail(PC_E_INVALID, "server is NULL");
  hipSetDevice(s->pool->device);
  hipStreamSynchronize(s->pool->stream);
  for (ServerReq& q : s->reqs)
    if (!q.slots.empty()) server_release_slots(s, q);
  if (s->d_slots) hipFree(s->d_slots);
  if (s->d_hash) hipFree(s->d_hash);
  delete s;
  return PC_OK;
}

int pc_server_submit(pc_server* s, const pc_request* req, int* out_id) {
  if (!s || !out_id) return fail(PC_E_INVALID, "invalid argument");
  CHK(server_validate(req));
  ServerReq q;
  q.req = *req;
  q.P = req->cells_per_side / req->patch_cells;
  q.tile_cells = (long long)(req->patch_cells + 2) * tile_stride(req->patch_cells) * q.P * q.P;
  if ((size_t)q.tile_cells * kBytesPerCell > s->pool->budget)
    return fail(PC_E_OOM, "request exceeds the pool's HBM budget");
  *out_id = (int)s->reqs.size();
  s->reqs.push_back(q);
  s->queue.push_back(*out_id);
  return PC_OK;
}

int pc_server_round(pc_server* s, int steps, int* active_out, int* queued_out) {
  if (!s || steps < 0) return fail(PC_E_INVALID, "invalid argument");
  pc_pool* p = s->pool;
  hipSetDevice(p->device);
  // admission: first fit in arrival order under the budget and active limit
  const long long budget_cells = (long long)(p->budget / kBytesPerCell);
  long long used = live_tile_cells(p);
  for (auto it = s->queue.begin(); it != s->queue.end();) {
    if ((int)s->active.size() >= s->max_active) break;
    ServerReq& q = s->reqs[*it];
    if (used + q.tile_cells > budget_cells) {
      ++it;
      continue;
    }
    const int rc = server_admit(s, *it);
    if (rc == PC_E_OOM) {  // fragmentation: leave it queued
      ++it;
      continue;
    }
    CHK(rc);
    used += q.tile_cells;
    it = s->queue.erase(it);
  }
  if (!s->active.empty() && steps > 0) {
    // Between admissions and completions the step is the same launch
    // sequence: replay it as a captured graph (state read on the device);
    // after a change, one plain step applies it and the graph is captured
    // again.
    const bool replay = !p->topo_dirty && p->graph_valid && !p->scn_dev_dirty &&
                        !(p->sides_any && p->edges_dirty) &&
                        !(fused_mode(p) && p->sides_any && !p->fused_ready);
    if (replay) {
      CHK(pc_graph_replay(p, steps));
    } else {
      CHK(pc_step_all(p, 1));
      if (steps > 1) {
        CHK(pc_graph_capture(p));
        CHK(pc_graph_replay(p, steps - 1));
      }
    }
  }
  // retire the requests that reached their end time
  CHK(sync_scn_host(p));
  std::vector<int> still;
  for (int id : s->active) {
    const ScnDev& S = p->scn[s->reqs[id].scenario];
    const double maxDt = S.end_time - (S.time + S.dt);
    if (!(maxDt - 0.0 > kLaEps)) {
      CHK(server_finish(s, id));
    } else {
      still.push_back(id);
    }
  }
  s->active.swap(still);
  ++s->round;
  if (active_out) *active_out = (int)s->active.size();
  if (queued_out) *queued_out = (int)s->queue.size();
  return PC_OK;
}

int pc_server_poll(pc_server* s, pc_request_result* out, int cap, int* n) {
  if (!s || !n || cap < 0 || (cap > 0 && !out)) return fail(PC_E_INVALID, "invalid argument");
  const int k = std::min(cap, (int)s->done.size());
  for (int i = 0; i < k; ++i) out[i] = s->reqs[s->done[i]].res;
  s->done.erase(s->done.begin(), s->done.begin() + k);
  *n = k;
  return PC_OK;
}

int pc_server_request_slots(pc_server* s, int id, int* slots, int cap, int* n) {
  if (!s || !n || id < 0 || id >= (int)s->reqs.size())
    return fail(PC_E_INVALID, "invalid request id");
  const ServerReq& q = s->reqs[id];
  *n = (int)q.slots.size();
  if (slots) {
    if (cap < *n) return fail(PC_E_INVALID, "output capacity too small");
    std::copy(q.slots.begin(), q.slots.end(), slots);
  }
  return PC_OK;
}

int pc_server_release(pc_server* s, int id) {
  if (!s || id < 0 || id >= (int)s->reqs.size()) return fail(PC_E_INVALID, "invalid request id");
  ServerReq& q = s->reqs[id];
  if (q.status != 2) return fail(PC_E_INVALID, "request is not a kept completed request");
  server_release_slots(s, q);
  q.status = 3;
  return PC_OK;
}

}  // extern "C"
