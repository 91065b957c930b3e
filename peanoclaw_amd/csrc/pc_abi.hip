// C ABI of the MI355X SWE patch path (declared in include/peanoclaw_amd.h).
//
// Host side of the patch pool: HBM allocation of the structure-of-arrays
// fields, slot registry mirroring CellDescription, scenario clocks, the
// ghost-strip table, the work list of the batched update, hipGraph capture
// of one step, and the per-patch Numerics-mirroring calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/peanoclaw_amd.h"
#include "pc_kernels.hpp"
#include "pc_patch_kernels.hpp"

using namespace pcamd;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                            \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess)                                                    \
      return fail(PC_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CHK(x)              \
  do {                      \
    int r_ = (x);           \
    if (r_ != PC_OK) return r_; \
  } while (0)

constexpr int kNumFields = 7;  // h0 h1 hu0 hu1 hv0 hv1 b
constexpr size_t kBytesPerCell = kNumFields * sizeof(double);
constexpr int kFallbackWgs = 2048;  // grid of the (usually idle) exact fallback pass

struct Slot {
  bool live = false;
  pc_cell_description desc{};
  long long off = 0;
  long long cells = 0;  // (nx + 2) * stride
  int stride = 0;       // row stride of the tile (tile_stride(ny))
  bool clock_from_scenario = false;
};

bool la_greater(double a, double b) { return a - b > kLaEps; }
bool la_smaller(double a, double b) { return b - a > kLaEps; }

}  // namespace

struct pc_pool {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t budget = 0;
  long long cap_cells = 0;
  long long top = 0;
  std::vector<std::pair<long long, long long>> free_ranges;
  double* f[kNumFields] = {};
  std::vector<Slot> slots;
  std::vector<ScnDev> scn;  // host mirror, indexed by scenario id
  bool scn_host_stale = false;
  bool scn_dev_dirty = true;
  bool topo_dirty = true;
  // device arrays
  PatchDev* d_pd = nullptr;
  size_t pd_cap = 0;
  double* d_smax = nullptr;
  double* d_smax_int = nullptr;
  double* d_est = nullptr;
  double* d_ebuf[2] = {nullptr, nullptr};
  double* d_bcol = nullptr;
  SideDev* d_sides = nullptr;
  std::vector<SideDev> h_sides;
  bool sides_any = false;
  bool edges_dirty = true;
  bool fused_ready = false;
  int* d_scn_of = nullptr;
  int nentries = 0;
  ScnDev* d_scn = nullptr;
  size_t scn_cap = 0;
  Task* d_tasks = nullptr;
  int ntasks = 0;
  std::vector<Task> h_tasks;
  // the update's work list: h_tasks, with the rows of every patch split into
  // row segments when the step has fewer tasks than the GPU holds
  Task* d_utasks = nullptr;
  int nutasks = 0;
  int row_segments = 0;  // 0 automatic, 1 off, 2 or 4 forced
  std::vector<int> h_rseg;  // row segments per slot of the current list
  int* d_plist = nullptr;
  int* d_plist_scn = nullptr;
  size_t plist_cap = 0;
  std::vector<int> h_plist;
  std::vector<int> h_live;  // live slots (for flags / sums)
  int* d_live = nullptr;
  BoundaryDesc* d_bdesc = nullptr;  // ghost-edge sources of the live slots
  size_t bdesc_cap = 0;
  size_t live_cap = 0;
  GhostStrip* d_strips = nullptr;
  int nstrips = 0;
  size_t strips_cap = 0;
  std::vector<GhostStrip> h_strips;
  std::vector<int> strip_begin;  // per slot (for the per-patch form)
  int* d_counter = nullptr;  // [0] persistent task counter, [1] pruned-mode redo flag
  int* d_redo = nullptr;     // = d_counter + 1
  int* d_iota = nullptr;   // d_iota[i] = i for every slot (per-patch kernels' lists)
  double* d_segmax = nullptr;  // row-segment speed maxima of the one-patch kernels
  double* d_keep = nullptr;    // k_patch_solve: uNew of the patch while it is updated
  double* d_result = nullptr;
  // results of the synchronous one-patch calls, written by their kernels
  // straight into pinned host memory (h_result; d_result_host is its device
  // address): no copy back after the kernel
  double* h_result = nullptr;
  double* d_result_host = nullptr;
  unsigned long long call_seq = 0;  // completion words of the per-call kernels
  double* d_aux = nullptr;
  size_t aux_cap = 0;
  int* d_flags = nullptr;
  size_t flags_cap = 0;
  // per-patch task list of pc_step_all_per_patch, kept between calls
  Task* d_single = nullptr;
  size_t single_cap = 0;
  int* d_single_slot = nullptr;
  size_t single_slot_cap = 0;
  // staging of pc_edge_batch (6 input and 5 output doubles per edge)
  double* d_edge_io = nullptr;
  size_t edge_io_cap = 0;
  // flux arrays of the flux correction, kFluxPerSlot doubles per slot
  double* d_flux = nullptr;
  size_t flux_slots = 0;
  // staging of the batched level transfers (pc_*_batch, pc_patches_*)
  char* d_batch = nullptr;
  size_t batch_cap = 0;
  // the table buffer of a destroyed AMR plan, kept for the next
  // pc_amr_create (a regrid replaces the plan)
  char* amr_tables = nullptr;
  size_t amr_tables_cap = 0;
  // pinned host staging for large synchronous uploads (plan tables)
  char* h_stage = nullptr;
  size_t h_stage_cap = 0;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_valid = false;
  // launch config
  int persistent_wgs = 0;
  int fused = 0;
  int task_cap = 0;
  int solver = 0;  // edge solver: 0 f-wave, 1 augmented Riemann
  double prune_slack = kPruneSlackDefault;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  float last_total_ms = 0.f, last_update_ms = 0.f;
  std::vector<hipEvent_t> upd_ev;  // per-step events around the update launch
  long long hwm = 0;               // cells below this may hold old data
};

namespace {

// Next-step speeds are fused into the update for the f-wave only; the
// augmented solver runs the separate exact speed pass.
inline int fused_mode(const pc_pool* p) { return p->solver ? 0 : p->fused; }

// Exact CFL speed pass over `ntasks` tasks (grid `g`), picking the kernel
// instance for the ghost-source mode and the edge solver.
void launch_speed(pc_pool* p, dim3 g, const Task* tasks, int ntasks, const int* plist,
                  Fields F, const SideDev* sides, EdgeBufs E, double* smax,
                  const int* only_if, int smax_stride = 1) {
  if (sides) {
    if (p->solver)
      hipLaunchKernelGGL((k_speed<true, 1>), g, dim3(64), 0, p->stream, tasks, ntasks, plist,
                         p->d_pd, F, p->d_scn, sides, E, smax, only_if,
                         smax_stride);
    else
      hipLaunchKernelGGL((k_speed<true, 0>), g, dim3(64), 0, p->stream, tasks, ntasks, plist,
                         p->d_pd, F, p->d_scn, sides, E, smax, only_if,
                         smax_stride);
  } else {
    if (p->solver)
      hipLaunchKernelGGL((k_speed<false, 1>), g, dim3(64), 0, p->stream, tasks, ntasks, plist,
                         p->d_pd, F, p->d_scn, sides, E, smax, only_if,
                         smax_stride);
    else
      hipLaunchKernelGGL((k_speed<false, 0>), g, dim3(64), 0, p->stream, tasks, ntasks, plist,
                         p->d_pd, F, p->d_scn, sides, E, smax, only_if,
                         smax_stride);
  }
}

// Update without fused next-step speeds.
void launch_update_plain(pc_pool* p, dim3 g, const Task* tasks, int ntasks, const int* plist,
                         Fields F, const SideDev* sides, EdgeBufs E, int pers) {
  if (sides) {
    if (p->solver)
      hipLaunchKernelGGL((k_update<0, true, 1>), g, dim3(64), 0, p->stream, tasks, ntasks,
                         plist, p->d_pd, F, p->d_scn, sides, E, p->d_smax_int, p->d_counter,
                         pers, 1.0);
    else
      hipLaunchKernelGGL((k_update<0, true, 0>), g, dim3(64), 0, p->stream, tasks, ntasks,
                         plist, p->d_pd, F, p->d_scn, sides, E, p->d_smax_int, p->d_counter,
                         pers, 1.0);
  } else {
    if (p->solver)
      hipLaunchKernelGGL((k_update<0, false, 1>), g, dim3(64), 0, p->stream, tasks, ntasks,
                         plist, p->d_pd, F, p->d_scn, sides, E, p->d_smax_int, p->d_counter,
                         pers, 1.0);
    else
      hipLaunchKernelGGL((k_update<0, false, 0>), g, dim3(64), 0, p->stream, tasks, ntasks,
                         plist, p->d_pd, F, p->d_scn, sides, E, p->d_smax_int, p->d_counter,
                         pers, 1.0);
  }
}

Fields fields_of(pc_pool* p) {
  Fields F;
  F.h[0] = p->f[0]; F.h[1] = p->f[1];
  F.hu[0] = p->f[2]; F.hu[1] = p->f[3];
  F.hv[0] = p->f[4]; F.hv[1] = p->f[5];
  F.b = p->f[6];
  return F;
}

template <class T>
int ensure_dev(T** ptr, size_t* cap, size_t need) {
  if (need <= *cap && *ptr) return PC_OK;
  size_t n = std::max<size_t>(need, std::max<size_t>(16, *cap * 2));
  if (*ptr) HIPCHK(hipFree(*ptr));
  *ptr = nullptr;
  HIPCHK(hipMalloc(reinterpret_cast<void**>(ptr), n * sizeof(T)));
  *cap = n;
  return PC_OK;
}

int invalidate_graph(pc_pool* p) {
  if (p->gexec) HIPCHK(hipGraphExecDestroy(p->gexec));
  if (p->graph) HIPCHK(hipGraphDestroy(p->graph));
  p->gexec = nullptr;
  p->graph = nullptr;
  p->graph_valid = false;
  return PC_OK;
}

int grow_fields(pc_pool* p, long long need_cells) {
  if (need_cells <= p->cap_cells) return PC_OK;
  long long ncap = std::max<long long>(need_cells, p->cap_cells * 2);
  ncap = std::max<long long>(ncap, 1 << 16);
  if ((size_t)ncap * kBytesPerCell > p->budget) {
    ncap = need_cells;
    if ((size_t)ncap * kBytesPerCell > p->budget)
      return fail(PC_E_OOM, "patch pool exceeds its HBM budget");
  }
  for (int k = 0; k < kNumFields; ++k) {
    double* nf = nullptr;
    HIPCHK(hipMalloc(&nf, (size_t)ncap * sizeof(double)));
    HIPCHK(hipMemsetAsync(nf, 0, (size_t)ncap * sizeof(double), p->stream));
    if (p->f[k]) {
      HIPCHK(hipMemcpyAsync(nf, p->f[k], (size_t)p->cap_cells * sizeof(double),
                            hipMemcpyDeviceToDevice, p->stream));
      HIPCHK(hipStreamSynchronize(p->stream));
      HIPCHK(hipFree(p->f[k]));
    }
    p->f[k] = nf;
  }
  p->cap_cells = ncap;
  CHK(invalidate_graph(p));
  return PC_OK;
}

// First fit of n cells; with phase >= 0 the offset is phase modulo
// kTileAlign (the part of a free range before it stays free).
int alloc_cells(pc_pool* p, long long n, long long* off, int phase = -1) {
  auto first_at = [&](long long a) {
    return phase < 0 ? a : a + ((phase - a) % kTileAlign + kTileAlign) % kTileAlign;
  };
  for (size_t i = 0; i < p->free_ranges.size(); ++i) {
    const long long a = p->free_ranges[i].first, len = p->free_ranges[i].second;
    const long long st = first_at(a);
    if (st + n > a + len) continue;
    p->free_ranges.erase(p->free_ranges.begin() + i);
    if (st + n < a + len) p->free_ranges.push_back({st + n, a + len - st - n});
    if (st > a) p->free_ranges.push_back({a, st - a});
    std::sort(p->free_ranges.begin(), p->free_ranges.end());
    *off = st;
    return PC_OK;
  }
  const long long st = first_at(p->top);
  CHK(grow_fields(p, st + n));
  if (st > p->top) {
    p->free_ranges.push_back({p->top, st - p->top});
    std::sort(p->free_ranges.begin(), p->free_ranges.end());
  }
  *off = st;
  p->top = st + n;
  return PC_OK;
}

void free_cells(pc_pool* p, long long off, long long n) {
  p->free_ranges.push_back({off, n});
  std::sort(p->free_ranges.begin(), p->free_ranges.end());
  std::vector<std::pair<long long, long long>> m;
  for (auto& r : p->free_ranges) {
    if (!m.empty() && m.back().first + m.back().second == r.first)
      m.back().second += r.second;
    else
      m.push_back(r);
  }
  if (!m.empty() && m.back().first + m.back().second == p->top) {
    p->top = m.back().first;
    m.pop_back();
  }
  p->free_ranges = m;
}

bool valid_slot(pc_pool* p, int s) {
  return p && s >= 0 && s < (int)p->slots.size() && p->slots[s].live;
}

// The state changed outside the batched step: edge buffers and the fused
// interior speed maxima must be rebuilt before the next batched step.
void mark_state_changed(pc_pool* p) {
  p->edges_dirty = true;
  p->fused_ready = false;
}

int sync_scn_host(pc_pool* p) {
  if (!p->scn_host_stale) return PC_OK;
  HIPCHK(hipMemcpyAsync(p->scn.data(), p->d_scn, p->scn.size() * sizeof(ScnDev),
                        hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  p->scn_host_stale = false;
  return PC_OK;
}

int upload_scn(pc_pool* p) {
  CHK(ensure_dev(&p->d_scn, &p->scn_cap, std::max<size_t>(p->scn.size(), 1)));
  if (!p->scn.empty())
    HIPCHK(hipMemcpyAsync(p->d_scn, p->scn.data(), p->scn.size() * sizeof(ScnDev),
                          hipMemcpyHostToDevice, p->stream));
  p->scn_dev_dirty = false;
  return PC_OK;
}

int cur_of(pc_pool* p, int slot) {
  return p->scn[p->slots[slot].desc.scenario].stepc & 1;
}

// Clock of a slot: the scenario clock while the batched path drives it.
void slot_clock(pc_pool* p, int slot, double* t, double* dt) {
  Slot& s = p->slots[slot];
  if (s.clock_from_scenario) {
    const ScnDev& S = p->scn[s.desc.scenario];
    *t = S.time;
    *dt = S.dt;
  } else {
    *t = s.desc.time;
    *dt = s.desc.timestep_size;
  }
}

// TimeIntervals::getTimeUOld / getTimeUNew (grid/TimeIntervals.cpp:95-115).
// A virtual patch reports its neighbour interval; the pool keeps that
// interval in (time, timestep_size) of its description.
double time_uold(pc_pool* p, int slot) {
  double t, dt;
  slot_clock(p, slot, &t, &dt);
  return t;
}
double time_unew(pc_pool* p, int slot) {
  double t, dt;
  slot_clock(p, slot, &t, &dt);
  return t + dt;
}

PatchDev make_patchdev(const Slot& s) {
  PatchDev d{};
  d.off = s.off;
  d.nx = s.desc.subdivision_factor[0];
  d.ny = s.desc.subdivision_factor[1];
  d.stride = s.stride;
  d.scn = s.desc.scenario;
  d.dx = s.desc.size[0] / (double)d.nx;
  d.dy = s.desc.size[1] / (double)d.ny;
  d.dxinv = 1.0 / d.dx;
  d.dyinv = 1.0 / d.dy;
  return d;
}

// Work list: tasks of same-size patches of one scenario, large sizes first.
int rebuild_topology(pc_pool* p) {
  if (!p->topo_dirty) return PC_OK;
  const int ns = (int)p->slots.size();
  std::vector<PatchDev> pd(std::max(ns, 1));
  for (int i = 0; i < ns; ++i)
    if (p->slots[i].live) pd[i] = make_patchdev(p->slots[i]);
  const size_t old_cap = p->pd_cap;
  bool realloc = false;
  CHK(ensure_dev(&p->d_pd, &p->pd_cap, pd.size()));
  HIPCHK(hipMemcpyAsync(p->d_pd, pd.data(), pd.size() * sizeof(PatchDev),
                        hipMemcpyHostToDevice, p->stream));
  // per-slot arrays
  if (p->pd_cap != old_cap || !p->d_smax) {
    const size_t cap = p->pd_cap;
    void** arrs[] = {(void**)&p->d_smax, (void**)&p->d_smax_int, (void**)&p->d_est,
                     (void**)&p->d_ebuf[0], (void**)&p->d_ebuf[1], (void**)&p->d_bcol,
                     (void**)&p->d_sides};
    const size_t sizes[] = {cap * sizeof(double), cap * kMaxRowSeg * sizeof(double),
                            cap * sizeof(double),
                            cap * 4 * kERows * sizeof(double), cap * 4 * kERows * sizeof(double),
                            cap * 4 * kERows * sizeof(double), cap * sizeof(SideDev)};
    for (int k = 0; k < 7; ++k) {
      if (*arrs[k]) HIPCHK(hipFree(*arrs[k]));
      *arrs[k] = nullptr;
      HIPCHK(hipMalloc(arrs[k], sizes[k]));
      HIPCHK(hipMemsetAsync(*arrs[k], 0, sizes[k], p->stream));
    }
    if (p->d_iota) HIPCHK(hipFree(p->d_iota));
    p->d_iota = nullptr;
    HIPCHK(hipMalloc(&p->d_iota, cap * sizeof(int)));
    hipLaunchKernelGGL(k_iota, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, p->stream,
                       p->d_iota, (int)cap);
    HIPCHK(hipGetLastError());
    p->edges_dirty = true;
    realloc = true;
  }
  p->h_sides.resize(ns);
  bool any = false;
  for (int i = 0; i < ns; ++i)
    for (int k = 0; k < 4; ++k) any |= p->slots[i].live && p->h_sides[i].kind[k] != kSideTile;
  // every neighbour side must name a live patch of the same size and
  // scenario (its rows and edge buffers are read with this patch's stride)
  for (int i = 0; i < ns; ++i) {
    if (!p->slots[i].live) continue;
    const auto& d = p->slots[i].desc;
    for (int k = 0; k < 4; ++k) {
      if (p->h_sides[i].kind[k] != kSideNeighbour) continue;
      const int q = p->h_sides[i].src[k];
      const bool ok = q >= 0 && q < ns && p->slots[q].live &&
                      p->slots[q].desc.subdivision_factor[0] == d.subdivision_factor[0] &&
                      p->slots[q].desc.subdivision_factor[1] == d.subdivision_factor[1] &&
                      p->slots[q].desc.scenario == d.scenario;
      if (!ok)
        return fail(PC_E_INVALID, "a grid side names a released patch or one of another "
                                  "size or scenario");
    }
  }
  p->sides_any = any;
  if (ns > 0)
    HIPCHK(hipMemcpyAsync(p->d_sides, p->h_sides.data(), ns * sizeof(SideDev),
                          hipMemcpyHostToDevice, p->stream));
  // group live slots by (nx, ny, scenario)
  std::vector<int> live;
  for (int i = 0; i < ns; ++i)
    if (p->slots[i].live) live.push_back(i);
  p->h_live = live;
  std::vector<int> order = live;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const auto& da = p->slots[a].desc;
    const auto& db = p->slots[b].desc;
    long long ca = (long long)da.subdivision_factor[0] * da.subdivision_factor[1];
    long long cb = (long long)db.subdivision_factor[0] * db.subdivision_factor[1];
    if (ca != cb) return ca > cb;
    if (da.subdivision_factor[0] != db.subdivision_factor[0])
      return da.subdivision_factor[0] > db.subdivision_factor[0];
    return da.scenario < db.scenario;
  });
  p->h_tasks.clear();
  p->h_plist.clear();
  size_t i = 0;
  while (i < order.size()) {
    const auto& d0 = p->slots[order[i]].desc;
    const int nx = d0.subdivision_factor[0], ny = d0.subdivision_factor[1];
    int W = 1;
    while (W < ny) W <<= 1;
    int G = std::min(64 / W, 64 / nx);
    if (p->task_cap > 0) G = std::min(G, p->task_cap);
    G = std::max(G, 1);
    Task t;
    t.first = (int)p->h_plist.size();
    t.nseg = 0;
    t.nx = nx;
    t.ny = ny;
    while (i < order.size() && t.nseg < G) {
      const auto& d = p->slots[order[i]].desc;
      if (d.subdivision_factor[0] != nx || d.subdivision_factor[1] != ny ||
          d.scenario != d0.scenario)
        break;
      p->h_plist.push_back(order[i]);
      ++t.nseg;
      ++i;
    }
    p->h_tasks.push_back(t);
  }
  p->ntasks = (int)p->h_tasks.size();
  // the update's tasks: below the number of wavefronts the GPU holds at the
  // update's occupancy, the rows of each patch are split into row segments
  // (at least 8 rows each), so a small step is not bound by the march of
  // whole patches
  std::vector<Task> utasks;
  std::vector<int> rseg_of(ns, 1);
  {
    constexpr int kFill = 3 * 1024;  // 3 waves on each of the 1024 SIMDs
    int want = 1;
    if (p->row_segments == 0) {
      while (want < kMaxRowSeg && (long long)p->ntasks * want * 2 <= kFill) want *= 2;
    } else {
      want = p->row_segments;
    }
    for (const Task& t : p->h_tasks) {
      int r = want;
      while (r > 1 && t.nx / r < 8) r /= 2;
      for (int k = 0; k < r; ++k) {
        Task u = t;
        u.r0 = seg_first_row(k, t.nx, r);
        u.r1 = seg_first_row(k + 1, t.nx, r) - 1;
        u.sidx = k;
        utasks.push_back(u);
      }
      for (int k = 0; k < t.nseg; ++k) rseg_of[p->h_plist[t.first + k]] = r;
    }
  }
  p->nutasks = (int)utasks.size();
  // interior maxima kept per row segment are only valid for the split they
  // were computed with: a changed split re-derives them all
  bool resplit = false;
  for (int i = 0; i < ns; ++i)
    if (p->slots[i].live && (i >= (int)p->h_rseg.size() || p->h_rseg[i] != rseg_of[i]))
      resplit = true;
  p->h_rseg = rseg_of;
  // scenario patch lists
  int nscn = 0;
  for (int s : live) nscn = std::max(nscn, p->slots[s].desc.scenario + 1);
  if ((int)p->scn.size() < nscn) p->scn.resize(nscn);
  std::vector<int> plist_scn, scn_of;
  for (size_t k = 0; k < p->scn.size(); ++k) {
    p->scn[k].pbeg = (int)plist_scn.size();
    for (int s : live)
      if (p->slots[s].desc.scenario == (int)k) {
        plist_scn.push_back(s);
        scn_of.push_back((int)k);
      }
    p->scn[k].pend = (int)plist_scn.size();
  }
  p->nentries = (int)plist_scn.size();
  size_t need = std::max<size_t>({p->h_plist.size(), plist_scn.size(), 1});
  if (need > p->plist_cap) {
    if (p->d_plist) HIPCHK(hipFree(p->d_plist));
    if (p->d_plist_scn) HIPCHK(hipFree(p->d_plist_scn));
    if (p->d_scn_of) HIPCHK(hipFree(p->d_scn_of));
    p->plist_cap = need * 2;
    HIPCHK(hipMalloc(&p->d_plist, p->plist_cap * sizeof(int)));
    HIPCHK(hipMalloc(&p->d_plist_scn, p->plist_cap * sizeof(int)));
    HIPCHK(hipMalloc(&p->d_scn_of, p->plist_cap * sizeof(int)));
  }
  if (!p->h_plist.empty())
    HIPCHK(hipMemcpyAsync(p->d_plist, p->h_plist.data(), p->h_plist.size() * sizeof(int),
                          hipMemcpyHostToDevice, p->stream));
  if (!plist_scn.empty()) {
    HIPCHK(hipMemcpyAsync(p->d_plist_scn, plist_scn.data(), plist_scn.size() * sizeof(int),
                          hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipMemcpyAsync(p->d_scn_of, scn_of.data(), scn_of.size() * sizeof(int),
                          hipMemcpyHostToDevice, p->stream));
  }
  if (p->d_tasks) HIPCHK(hipFree(p->d_tasks));
  HIPCHK(hipMalloc(&p->d_tasks, std::max<size_t>(p->h_tasks.size(), 1) * sizeof(Task)));
  if (!p->h_tasks.empty())
    HIPCHK(hipMemcpyAsync(p->d_tasks, p->h_tasks.data(), p->h_tasks.size() * sizeof(Task),
                          hipMemcpyHostToDevice, p->stream));
  if (p->d_utasks) HIPCHK(hipFree(p->d_utasks));
  HIPCHK(hipMalloc(&p->d_utasks, std::max<size_t>(utasks.size(), 1) * sizeof(Task)));
  if (!utasks.empty())
    HIPCHK(hipMemcpyAsync(p->d_utasks, utasks.data(), utasks.size() * sizeof(Task),
                          hipMemcpyHostToDevice, p->stream));
  {
    // ghost-edge sources of every live patch for the boundary speed pass
    std::vector<BoundaryDesc> bdv(live.size());
    for (size_t k = 0; k < live.size(); ++k) {
      const int i = live[k];
      const PatchDev& P = pd[i];
      const SideDev& sd = p->h_sides[i];
      const long long stride = P.stride;  // a same-size neighbour's too
      BoundaryDesc& D = bdv[k];
      D.off = P.off;
      D.stride = P.stride;
      D.nx = P.nx;
      D.ny = P.ny;
      D.scn = P.scn;
      D.slot = i;
      D.dx = P.dx;
      D.dy = P.dy;
      D.rseg = rseg_of[i];
      for (int side = 0; side < 2; ++side) {
        const int kx = sd.kind[side];
        D.xwall[side] = kx == kSideWall;
        if (kx == kSideNeighbour) {
          const PatchDev& Q = pd[sd.src[side]];
          D.xrow[side] = Q.off + (side ? 1 : Q.nx) * stride;
        } else if (kx == kSideWall) {
          D.xrow[side] = P.off + (side ? P.nx : 1) * stride;
        } else {
          D.xrow[side] = P.off + (side ? P.nx + 1 : 0) * stride;
        }
        const int ky = sd.kind[2 + side];
        D.ebo[side] = (long long)((i * 2 + side) * 2) * kERows;
        if (ky == kSideNeighbour)
          D.ebg[side] = (long long)((sd.src[2 + side] * 2 + (1 - side)) * 2) * kERows;
        else
          D.ebg[side] = ky == kSideWall ? -1 : -2;
      }
      for (int d = 0; d < 2; ++d) {
        // lower side d shared with a same-size neighbour that sees this patch
        // across its upper side
        D.shared[d] = 0;
        if (sd.kind[2 * d] != kSideNeighbour) continue;
        const int q = sd.src[2 * d];
        const SideDev& sq = p->h_sides[q];
        const PatchDev& Q = pd[q];
        D.shared[d] = sq.kind[2 * d + 1] == kSideNeighbour && sq.src[2 * d + 1] == i &&
                      Q.dx == P.dx && Q.dy == P.dy && Q.nx == P.nx && Q.ny == P.ny &&
                      Q.scn == P.scn && p->slots[q].live;
      }
    }
    if (bdv.size() > p->bdesc_cap) {
      if (p->d_bdesc) HIPCHK(hipFree(p->d_bdesc));
      p->bdesc_cap = bdv.size() * 2;
      HIPCHK(hipMalloc(&p->d_bdesc, p->bdesc_cap * sizeof(BoundaryDesc)));
    }
    if (!bdv.empty())
      HIPCHK(hipMemcpyAsync(p->d_bdesc, bdv.data(), bdv.size() * sizeof(BoundaryDesc),
                            hipMemcpyHostToDevice, p->stream));
  }
  CHK(ensure_dev(&p->d_live, &p->live_cap, std::max<size_t>(live.size(), 1)));
  if (!live.empty())
    HIPCHK(hipMemcpyAsync(p->d_live, live.data(), live.size() * sizeof(int),
                          hipMemcpyHostToDevice, p->stream));
  CHK(upload_scn(p));
  HIPCHK(hipStreamSynchronize(p->stream));
  p->topo_dirty = false;
  // The interior speed maxima of unchanged patches stay valid; the pruned
  // mode re-derives those of scenarios whose patch set changed through its
  // exact fallback (their estimate is reset on register / release).
  if (fused_mode(p) != 2 || realloc || resplit) p->fused_ready = false;
  CHK(invalidate_graph(p));
  return PC_OK;
}

int prepare(pc_pool* p) {
  CHK(sync_scn_host(p));
  CHK(rebuild_topology(p));
  if (p->scn_dev_dirty) CHK(upload_scn(p));
  return PC_OK;
}

EdgeBufs edges_of(pc_pool* p) {
  EdgeBufs E;
  E.e[0] = p->d_ebuf[0];
  E.e[1] = p->d_ebuf[1];
  E.bcol = p->d_bcol;
  return E;
}

// Everything the batched step needs that is derived from the current
// state: edge buffers (after uploads / per-patch calls), and in fused mode
// the interior speed maxima of the current state.
int prepare_batched(pc_pool* p) {
  CHK(prepare(p));
  if (p->sides_any && p->edges_dirty && !p->h_live.empty()) {
    hipLaunchKernelGGL(k_edges_init, dim3((unsigned)p->h_live.size()), dim3(64), 0, p->stream,
                       p->d_live, (int)p->h_live.size(), p->d_pd, fields_of(p), p->d_scn,
                       edges_of(p), p->d_bcol);
    HIPCHK(hipGetLastError());
    p->edges_dirty = false;
    p->fused_ready = false;
  }
  if (fused_mode(p) && !p->fused_ready && p->ntasks > 0) {
    if (p->nstrips > 0)
      hipLaunchKernelGGL(k_gather, dim3((p->nstrips + 3) / 4), dim3(256), 0, p->stream,
                         p->d_strips, p->nstrips, p->d_pd, fields_of(p), p->d_scn);
    launch_speed(p, dim3(p->ntasks), p->d_tasks, p->ntasks, p->d_plist, fields_of(p),
                 p->sides_any ? p->d_sides : nullptr, edges_of(p), p->d_smax_int, nullptr,
                 kMaxRowSeg);
    HIPCHK(hipGetLastError());
    p->fused_ready = true;
  }
  return PC_OK;
}

// One step: [ghost strips] -> CFL speeds -> dt -> update.
int launch_step(pc_pool* p, bool timed_update, hipEvent_t eb = nullptr,
                hipEvent_t ee = nullptr) {
  Fields F = fields_of(p);
  EdgeBufs E = edges_of(p);
  const SideDev* sides = p->sides_any ? p->d_sides : nullptr;
  const bool fused = fused_mode(p) && p->sides_any;
  const bool prune = fused && fused_mode(p) == 2;
  if (p->nstrips > 0) {
    hipLaunchKernelGGL(k_gather, dim3((p->nstrips + 3) / 4), dim3(256), 0, p->stream,
                       p->d_strips, p->nstrips, p->d_pd, F, p->d_scn);
  }
  if (p->ntasks > 0 && !fused)
    launch_speed(p, dim3(p->ntasks), p->d_tasks, p->ntasks, p->d_plist, F, sides, E,
                 p->d_smax, nullptr);
  const int nscn = (int)p->scn.size();
  if (nscn > 0) {
    if (p->nentries > 0 && fused) {
      const int nl = (int)p->h_live.size();
      const dim3 g((nl + 3) / 4);
      if (prune)
        hipLaunchKernelGGL(k_boundary_speed<true>, g, dim3(256), 0, p->stream, p->d_scn,
                           p->d_bdesc, nl, F, E, p->d_smax_int, p->d_smax, p->prune_slack);
      else
        hipLaunchKernelGGL(k_boundary_speed<false>, g, dim3(256), 0, p->stream, p->d_scn,
                           p->d_bdesc, nl, F, E, p->d_smax_int, p->d_smax, 1.0);
    }
    // few patch entries: both dt stages in one workgroup
    const bool block = p->nentries <= kDtBlockEntries;
    if (block) {
      hipLaunchKernelGGL(k_dt_block, dim3(1), dim3(1024), 0, p->stream, p->d_scn, nscn,
                         p->d_plist_scn, p->d_scn_of, p->nentries, p->d_pd, p->d_smax, p->d_est,
                         p->d_counter, prune ? 1 : 0, p->d_redo, p->prune_slack);
    } else {
      if (p->nentries > 0)
        hipLaunchKernelGGL(k_dt_partial, dim3((p->nentries + 1023) / 1024), dim3(256), 0,
                           p->stream, p->d_scn, p->d_plist_scn, p->d_scn_of, p->nentries,
                           p->d_pd, p->d_smax, p->d_est, prune ? p->d_redo : nullptr, nullptr);
      hipLaunchKernelGGL(k_dt_final, dim3((nscn + 63) / 64), dim3(64), 0, p->stream, p->d_scn,
                         nscn, p->d_counter, prune ? 1 : 0, p->d_redo, p->prune_slack);
    }
    if (prune) {
      // exact fallback, a no-op unless the pruned minimum was not proven exact
      if (p->ntasks > 0)
        hipLaunchKernelGGL((k_speed<true, 0>), dim3(std::min(p->ntasks, kFallbackWgs)),
                           dim3(64), 0, p->stream, p->d_tasks, p->ntasks, p->d_plist, p->d_pd,
                           F, p->d_scn, sides, E, p->d_smax, p->d_redo, 1);
      if (block) {
        hipLaunchKernelGGL(k_dt_block, dim3(1), dim3(1024), 0, p->stream, p->d_scn, nscn,
                           p->d_plist_scn, p->d_scn_of, p->nentries, p->d_pd, p->d_smax,
                           p->d_est, p->d_counter, 2, p->d_redo, p->prune_slack);
      } else {
        if (p->nentries > 0)
          hipLaunchKernelGGL(k_dt_partial, dim3((p->nentries + 1023) / 1024), dim3(256), 0,
                             p->stream, p->d_scn, p->d_plist_scn, p->d_scn_of, p->nentries,
                             p->d_pd, p->d_smax, p->d_est, nullptr, p->d_redo);
        hipLaunchKernelGGL(k_dt_final, dim3((nscn + 63) / 64), dim3(64), 0, p->stream,
                           p->d_scn, nscn, p->d_counter, 2, p->d_redo, p->prune_slack);
      }
    }
  }
  if (p->ntasks > 0) {
    if (timed_update) HIPCHK(hipEventRecord(eb ? eb : p->ev2, p->stream));
    // a persistent grid only pays when there are more tasks than workgroups:
    // below that every workgroup would take one task through the shared
    // counter (contended atomics) instead of its own block index
    const int pers = (p->persistent_wgs > 0 && p->nutasks > p->persistent_wgs) ? 1 : 0;
    const int grid = pers ? p->persistent_wgs : p->nutasks;
    // field arrays below 2^29 doubles: 32-bit tile offsets (TileMem)
    const bool o32 = p->cap_cells <= (1LL << 29);
    if (prune && o32)
      hipLaunchKernelGGL((k_update<2, true, 0, true>), dim3(grid), dim3(64), 0, p->stream,
                         p->d_utasks, p->nutasks, p->d_plist, p->d_pd, F, p->d_scn, sides, E,
                         p->d_smax_int, p->d_counter, pers, p->prune_slack);
    else if (prune)
      hipLaunchKernelGGL((k_update<2, true, 0>), dim3(grid), dim3(64), 0, p->stream,
                         p->d_utasks, p->nutasks, p->d_plist, p->d_pd, F, p->d_scn, sides, E,
                         p->d_smax_int, p->d_counter, pers, p->prune_slack);
    else if (fused && o32)
      hipLaunchKernelGGL((k_update<1, true, 0, true>), dim3(grid), dim3(64), 0, p->stream,
                         p->d_utasks, p->nutasks, p->d_plist, p->d_pd, F, p->d_scn, sides, E,
                         p->d_smax_int, p->d_counter, pers, p->prune_slack);
    else if (fused)
      hipLaunchKernelGGL((k_update<1, true, 0>), dim3(grid), dim3(64), 0, p->stream,
                         p->d_utasks, p->nutasks, p->d_plist, p->d_pd, F, p->d_scn, sides, E,
                         p->d_smax_int, p->d_counter, pers, p->prune_slack);
    else
      launch_update_plain(p, dim3(grid), p->d_utasks, p->nutasks, p->d_plist, F, sides, E,
                          pers);
    if (timed_update) HIPCHK(hipEventRecord(ee ? ee : p->ev3, p->stream));
  }
  HIPCHK(hipGetLastError());
  return PC_OK;
}

}  // namespace

extern "C" {

const char* pc_last_error_string(void) { return g_err.c_str(); }

// Staging of pc_edge_batch_static: a file-scope device array (one copy per
// GPU, held until the process ends) of kStaticEdgeMax edges, 6 input and 5
// output doubles each.  Host code reaches it only through the device
// address hipGetSymbolAddress returns on that GPU, kept per device.
constexpr long long kStaticEdgeMax = 65536;
static __device__ __align__(256) double scratch[kStaticEdgeMax * 11];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Once per device, with `device` current: the device address of `scratch`.
static int scratch_setup(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) return fail(PC_E_INVALID, "device index too large");
  if (g_scratch_ptr[device]) return PC_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(PC_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = static_cast<double*>(ptr);
  return PC_OK;
}

int pc_pool_create(int device, size_t hbm_budget_bytes, pc_pool** out) {
  if (!out) return fail(PC_E_INVALID, "out is NULL");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0)
    return fail(PC_E_NODEVICE, "no HIP device with that index");
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(PC_E_NODEVICE, std::string("device is ") + prop.gcnArchName +
                                   ", this library is built for gfx950");
  HIPCHK(hipSetDevice(device));
  CHK(scratch_setup(device));
  pc_pool* p = new pc_pool();
  p->device = device;
  p->budget = hbm_budget_bytes ? hbm_budget_bytes : (size_t)prop.totalGlobalMem / 10 * 9;
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
    delete p;
    return fail(PC_E_HIP, "hipStreamCreate failed");
  }
  HIPCHK(hipMalloc(&p->d_counter, 2 * sizeof(int)));
  HIPCHK(hipMemsetAsync(p->d_counter, 0, 2 * sizeof(int), p->stream));
  p->d_redo = p->d_counter + 1;
  HIPCHK(hipMalloc(&p->d_result, 8 * sizeof(double)));
  // fine-grained (coherent), so a kernel's stores are visible while it runs
  HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&p->h_result), 8 * sizeof(double),
                       hipHostMallocMapped | hipHostMallocCoherent));
  std::memset(p->h_result, 0, 8 * sizeof(double));
  HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void**>(&p->d_result_host), p->h_result, 0));
  HIPCHK(hipMalloc(&p->d_segmax, kPatchSeg * sizeof(double)));
  HIPCHK(hipMalloc(&p->d_keep, 3 * 64 * 64 * sizeof(double)));
  HIPCHK(hipEventCreate(&p->ev0));
  HIPCHK(hipEventCreate(&p->ev1));
  HIPCHK(hipEventCreate(&p->ev2));
  HIPCHK(hipEventCreate(&p->ev3));
  *out = p;
  return PC_OK;
}

int pc_pool_destroy(pc_pool* p) {
  if (!p) return fail(PC_E_INVALID, "pool is NULL");
  hipSetDevice(p->device);
  hipStreamSynchronize(p->stream);
  invalidate_graph(p);
  for (int k = 0; k < kNumFields; ++k)
    if (p->f[k]) hipFree(p->f[k]);
  void* ptrs[] = {p->d_pd, p->d_smax, p->d_smax_int, p->d_est, p->d_scn, p->d_tasks, p->d_utasks,
                  p->d_plist, p->d_plist_scn, p->d_scn_of, p->d_live, p->d_bdesc, p->d_strips,
                  p->d_counter, p->d_iota, p->d_segmax, p->d_keep, p->d_result, p->d_aux, p->d_flags, p->d_ebuf[0],
                  p->d_ebuf[1], p->d_bcol, p->d_sides, p->d_single, p->d_single_slot,
                  p->d_edge_io, p->d_flux, p->d_batch, p->amr_tables};
  for (void* q : ptrs)
    if (q) hipFree(q);
  if (p->h_result) hipHostFree(p->h_result);
  if (p->h_stage) hipHostFree(p->h_stage);
  for (hipEvent_t e : p->upd_ev) hipEventDestroy(e);
  hipEventDestroy(p->ev0);
  hipEventDestroy(p->ev1);
  hipEventDestroy(p->ev2);
  hipEventDestroy(p->ev3);
  hipStreamDestroy(p->stream);
  delete p;
  return PC_OK;
}

int pc_pool_reserve(pc_pool* p, long long cells) {
  if (!p || cells < 0) return fail(PC_E_INVALID, "invalid argument");
  hipSetDevice(p->device);
  CHK(grow_fields(p, cells));
  HIPCHK(hipStreamSynchronize(p->stream));
  return PC_OK;
}

int pc_pool_bytes_in_use(pc_pool* p, size_t* out) {
  if (!p || !out) return fail(PC_E_INVALID, "NULL argument");
  long long freec = 0;
  for (auto& r : p->free_ranges) freec += r.second;
  *out = (size_t)(p->top - freec) * kBytesPerCell;
  return PC_OK;
}

int pc_pool_synchronize(pc_pool* p) {
  if (!p) return fail(PC_E_INVALID, "pool is NULL");
  HIPCHK(hipStreamSynchronize(p->stream));
  return PC_OK;
}

int pc_patch_register(pc_pool* p, const pc_cell_description* d, int* out_slot) {
  if (!p || !d || !out_slot) return fail(PC_E_INVALID, "NULL argument");
  const int nx = d->subdivision_factor[0], ny = d->subdivision_factor[1];
  if (nx < 2 || ny < 2 || nx > 64 || ny > 64)
    return fail(PC_E_INVALID, "subdivision factor must be in [2, 64]");
  if (d->ghostlayer_width != 1)
    return fail(PC_E_INVALID, "the SWE kernel uses ghost layer width 1");
  if (d->scenario < 0 || d->scenario > (1 << 20))
    return fail(PC_E_INVALID, "scenario id out of range");
  if (!(d->size[0] > 0) || !(d->size[1] > 0))
    return fail(PC_E_INVALID, "patch size must be positive");
  hipSetDevice(p->device);
  CHK(sync_scn_host(p));
  Slot s;
  s.live = true;
  s.desc = *d;
  s.stride = tile_stride(ny);
  s.cells = (long long)(nx + 2) * s.stride;
  CHK(alloc_cells(p, s.cells, &s.off, tile_phase(ny)));
  // cells above the high-water mark are still zero from the allocation
  if (s.off < p->hwm)
    for (int k = 0; k < kNumFields; ++k)
      HIPCHK(hipMemsetAsync(p->f[k] + s.off, 0, (size_t)s.cells * sizeof(double), p->stream));
  p->hwm = std::max(p->hwm, s.off + s.cells);
  int slot = -1;
  for (size_t i = 0; i < p->slots.size(); ++i)
    if (!p->slots[i].live) {
      slot = (int)i;
      break;
    }
  if (slot < 0) {
    slot = (int)p->slots.size();
    p->slots.push_back(s);
  } else {
    p->slots[slot] = s;
  }
  if ((int)p->scn.size() <= d->scenario) {
    size_t old = p->scn.size();
    p->scn.resize(d->scenario + 1);
    for (size_t k = old; k < p->scn.size(); ++k) {
      ScnDev z{};
      z.end_time = INFINITY;
      z.minbits = 0x7FEFFFFFFFFFFFFFull;  // DBL_MAX
      p->scn[k] = z;
    }
    p->scn[d->scenario].time = d->time;
    p->scn[d->scenario].dt = d->timestep_size;
  }
  p->scn[d->scenario].est = 0.0;  // patch set changed: no pruning until exact again
  p->scn_dev_dirty = true;
  p->topo_dirty = true;
  *out_slot = slot;
  return PC_OK;
}

int pc_patch_release(pc_pool* p, int slot) {
  if (!valid_slot(p, slot)) return fail(PC_E_INVALID, "invalid slot");
  CHK(sync_scn_host(p));
  p->scn[p->slots[slot].desc.scenario].est = 0.0;
  p->scn_dev_dirty = true;
  free_cells(p, p->slots[slot].off, p->slots[slot].cells);
  p->slots[slot].live = false;
  if ((size_t)slot < p->h_sides.size()) p->h_sides[slot] = SideDev{};
  // A side that read its ghosts from the released patch now faces a hole
  // in the grid: it becomes a reflective wall, as a grid with that patch
  // missing would have been given by pc_set_grid_sides.  The slot id may be
  // reused by a patch of another size, which must never be read as a
  // neighbour.
  for (size_t i = 0; i < p->h_sides.size(); ++i)
    for (int k = 0; k < 4; ++k)
      if (p->h_sides[i].kind[k] == kSideNeighbour && p->h_sides[i].src[k] == slot) {
        p->h_sides[i].kind[k] = kSideWall;
        p->h_sides[i].src[k] = (int)i;
      }
  p->topo_dirty = true;
  return PC_OK;
}

int pc_patch_get_description(pc_pool* p, int slot, pc_cell_description* out) {
  if (!valid_slot(p, slot) || !out) return fail(PC_E_INVALID, "invalid slot");
  CHK(sync_scn_host(p));
  *out = p->slots[slot].desc;
  if (p->slots[slot].clock_from_scenario) {
    const ScnDev& S = p->scn[out->scenario];
    out->time = S.time;
    out->timestep_size = S.dt;
    double e = 0.0;
    if (fused_mode(p) == 2 && p->sides_any) {
      // pruned speeds: only the scenario minimum is exact
      e = S.est;
    } else if (p->d_est && !p->topo_dirty) {
      HIPCHK(hipMemcpyAsync(&e, p->d_est + slot, sizeof(double), hipMemcpyDeviceToHost,
                            p->stream));
      HIPCHK(hipStreamSynchronize(p->stream));
    }
    out->estimated_next_timestep_size = e;
  }
  return PC_OK;
}

int pc_patch_set_description(pc_pool* p, int slot, const pc_cell_description* d) {
  if (!valid_slot(p, slot) || !d) return fail(PC_E_INVALID, "invalid slot");
  Slot& s = p->slots[slot];
  if (d->subdivision_factor[0] != s.desc.subdivision_factor[0] ||
      d->subdivision_factor[1] != s.desc.subdivision_factor[1] ||
      d->scenario != s.desc.scenario || d->ghostlayer_width != 1)
    return fail(PC_E_INVALID, "size, ghost width and scenario are fixed per slot");
  bool geom = d->size[0] != s.desc.size[0] || d->size[1] != s.desc.size[1];
  s.desc = *d;
  s.clock_from_scenario = false;
  if (geom) {
    p->topo_dirty = true;
    CHK(sync_scn_host(p));
    p->scn[d->scenario].est = 0.0;
    p->scn_dev_dirty = true;
  }
  return PC_OK;
}

int pc_patch_upload(pc_pool* p, int slot, const double* unew, const double* uold,
                    const double* b) {
  if (!valid_slot(p, slot)) return fail(PC_E_INVALID, "invalid slot");
  hipSetDevice(p->device);
  CHK(sync_scn_host(p));
  const Slot& s = p->slots[slot];
  const int nx = s.desc.subdivision_factor[0], ny = s.desc.subdivision_factor[1];
  const int cur = cur_of(p, slot);
  // host arrays in the reference layout (rows of ny, or ny + 2 with the
  // ghost ring), tile rows of s.stride
  const size_t w = (size_t)(ny + 2), area = (size_t)(nx + 2) * w, pitch = s.stride * sizeof(double);
  if (unew) {
    for (int q = 0; q < 3; ++q) {
      double* dst = p->f[2 * q + cur] + s.off + s.stride + 1;
      HIPCHK(hipMemcpy2DAsync(dst, pitch, unew + (size_t)q * nx * ny, ny * sizeof(double),
                              ny * sizeof(double), nx, hipMemcpyHostToDevice, p->stream));
    }
  }
  if (uold) {
    for (int q = 0; q < 3; ++q)
      HIPCHK(hipMemcpy2DAsync(p->f[2 * q + 1 - cur] + s.off, pitch, uold + q * area,
                              w * sizeof(double), w * sizeof(double), nx + 2,
                              hipMemcpyHostToDevice, p->stream));
  }
  if (b)
    HIPCHK(hipMemcpy2DAsync(p->f[6] + s.off, pitch, b, w * sizeof(double), w * sizeof(double),
                            nx + 2, hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  mark_state_changed(p);
  return PC_OK;
}

int pc_patch_download(pc_pool* p, int slot, double* unew, double* uold, double* b) {
  if (!valid_slot(p, slot)) return fail(PC_E_INVALID, "invalid slot");
  hipSetDevice(p->device);
  CHK(sync_scn_host(p));
  const Slot& s = p->slots[slot];
  const int nx = s.desc.subdivision_factor[0], ny = s.desc.subdivision_factor[1];
  const int cur = cur_of(p, slot);
  const size_t w = (size_t)(ny + 2), area = (size_t)(nx + 2) * w, pitch = s.stride * sizeof(double);
  if (unew) {
    for (int q = 0; q < 3; ++q) {
      const double* src = p->f[2 * q + cur] + s.off + s.stride + 1;
      HIPCHK(hipMemcpy2DAsync(unew + (size_t)q * nx * ny, ny * sizeof(double), src, pitch,
                              ny * sizeof(double), nx, hipMemcpyDeviceToHost, p->stream));
    }
  }
  if (uold) {
    for (int q = 0; q < 3; ++q)
      HIPCHK(hipMemcpy2DAsync(uold + q * area, w * sizeof(double), p->f[2 * q + 1 - cur] + s.off,
                              pitch, w * sizeof(double), nx + 2, hipMemcpyDeviceToHost,
                              p->stream));
  }
  if (b)
    HIPCHK(hipMemcpy2DAsync(b, w * sizeof(double), p->f[6] + s.off, pitch, w * sizeof(double),
                            nx + 2, hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return PC_OK;
}

int pc_patch_offset(pc_pool* p, int slot, long long* off) {
  if (!valid_slot(p, slot) || !off) return fail(PC_E_INVALID, "invalid slot");
  *off = p->slots[slot].off;
  return PC_OK;
}

int pc_patch_stride(pc_pool* p, int slot, int* stride) {
  if (!valid_slot(p, slot) || !stride) return fail(PC_E_INVALID, "invalid slot");
  *stride = p->slots[slot].stride;
  return PC_OK;
}

int pc_tile_stride(int ny) { return ny >= 2 && ny <= 64 ? tile_stride(ny) : -1; }

int pc_scenario_current_state(pc_pool* p, int scenario, int* state) {
  if (!p || !state || scenario < 0) return fail(PC_E_INVALID, "invalid argument");
  CHK(sync_scn_host(p));
  if (scenario >= (int)p->scn.size()) return fail(PC_E_INVALID, "unknown scenario");
  *state = p->scn[scenario].stepc & 1;
  return PC_OK;
}

int pc_pool_copy_fields(pc_pool* p, int to_device, int state, long long begin,
                        long long count, double* h, double* hu, double* hv, double* b) {
  if (!p || (state != 0 && state != 1) || begin < 0 || count < 0 ||
      begin + count > p->cap_cells)
    return fail(PC_E_INVALID, "invalid range");
  if (count == 0) return PC_OK;
  hipSetDevice(p->device);
  double* host[4] = {h, hu, hv, b};
  double* dev[4] = {p->f[0 + state], p->f[2 + state], p->f[4 + state], p->f[6]};
  for (int k = 0; k < 4; ++k) {
    if (!host[k]) continue;
    if (to_device)
      HIPCHK(hipMemcpyAsync(dev[k] + begin, host[k], (size_t)count * sizeof(double),
                            hipMemcpyHostToDevice, p->stream));
    else
      HIPCHK(hipMemcpyAsync(host[k], dev[k] + begin, (size_t)count * sizeof(double),
                            hipMemcpyDeviceToHost, p->stream));
  }
  HIPCHK(hipStreamSynchronize(p->stream));
  if (to_device) mark_state_changed(p);
  return PC_OK;
}

int pc_get_number_of_unknowns_per_cell(void) { return 3; }
int pc_get_number_of_parameter_fields_without_ghostlayer(void) { return 0; }
int pc_get_number_of_parameter_fields_with_ghostlayer(void) { return 1; }
int pc_get_ghostlayer_width(void) { return 1; }

}  // extern "C"

namespace {
// Wait for a synchronous per-call kernel: spin on its completion word in
// pinned host memory (visible as soon as the kernel has stored its results,
// without waiting for the kernel's retirement to reach the stream), for at
// most 2 ms; then, or if the word does not come, synchronise the stream,
// which also reports any error.
int wait_call(pc_pool* p, unsigned long long seq) {
  const unsigned long long* word =
      reinterpret_cast<const unsigned long long*>(p->h_result + kDoneWord);
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned it = 0;; ++it) {
    if (__atomic_load_n(word, __ATOMIC_ACQUIRE) == seq) return PC_OK;
    if ((it & 255) == 255 &&
        std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2))
      break;
  }
  HIPCHK(hipStreamSynchronize(p->stream));
  if (__atomic_load_n(word, __ATOMIC_ACQUIRE) != seq)
    return fail(PC_E_HIP, "a per-call kernel did not signal its completion");
  return PC_OK;
}
}  // namespace

extern "C" {

int pc_solve_timestep(pc_pool* p, int slot, double maxDt, const int flags[4],
                      double* out_dt, double* out_est) {
  (void)flags;
  if (!valid_slot(p, slot)) return fail(PC_E_INVALID, "invalid slot");
  hipSetDevice(p->device);
  CHK(prepare(p));
  mark_state_changed(p);
  Slot& s = p->slots[slot];
  if (s.clock_from_scenario) {
    slot_clock(p, slot, &s.desc.time, &s.desc.timestep_size);
    s.clock_from_scenario = false;
  }
  const int nseg = patch_row_segments(s.desc.subdivision_factor[0]);
  const unsigned long long seq = ++p->call_seq;
  if (p->solver)
    hipLaunchKernelGGL(k_patch_solve<1>, dim3(1), dim3(64 * nseg), 0, p->stream, slot,
                       p->d_iota, p->d_pd, fields_of(p), cur_of(p, slot), maxDt, p->d_segmax,
                       p->d_result_host, p->d_keep, seq);
  else
    hipLaunchKernelGGL(k_patch_solve<0>, dim3(1), dim3(64 * nseg), 0, p->stream, slot,
                       p->d_iota, p->d_pd, fields_of(p), cur_of(p, slot), maxDt, p->d_segmax,
                       p->d_result_host, p->d_keep, seq);
  HIPCHK(hipGetLastError());
  CHK(wait_call(p, seq));
  const double res[2] = {p->h_result[0], p->h_result[1]};
  // SWEKernel.cpp:83-87
  if (la_greater(res[0], 0.0)) {
    s.desc.time = s.desc.time + s.desc.timestep_size;
    s.desc.timestep_size = res[0];
  }
  s.desc.estimated_next_timestep_size = res[1];
  if (out_dt) *out_dt = res[0];
  if (out_est) *out_est = res[1];
  return PC_OK;
}

int pc_patch_cfl(pc_pool* p, int slot, double* out) {
  if (!valid_slot(p, slot) || !out) return fail(PC_E_INVALID, "invalid slot or output");
  hipSetDevice(p->device);
  CHK(prepare(p));
  const int nseg = patch_row_segments(p->slots[slot].desc.subdivision_factor[0]);
  const unsigned long long seq = ++p->call_seq;
  if (p->solver)
    hipLaunchKernelGGL(k_patch_cfl<1>, dim3(1), dim3(64 * nseg), 0, p->stream, slot,
                       p->d_iota, p->d_pd, fields_of(p), cur_of(p, slot), p->d_segmax,
                       p->d_result_host, seq);
  else
    hipLaunchKernelGGL(k_patch_cfl<0>, dim3(1), dim3(64 * nseg), 0, p->stream, slot,
                       p->d_iota, p->d_pd, fields_of(p), cur_of(p, slot), p->d_segmax,
                       p->d_result_host, seq);
  HIPCHK(hipGetLastError());
  CHK(wait_call(p, seq));
  *out = p->h_result[0];
  return PC_OK;
}

int pc_fill_boundary_layer(pc_pool* p, int slot, int dim, int upper) {
  if (!valid_slot(p, slot) || (dim != 0 && dim != 1))
    return fail(PC_E_INVALID, "invalid slot or dimension");
  hipSetDevice(p->device);
  CHK(prepare(p));
  mark_state_changed(p);
  hipLaunchKernelGGL(k_fill_boundary, dim3(1), dim3(128), 0, p->stream, p->d_pd, slot,
                     fields_of(p), 1 - cur_of(p, slot), dim, upper ? 1 : 0);
  HIPCHK(hipGetLastError());
  return PC_OK;
}

int pc_transfer_ghostlayer(pc_pool* p, const int size[2], const int so[2],
                           const int dof[2], int src, int dst) {
  if (!valid_slot(p, src) || !valid_slot(p, dst) || !size || !so || !dof)
    return fail(PC_E_INVALID, "invalid argument");
  hipSetDevice(p->device);
  CHK(prepare(p));
  mark_state_changed(p);
  const auto& sd = p->slots[src].desc;
  const auto& dd = p->slots[dst].desc;
  for (int d = 0; d < 2; ++d) {
    if (so[d] < 0 || so[d] + size[d] > sd.subdivision_factor[d] || size[d] < 0 ||
        dof[d] < -1 || dof[d] + size[d] > dd.subdivision_factor[d] + 1)
      return fail(PC_E_INVALID, "transfer region out of bounds");
  }
  // DefaultTransfer.cpph:27-51
  double tf;
  const double dstTimeUNew = time_unew(p, dst);
  if (sd.is_virtual) {
    tf = (dstTimeUNew - 0.0) / 1.0;
  } else {
    const double sNew = time_unew(p, src), sOld = time_uold(p, src);
    if (la_greater(sNew - sOld, 0.0))
      tf = (dstTimeUNew - sOld) / (sNew - sOld);
    else
      tf = 1.0;
  }
  const double inv = 1.0 - tf;
  const int n = size[0] * size[1];
  if (n == 0) return PC_OK;
  hipLaunchKernelGGL(k_transfer, dim3((n + 255) / 256), dim3(256), 0, p->stream, p->d_pd,
                     src, dst, fields_of(p), cur_of(p, src), cur_of(p, dst), size[0], size[1],
                     so[0], so[1], dof[0], dof[1], tf, inv);
  HIPCHK(hipGetLastError());
  return PC_OK;
}

// SWEKernel::transformWaterHeight over a region of one state (uold: the
// previous state, else uNew with its layout), planned and checked before
// anything runs; `out.sx * out.sy == 0` is a no-op.
static int plan_transform(pc_pool* p, int slot, const int off[2], const int size[2], int uold,
                          int to_above_sea_floor, TransformItem& out) {
  out = TransformItem{};
  out.slot = slot;
  if (size[0] * size[1] <= 0) return PC_OK;
  const int cur = cur_of(p, slot);
  const auto& d = p->slots[slot].desc;
  const int nx = d.subdivision_factor[0], ny = d.subdivision_factor[1];
  for (int k = 0; k < 2; ++k)
    if (off[k] < -1 || off[k] + size[k] > (k ? ny : nx) + 1)
      return fail(PC_E_INVALID, "transform region out of bounds");
  if (!uold && (off[0] < 0 || off[1] < 0 ||
                (long long)(off[0] + size[0] - 1) * ny + (off[1] + size[1] - 1) >= 3LL * nx * ny))
    return fail(PC_E_INVALID, "uNew transform region outside the unknowns");
  out.st = uold ? 1 - cur : cur;
  out.ox = off[0];
  out.oy = off[1];
  out.sx = size[0];
  out.sy = size[1];
  out.sign = to_above_sea_floor ? -1.0 : 1.0;
  out.unew_layout = uold ? 0 : 1;
  return PC_OK;
}

static int launch_transform(pc_pool* p, const TransformItem& it) {
  const int n = it.sx * it.sy;
  if (n <= 0) return PC_OK;
  hipLaunchKernelGGL(k_transform, dim3((n + 255) / 256), dim3(256), 0, p->stream, p->d_pd,
                     it.slot, fields_of(p), it.st, it.ox, it.oy, it.sx, it.sy, it.sign,
                     it.unew_layout);
  HIPCHK(hipGetLastError());
  return PC_OK;
}

static int transform(pc_pool* p, int slot, const int off[2], const int size[2], int uold,
                     int to_above_sea_floor) {
  TransformItem it;
  CHK(plan_transform(p, slot, off, size, uold, to_above_sea_floor, it));
  return launch_transform(p, it);
}

// Region::mapToPatch (Region.cpp:174-206)
static void map_to_patch(const int off[2], const int size[2], const PatchDev& src,
                         const pc_cell_description& sd, const PatchDev& dst,
                         const pc_cell_description& dd, double eps, int o[2], int s[2]) {
  const double ssub[2] = {src.dx, src.dy}, dsub[2] = {dst.dx, dst.dy};
  for (int d = 0; d < 2; ++d) {
    double position = ssub[d] * (double)off[d];
    position += sd.position[d] - dd.position[d];
    double ot = position + eps;
    ot /= dsub[d];
    o[d] = (int)ot;
    double sz = ssub[d] * (double)(off[d] + size[d]);
    sz += sd.position[d] - dd.position[d];
    double st = sz - eps;
    st /= dsub[d];
    st -= (double)o[d];
    st += 1.0;
    s[d] = (int)st;
  }
}

// One interpolateSolution call (SWEKernel.cpp:239-292 around
// DefaultInterpolation.cpph:29-270), planned and checked before anything
// runs: the source region (grown by one cell) to h + b in both states, the
// interpolation, both back to h, the destination region back to h.
struct InterpPlan {
  InterpArgs a;
  int n;  // destination cells
  TransformItem fwd[2], back[3];
};

static int plan_interpolation(pc_pool* p, const int dsize[2], const int doff[2], int src,
                              int dst, int toUOld, int toCurrentTime, int useTimeUNewOrUOld,
                              InterpPlan& out) {
  if (!valid_slot(p, src) || !valid_slot(p, dst) || !dsize || !doff)
    return fail(PC_E_INVALID, "invalid argument");
  const auto& sd = p->slots[src].desc;
  const auto& dd = p->slots[dst].desc;
  const PatchDev S = make_patchdev(p->slots[src]), D = make_patchdev(p->slots[dst]);
  for (int d = 0; d < 2; ++d)
    if (doff[d] < -1 || dsize[d] < 0 || doff[d] + dsize[d] > dd.subdivision_factor[d] + 1 ||
        (!toUOld && (doff[d] < 0 || doff[d] + dsize[d] > dd.subdivision_factor[d])))
      return fail(PC_E_INVALID, "interpolation region out of bounds");
  // SWEKernel.cpp:252-263: source region grown by one cell
  int so[2], ss[2];
  map_to_patch(doff, dsize, D, dd, S, sd, 1e-12, so, ss);
  const int sn[2] = {sd.subdivision_factor[0], sd.subdivision_factor[1]};
  for (int d = 0; d < 2; ++d) {
    if (so[d] > 0) {
      so[d] = so[d] - 1;
      ss[d] = std::min(sn[d], ss[d] + 2);
    } else {
      ss[d] = std::min(sn[d], ss[d] + 1);
    }
    // the grown region may reach index n (the reference transforms it too)
    if (so[d] < 0 || so[d] + ss[d] > sn[d] + 1)
      return fail(PC_E_INVALID, "destination region is not covered by the source patch");
  }
  // DefaultInterpolation.cpph:45-67 time factor
  double st, sdt, dt_, ddt;
  slot_clock(p, src, &st, &sdt);
  slot_clock(p, dst, &dt_, &ddt);
  double tf;
  if (std::fabs(sdt - 0.0) <= kLaEps) {
    tf = 1.0;
  } else if (useTimeUNewOrUOld) {
    const double sOld = time_uold(p, src), sNew = time_unew(p, src);
    tf = toCurrentTime ? (time_uold(p, dst) - sOld) / (sNew - sOld)
                       : (time_unew(p, dst) - sOld) / (sNew - sOld);
  } else {
    tf = toCurrentTime ? (dt_ - st) / (st + sdt - st) : (dt_ + ddt - st) / (st + sdt - st);
  }
  tf = std::max(0.0, std::min(1.0, tf));
  CHK(plan_transform(p, src, so, ss, 1, 0, out.fwd[0]));
  CHK(plan_transform(p, src, so, ss, 0, 0, out.fwd[1]));
  CHK(plan_transform(p, src, so, ss, 1, 1, out.back[0]));
  CHK(plan_transform(p, src, so, ss, 0, 1, out.back[1]));
  CHK(plan_transform(p, dst, doff, dsize, toUOld, 1, out.back[2]));
  InterpArgs& a = out.a;
  a = InterpArgs{};
  a.srcSlot = src; a.dstSlot = dst;
  a.srcCur = cur_of(p, src); a.dstCur = cur_of(p, dst);
  a.sx = dsize[0]; a.sy = dsize[1]; a.ox = doff[0]; a.oy = doff[1];
  a.toUOld = toUOld ? 1 : 0;
  a.tf = tf;
  a.spos[0] = sd.position[0]; a.spos[1] = sd.position[1];
  a.ssub[0] = S.dx; a.ssub[1] = S.dy;
  a.sinv[0] = 1.0 / S.dx; a.sinv[1] = 1.0 / S.dy;
  a.dpos[0] = dd.position[0]; a.dpos[1] = dd.position[1];
  a.dsub[0] = D.dx; a.dsub[1] = D.dy;
  a.sn[0] = sn[0]; a.sn[1] = sn[1];
  out.n = dsize[0] * dsize[1];
  return PC_OK;
}

int pc_interpolate_solution(pc_pool* p, const int dsize[2], const int doff[2], int src,
                            int dst, int toUOld, int toCurrentTime, int useTimeUNewOrUOld) {
  if (!valid_slot(p, src) || !valid_slot(p, dst) || !dsize || !doff)
    return fail(PC_E_INVALID, "invalid argument");
  hipSetDevice(p->device);
  CHK(prepare(p));
  InterpPlan pl;
  CHK(plan_interpolation(p, dsize, doff, src, dst, toUOld, toCurrentTime, useTimeUNewOrUOld,
                         pl));
  mark_state_changed(p);
  CHK(launch_transform(p, pl.fwd[0]));
  CHK(launch_transform(p, pl.fwd[1]));
  if (pl.n > 0) {
    hipLaunchKernelGGL(k_interpolate, dim3((pl.n + 127) / 128), dim3(128), 0, p->stream,
                       p->d_pd, fields_of(p), pl.a);
    HIPCHK(hipGetLastError());
  }
  for (const TransformItem& it : pl.back) CHK(launch_transform(p, it));
  return PC_OK;
}

// Region::getRegionsOverlappedByNeighboringGhostlayers (Region.cpp:247-315)
static int overlapped_regions(const double lower[2], const double upper[2],
                              const double pos[2], const double size[2],
                              const double sub[2], const int n[2], int off[8], int sz[8]) {
  auto gt = [](double a, double b) { return a - b > kLaEps; };
  std::memset(off, 0, 8 * sizeof(int));
  std::memset(sz, 0, 8 * sizeof(int));
  const bool oneGreater = gt(upper[0] + sub[0], lower[0]) || gt(upper[1] + sub[1], lower[1]);
  const bool oneGE = (upper[0] - (pos[0] + size[0]) >= -kLaEps) ||
                     (upper[1] - (pos[1] + size[1]) >= -kLaEps);
  const bool allGreater = gt(lower[0], pos[0]) && gt(lower[1], pos[1]);
  if (oneGreater || oneGE || !allGreater) {
    sz[0] = n[0];
    sz[1] = n[1];
    return 1;
  }
  const double eps = 1e-12;
  int lo[2], hi[2];
  for (int d = 0; d < 2; ++d) {
    double dl = (lower[d] - pos[d]) * (1.0 / sub[d]) + eps;
    double du = (upper[d] - pos[d]) * (1.0 / sub[d]) - eps;
    dl = std::min((double)2147483647, dl);
    du = std::max((double)(-2147483647 - 1), du);
    lo[d] = std::max(std::min((int)std::floor(dl), n[d]), -1);
    hi[d] = std::max(std::min((int)std::floor(du), n[d]), -1);
  }
  for (int d = 0; d < 2; ++d) {
    const int r0 = 2 * d, r1 = 2 * d + 1;
    sz[2 * r0 + d] = hi[d] + 1;
    off[2 * r0 + d] = 0;
    sz[2 * r1 + d] = n[d] - lo[d];
    off[2 * r1 + d] = lo[d];
    for (int i = 0; i < d; ++i) {
      sz[2 * r0 + i] = -hi[i] - 1 + lo[i];
      off[2 * r0 + i] = hi[i] + 1;
      sz[2 * r1 + i] = sz[2 * r0 + i];
      off[2 * r1 + i] = off[2 * r0 + i];
    }
    for (int i = d + 1; i < 2; ++i) {
      sz[2 * r0 + i] = n[i];
      off[2 * r0 + i] = 0;
      sz[2 * r1 + i] = n[i];
      off[2 * r1 + i] = 0;
    }
  }
  return 4;
}

// One restrictSolution call (SWEKernel.cpp:294-310 around
// DefaultRestriction.cpph:16-142), planned and checked before anything runs:
// the source to h + b in both states, its (up to four) regions accumulated
// into the destination in order, the source back to h.
struct RestrictPlan {
  TransformItem fwd[2], back[2];
  RestrictArgs r[4];
  int nr;  // non-empty regions, in the reference's order
};

static int plan_restriction(pc_pool* p, int src, int dst, int onlyOverlapped,
                            RestrictPlan& out) {
  if (!valid_slot(p, src) || !valid_slot(p, dst))
    return fail(PC_E_INVALID, "invalid slot");
  const auto& sd = p->slots[src].desc;
  const auto& dd = p->slots[dst].desc;
  if (!dd.is_virtual) return fail(PC_E_INVALID, "restriction destination must be virtual");
  const PatchDev S = make_patchdev(p->slots[src]), D = make_patchdev(p->slots[dst]);
  const int sn[2] = {sd.subdivision_factor[0], sd.subdivision_factor[1]};
  const int dn[2] = {dd.subdivision_factor[0], dd.subdivision_factor[1]};
  const int z[2] = {0, 0};
  // SWEKernel.cpp:294-297
  CHK(plan_transform(p, src, z, sn, 1, 0, out.fwd[0]));
  CHK(plan_transform(p, src, z, sn, 0, 0, out.fwd[1]));
  CHK(plan_transform(p, src, z, sn, 1, 1, out.back[0]));
  CHK(plan_transform(p, src, z, sn, 0, 1, out.back[1]));
  RestrictArgs a{};
  a.srcSlot = src; a.dstSlot = dst;
  a.srcCur = cur_of(p, src); a.dstCur = cur_of(p, dst);
  a.eps = 1e-12;
  for (int d = 0; d < 2; ++d) {
    a.srcSub[d] = sd.size[d] * (1.0 / (double)sn[d]);
    a.dstSub[d] = dd.size[d] * (1.0 / (double)dn[d]);
    a.spos[d] = sd.position[d];
    a.dpos[d] = dd.position[d];
  }
  a.dstRegion = a.dstSub[0] * a.dstSub[1];
  double st, sdt;
  slot_clock(p, src, &st, &sdt);
  a.tfOld = 1.0;
  a.tfNew = 1.0;
  if (la_greater(sdt, 0.0)) {
    const double sOld = time_uold(p, src), sNew = time_unew(p, src);
    a.tfOld = (0.0 - sOld) / (sNew - sOld);
    a.tfNew = (1.0 - sOld) / (sNew - sOld);
  }
  int roff[8], rsz[8];
  int nr;
  if (onlyOverlapped) {
    // DefaultRestriction.cpph:121-142
    const double spos[2] = {sd.position[0], sd.position[1]};
    const double ssz[2] = {sd.size[0], sd.size[1]};
    const double ssub[2] = {S.dx, S.dy};
    nr = overlapped_regions(dd.restriction_lower_bounds, dd.restriction_upper_bounds, spos,
                            ssz, ssub, sn, roff, rsz);
  } else {
    roff[0] = roff[1] = 0;
    rsz[0] = sn[0];
    rsz[1] = sn[1];
    nr = 1;
  }
  out.nr = 0;
  for (int r = 0; r < nr; ++r) {
    if (rsz[2 * r] <= 0 || rsz[2 * r + 1] <= 0) continue;
    a.rOff[0] = roff[2 * r]; a.rOff[1] = roff[2 * r + 1];
    a.rSize[0] = rsz[2 * r]; a.rSize[1] = rsz[2 * r + 1];
    map_to_patch(a.rOff, a.rSize, S, sd, D, dd, a.eps, a.dOff, a.dSize);
    for (int d = 0; d < 2; ++d)
      if (a.dOff[d] < 0 || a.dOff[d] + a.dSize[d] > dn[d])
        return fail(PC_E_INVALID, "restriction source is not inside the destination");
    if (a.dSize[0] * a.dSize[1] > 0) out.r[out.nr++] = a;
  }
  return PC_OK;
}

int pc_restrict_solution(pc_pool* p, int src, int dst, int onlyOverlapped) {
  if (!valid_slot(p, src) || !valid_slot(p, dst))
    return fail(PC_E_INVALID, "invalid slot");
  hipSetDevice(p->device);
  CHK(prepare(p));
  RestrictPlan pl;
  CHK(plan_restriction(p, src, dst, onlyOverlapped, pl));
  mark_state_changed(p);
  CHK(launch_transform(p, pl.fwd[0]));
  CHK(launch_transform(p, pl.fwd[1]));
  for (int r = 0; r < pl.nr; ++r) {
    const int n = pl.r[r].dSize[0] * pl.r[r].dSize[1];
    hipLaunchKernelGGL(k_restrict, dim3((n + 63) / 64), dim3(64), 0, p->stream, p->d_pd,
                       fields_of(p), pl.r[r]);
    HIPCHK(hipGetLastError());
  }
  CHK(launch_transform(p, pl.back[0]));
  CHK(launch_transform(p, pl.back[1]));
  return PC_OK;
}

}  // extern "C"

// ---- batched level transfers ----
//
// A batch is exactly the sequence of its per-call entry points.  Calls that
// share no patch commute, so each call goes to the first wave after every
// earlier call it shares a source or destination with; a wave runs as a few
// launches with one item per call (blockIdx.y).  Everything is checked
// before anything runs.

// Wave of every call: one more than the latest wave of an earlier call that
// touches either of its two patches.
static std::vector<int> conflict_waves(int count, const int* a, const int* b, int* nwaves) {
  std::vector<int> wave(count);
  std::map<int, int> last;
  int top = -1;
  for (int i = 0; i < count; ++i) {
    int w = 0;
    for (int s : {a[i], b[i]}) {
      auto it = last.find(s);
      if (it != last.end()) w = std::max(w, it->second + 1);
    }
    wave[i] = w;
    last[a[i]] = w;
    last[b[i]] = w;
    top = std::max(top, w);
  }
  *nwaves = top + 1;
  return wave;
}

// Host image of the staging buffer: arrays of items appended at 16-byte
// aligned offsets, uploaded in one copy.
struct Staging {
  std::vector<char> bytes;
  template <class T>
  size_t add(const std::vector<T>& v) {
    const size_t off = (bytes.size() + 15) / 16 * 16;
    bytes.resize(off + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(bytes.data() + off, v.data(), v.size() * sizeof(T));
    return off;
  }
};

// Tables gathered in the pool's pinned staging buffer (kept between calls,
// so its pages are resident) and uploaded with one DMA: add() places a
// host array (or zeros) at a 256-byte aligned offset; upload() copies
// everything to dst, synchronously.
struct PinnedTables {
  struct Piece {
    const void* src;  // nullptr: zeros
    size_t bytes, off;
  };
  std::vector<Piece> pieces;
  size_t total = 0;
  size_t add(const void* src, size_t bytes) {
    const size_t off = (total + 255) / 256 * 256;
    pieces.push_back({src, bytes, off});
    total = off + std::max<size_t>(bytes, 16);  // an empty table still gets an address
    return off;
  }
  template <class T>
  size_t add(const std::vector<T>& v) { return add(v.data(), v.size() * sizeof(T)); }
  int upload(pc_pool* p, char* dst) const {
    if (total == 0) return PC_OK;
    if (p->h_stage_cap < total) {
      HIPCHK(hipStreamSynchronize(p->stream));
      if (p->h_stage) HIPCHK(hipHostFree(p->h_stage));
      p->h_stage = nullptr;
      p->h_stage_cap = 0;
      const size_t cap = std::max<size_t>(2 * total, size_t(4) << 20);
      HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&p->h_stage), cap, hipHostMallocDefault));
      p->h_stage_cap = cap;
    }
    for (const Piece& c : pieces) {
      if (c.src)
        std::memcpy(p->h_stage + c.off, c.src, c.bytes);
      else
        std::memset(p->h_stage + c.off, 0, c.bytes);
    }
    HIPCHK(hipMemcpyAsync(dst, p->h_stage, total, hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return PC_OK;
  }
};

static int upload_staging(pc_pool* p, const Staging& st) {
  if (st.bytes.empty()) return PC_OK;
  CHK(ensure_dev(&p->d_batch, &p->batch_cap, st.bytes.size()));
  HIPCHK(hipMemcpyAsync(p->d_batch, st.bytes.data(), st.bytes.size(), hipMemcpyHostToDevice,
                        p->stream));
  return PC_OK;
}

constexpr int kMaxBatchY = 65535;  // grid y limit: items per launch

static int launch_transform_batch(pc_pool* p, size_t off, const std::vector<TransformItem>& v) {
  int most = 0;
  for (const auto& it : v) most = std::max(most, it.sx * it.sy);
  if (most == 0) return PC_OK;
  const TransformItem* d = reinterpret_cast<const TransformItem*>(p->d_batch + off);
  for (size_t b = 0; b < v.size(); b += kMaxBatchY) {
    const unsigned ny = (unsigned)std::min<size_t>(kMaxBatchY, v.size() - b);
    hipLaunchKernelGGL(k_transform_batch, dim3((most + 255) / 256, ny), dim3(256), 0,
                       p->stream, p->d_pd, fields_of(p), d + b);
    HIPCHK(hipGetLastError());
  }
  return PC_OK;
}

// Slots of a one-launch patch batch: live, each at most once, at most
// kMaxBatchY of them.
static int check_batch_slots(pc_pool* p, int count, const int* slots) {
  if (count > kMaxBatchY) return fail(PC_E_INVALID, "at most 65535 patches per call");
  std::vector<char> seen(p->slots.size(), 0);
  for (int i = 0; i < count; ++i) {
    if (!valid_slot(p, slots[i])) return fail(PC_E_INVALID, "invalid slot");
    if (seen[slots[i]]) return fail(PC_E_INVALID, "a slot appears twice in the batch");
    seen[slots[i]] = 1;
  }
  return PC_OK;
}

extern "C" {

int pc_interpolate_batch(pc_pool* p, int count, const int* dsize, const int* doff,
                         const int* src, const int* dst, const int* opts) {
  if (!p || count < 0 || (count > 0 && (!dsize || !doff || !src || !dst || !opts)))
    return fail(PC_E_INVALID, "invalid argument");
  if (count == 0) return PC_OK;
  hipSetDevice(p->device);
  CHK(prepare(p));
  std::vector<InterpPlan> pl(count);
  for (int i = 0; i < count; ++i)
    CHK(plan_interpolation(p, dsize + 2 * i, doff + 2 * i, src[i], dst[i], opts[3 * i],
                           opts[3 * i + 1], opts[3 * i + 2], pl[i]));
  int nw = 0;
  const std::vector<int> wave = conflict_waves(count, src, dst, &nw);
  std::vector<std::vector<TransformItem>> fwd(nw), back(nw);
  std::vector<std::vector<InterpArgs>> interp(nw);
  std::vector<int> most(nw, 0);
  for (int i = 0; i < count; ++i) {
    const int w = wave[i];
    for (const auto& it : pl[i].fwd)
      if (it.sx * it.sy > 0) fwd[w].push_back(it);
    if (pl[i].n > 0) {
      interp[w].push_back(pl[i].a);
      most[w] = std::max(most[w], pl[i].n);
    }
    for (const auto& it : pl[i].back)
      if (it.sx * it.sy > 0) back[w].push_back(it);
  }
  Staging st;
  std::vector<size_t> ofw(nw), oin(nw), obk(nw);
  for (int w = 0; w < nw; ++w) {
    ofw[w] = st.add(fwd[w]);
    oin[w] = st.add(interp[w]);
    obk[w] = st.add(back[w]);
  }
  mark_state_changed(p);
  CHK(upload_staging(p, st));
  for (int w = 0; w < nw; ++w) {
    CHK(launch_transform_batch(p, ofw[w], fwd[w]));
    const InterpArgs* d = reinterpret_cast<const InterpArgs*>(p->d_batch + oin[w]);
    for (size_t b = 0; b < interp[w].size(); b += kMaxBatchY) {
      const unsigned ny = (unsigned)std::min<size_t>(kMaxBatchY, interp[w].size() - b);
      hipLaunchKernelGGL(k_interpolate_batch, dim3((most[w] + 127) / 128, ny), dim3(128), 0,
                         p->stream, p->d_pd, fields_of(p), d + b);
      HIPCHK(hipGetLastError());
    }
    CHK(launch_transform_batch(p, obk[w], back[w]));
  }
  // the host image of the items lives until the copy has been consumed
  HIPCHK(hipStreamSynchronize(p->stream));
  return PC_OK;
}

int pc_restrict_batch(pc_pool* p, int count, const int* src, const int* dst,
                      int onlyOverlapped) {
  if (!p || count < 0 || (count > 0 && (!src || !dst)))
    return fail(PC_E_INVALID, "invalid argument");
  if (count == 0) return PC_OK;
  hipSetDevice(p->device);
  CHK(prepare(p));
  std::vector<RestrictPlan> pl(count);
  for (int i = 0; i < count; ++i) CHK(plan_restriction(p, src[i], dst[i], onlyOverlapped, pl[i]));
  int nw = 0;
  const std::vector<int> wave = conflict_waves(count, src, dst, &nw);
  // per wave: forward transforms, region k of every call (k = 0..3, in the
  // order of each call's regions), back transforms
  std::vector<std::vector<TransformItem>> fwd(nw), back(nw);
  std::vector<std::vector<RestrictArgs>> reg(4 * nw);
  std::vector<int> most(4 * nw, 0);
  for (int i = 0; i < count; ++i) {
    const int w = wave[i];
    for (const auto& it : pl[i].fwd)
      if (it.sx * it.sy > 0) fwd[w].push_back(it);
    for (int r = 0; r < pl[i].nr; ++r) {
      reg[4 * w + r].push_back(pl[i].r[r]);
      most[4 * w + r] = std::max(most[4 * w + r], pl[i].r[r].dSize[0] * pl[i].r[r].dSize[1]);
    }
    for (const auto& it : pl[i].back)
      if (it.sx * it.sy > 0) back[w].push_back(it);
  }
  Staging st;
  std::vector<size_t> ofw(nw), obk(nw), org(4 * nw);
  for (int w = 0; w < nw; ++w) {
    ofw[w] = st.add(fwd[w]);
    for (int r = 0; r < 4; ++r) org[4 * w + r] = st.add(reg[4 * w + r]);
    obk[w] = st.add(back[w]);
  }
  mark_state_changed(p);
  CHK(upload_staging(p, st));
  for (int w = 0; w < nw; ++w) {
    CHK(launch_transform_batch(p, ofw[w], fwd[w]));
    for (int r = 0; r < 4; ++r) {
      const auto& v = reg[4 * w + r];
      const RestrictArgs* d = reinterpret_cast<const RestrictArgs*>(p->d_batch + org[4 * w + r]);
      for (size_t b = 0; b < v.size(); b += kMaxBatchY) {
        const unsigned ny = (unsigned)std::min<size_t>(kMaxBatchY, v.size() - b);
        hipLaunchKernelGGL(k_restrict_batch, dim3((most[4 * w + r] + 63) / 64, ny), dim3(64), 0,
                           p->stream, p->d_pd, fields_of(p), d + b);
        HIPCHK(hipGetLastError());
      }
    }
    CHK(launch_transform_batch(p, obk[w], back[w]));
  }
  HIPCHK(hipStreamSynchronize(p->stream));
  return PC_OK;
}

// Patch::switchToLeaf (Patch.cpp:793-830) for a batch of virtual patches:
// the new interval [time[i], time[i] + dt[i]] and the restricted [0, 1]-basis
// values evaluated at its ends.
int pc_switch_to_leaf(pc_pool* p, int count, const int* slots, const double* time,
                      const double* dt) {
  if (!p || count < 0 || (count > 0 && (!slots || !time || !dt)))
    return fail(PC_E_INVALID, "invalid argument");
  if (count == 0) return PC_OK;
  CHK(check_batch_slots(p, count, slots));
  hipSetDevice(p->device);
  CHK(prepare(p));
  std::vector<LeafSwitch> v(count);
  int most = 0;
  for (int i = 0; i < count; ++i) {
    const auto& d = p->slots[slots[i]].desc;
    // Patch.cpp:55-76 with the [0, 1] basis (time 0, step 1)
    v[i].slot = slots[i];
    v[i].cur = cur_of(p, slots[i]);
    v[i].fOld = (time[i] - 0.0) / 1.0;
    v[i].fNew = (time[i] + dt[i] - 0.0) / 1.0;
    most = std::max(most, d.subdivision_factor[0] * d.subdivision_factor[1]);
  }
  Staging st;
  const size_t off = st.add(v);
  mark_state_changed(p);
  CHK(upload_staging(p, st));
  hipLaunchKernelGGL(k_switch_to_leaf, dim3((most + 255) / 256, (unsigned)count), dim3(256), 0,
                     p->stream, p->d_pd, fields_of(p),
                     reinterpret_cast<const LeafSwitch*>(p->d_batch + off));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(p->stream));
  for (int i = 0; i < count; ++i) {
    Slot& s = p->slots[slots[i]];
    s.desc.is_virtual = 0;
    s.desc.time = time[i];
    s.desc.timestep_size = dt[i];
    s.clock_from_scenario = false;
  }
  return PC_OK;
}

int pc_patches_clear(pc_pool* p, int count, const int* slots) {
  if (!p || count < 0 || (count > 0 && !slots)) return fail(PC_E_INVALID, "invalid argument");
  if (count == 0) return PC_OK;
  CHK(check_batch_slots(p, count, slots));
  hipSetDevice(p->device);
  CHK(prepare(p));
  std::vector<int> sl(slots, slots + count), cur(count);
  int most = 0;
  for (int i = 0; i < count; ++i) {
    const auto& d = p->slots[slots[i]].desc;
    cur[i] = cur_of(p, slots[i]);
    most = std::max(most, (d.subdivision_factor[0] + 2) * p->slots[slots[i]].stride);
  }
  Staging st;
  const size_t o1 = st.add(sl), o2 = st.add(cur);
  mark_state_changed(p);
  CHK(upload_staging(p, st));
  hipLaunchKernelGGL(k_clear_patches, dim3((most + 255) / 256, (unsigned)count), dim3(256), 0,
                     p->stream, p->d_pd, fields_of(p),
                     reinterpret_cast<const int*>(p->d_batch + o1),
                     reinterpret_cast<const int*>(p->d_batch + o2));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(p->stream));
  return PC_OK;
}

int pc_patches_set_b(pc_pool* p, int count, const int* slots, const double* b) {
  if (!p || count < 0 || (count > 0 && (!slots || !b)))
    return fail(PC_E_INVALID, "invalid argument");
  if (count == 0) return PC_OK;
  CHK(check_batch_slots(p, count, slots));
  hipSetDevice(p->device);
  CHK(prepare(p));
  std::vector<int> sl(slots, slots + count);
  std::vector<long long> soff(count);
  long long total = 0;
  int most = 0;
  for (int i = 0; i < count; ++i) {
    const auto& d = p->slots[slots[i]].desc;
    const int area = (d.subdivision_factor[0] + 2) * (d.subdivision_factor[1] + 2);
    soff[i] = total;
    total += area;
    most = std::max(most, area);
  }
  Staging st;
  const size_t o1 = st.add(sl), o2 = st.add(soff);
  const size_t o3 = st.add(std::vector<double>(b, b + total));
  mark_state_changed(p);
  CHK(upload_staging(p, st));
  hipLaunchKernelGGL(k_scatter_b, dim3((most + 255) / 256, (unsigned)count), dim3(256), 0,
                     p->stream, p->d_pd, fields_of(p),
                     reinterpret_cast<const int*>(p->d_batch + o1),
                     reinterpret_cast<const long long*>(p->d_batch + o2),
                     reinterpret_cast<const double*>(p->d_batch + o3));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(p->stream));
  return PC_OK;
}

// ---- flux correction (Numerics::computeFluxes / applyFluxCorrection) ----

// Flux arrays for slots [0, slots.size()), grown with their contents kept
// (fluxes computed for one patch are read when its neighbours are corrected).
static int ensure_flux(pc_pool* p) {
  const size_t need = std::max<size_t>(p->slots.size(), 1);
  if (p->d_flux && need <= p->flux_slots) return PC_OK;
  const size_t cap = std::max(need, 2 * p->flux_slots);
  double* n = nullptr;
  HIPCHK(hipMalloc(&n, cap * kFluxPerSlot * sizeof(double)));
  HIPCHK(hipMemsetAsync(n, 0, cap * kFluxPerSlot * sizeof(double), p->stream));
  if (p->d_flux) {
    HIPCHK(hipMemcpyAsync(n, p->d_flux, p->flux_slots * kFluxPerSlot * sizeof(double),
                          hipMemcpyDeviceToDevice, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    HIPCHK(hipFree(p->d_flux));
  }
  p->d_flux = n;
  p->flux_slots = cap;
  return PC_OK;
}

static int flux_slot_ok(pc_pool* p, int slot) {
  if (!valid_slot(p, slot)) return fail(PC_E_INVALID, "invalid slot");
  const auto& d = p->slots[slot].desc;
  if (d.subdivision_factor[0] != d.subdivision_factor[1])
    return fail(PC_E_INVALID, "flux correction needs square patches (the reference's face "
                              "iterator places every upper face at subdivisionFactor[1])");
  return PC_OK;
}

int pc_compute_fluxes(pc_pool* p, int slot) {
  CHK(flux_slot_ok(p, slot));
  hipSetDevice(p->device);
  CHK(prepare(p));
  CHK(ensure_flux(p));
  const PatchDev P = make_patchdev(p->slots[slot]);
  const int n = 2 * (P.nx + P.ny);
  hipLaunchKernelGGL(k_compute_fluxes, dim3((n + 63) / 64), dim3(64), 0, p->stream, P,
                     fields_of(p), 1 - cur_of(p, slot), p->d_flux + (size_t)slot * kFluxPerSlot);
  HIPCHK(hipGetLastError());
  return PC_OK;
}

int pc_apply_flux_correction(pc_pool* p, int src, int dst, int dimension, int direction) {
  CHK(flux_slot_ok(p, src));
  CHK(flux_slot_ok(p, dst));
  if ((dimension != 0 && dimension != 1) || (direction != -1 && direction != 1))
    return fail(PC_E_INVALID, "dimension must be 0 or 1 and direction -1 or +1");
  hipSetDevice(p->device);
  CHK(prepare(p));
  CHK(ensure_flux(p));
  const auto& sd = p->slots[src].desc;
  const auto& dd = p->slots[dst].desc;
  // HyperplaneRegion::getInterfaceRegion: only its dimension is used, the
  // last coordinate in which the positions differ
  int rdim = -1;
  for (int d = 0; d < 2; ++d)
    if (la_smaller(sd.position[d], dd.position[d]) || la_greater(sd.position[d], dd.position[d]))
      rdim = d;
  if (rdim < 0)
    return fail(PC_E_INVALID, "source and destination share their position (the reference's "
                              "interface region is undefined then)");
  double st, sdt, dt0, ddt;
  slot_clock(p, src, &st, &sdt);
  slot_clock(p, dst, &dt0, &ddt);
  // DefaultFluxCorrection.cpph:96-100 (std::max / std::min)
  const double lower = (st < dt0) ? dt0 : st;
  const double a = st + sdt, b = dt0 + ddt;
  const double upper = (b < a) ? b : a;
  const double overlap = upper - lower;
  FluxGeom S{}, D{};
  const PatchDev SP = make_patchdev(p->slots[src]), DP = make_patchdev(p->slots[dst]);
  for (int d = 0; d < 2; ++d) {
    S.pos[d] = sd.position[d];
    D.pos[d] = dd.position[d];
  }
  S.sub[0] = SP.dx; S.sub[1] = SP.dy; S.nx = SP.nx; S.ny = SP.ny;
  D.sub[0] = DP.dx; D.sub[1] = DP.dy; D.nx = DP.nx; D.ny = DP.ny;
  mark_state_changed(p);
  hipLaunchKernelGGL(k_apply_flux_correction, dim3(1), dim3(64), 0, p->stream, S, SP,
                     p->d_flux + (size_t)src * kFluxPerSlot, D, DP,
                     p->d_flux + (size_t)dst * kFluxPerSlot, fields_of(p), cur_of(p, dst),
                     dimension, direction, rdim, overlap);
  HIPCHK(hipGetLastError());
  return PC_OK;
}

int pc_patch_fluxes(pc_pool* p, int slot, int to_device, double* fluxes) {
  CHK(flux_slot_ok(p, slot));
  if (!fluxes) return fail(PC_E_INVALID, "fluxes is NULL");
  hipSetDevice(p->device);
  CHK(ensure_flux(p));
  const auto& d = p->slots[slot].desc;
  const size_t n = (size_t)2 * 3 * (d.subdivision_factor[0] + d.subdivision_factor[1]);
  double* dev = p->d_flux + (size_t)slot * kFluxPerSlot;
  if (to_device)
    HIPCHK(hipMemcpyAsync(dev, fluxes, n * sizeof(double), hipMemcpyHostToDevice, p->stream));
  else
    HIPCHK(hipMemcpyAsync(fluxes, dev, n * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return PC_OK;
}

int pc_post_process_restriction(pc_pool* p, int dst, int onlyOverlapped) {
  (void)onlyOverlapped;
  if (!valid_slot(p, dst)) return fail(PC_E_INVALID, "invalid slot");
  hipSetDevice(p->device);
  CHK(prepare(p));
  mark_state_changed(p);
  const int z[2] = {0, 0};
  const int dn[2] = {p->slots[dst].desc.subdivision_factor[0],
                     p->slots[dst].desc.subdivision_factor[1]};
  CHK(transform(p, dst, z, dn, 1, 1));
  CHK(transform(p, dst, z, dn, 0, 1));
  return PC_OK;
}

int pc_copy_unew_to_uold(pc_pool* p, int slot) {
  if (!valid_slot(p, slot)) return fail(PC_E_INVALID, "invalid slot");
  hipSetDevice(p->device);
  CHK(prepare(p));
  mark_state_changed(p);
  hipLaunchKernelGGL(k_copy_swap, dim3(4), dim3(256), 0, p->stream, p->d_pd, slot,
                     fields_of(p), cur_of(p, slot), 0);
  HIPCHK(hipGetLastError());
  return PC_OK;
}

int pc_swap_unew_and_uold(pc_pool* p, int slot) {
  if (!valid_slot(p, slot)) return fail(PC_E_INVALID, "invalid slot");
  hipSetDevice(p->device);
  CHK(prepare(p));
  mark_state_changed(p);
  hipLaunchKernelGGL(k_copy_swap, dim3(4), dim3(256), 0, p->stream, p->d_pd, slot,
                     fields_of(p), cur_of(p, slot), 1);
  HIPCHK(hipGetLastError());
  return PC_OK;
}

static int refine_flags_slots(pc_pool* p, const std::vector<int>& slots, double thr,
                              double minW, double maxW, int32_t* flags, double* demanded) {
  const int n = (int)slots.size();
  if (n == 0) return PC_OK;
  CHK(ensure_dev(&p->d_flags, &p->flags_cap, (size_t)n + slots.size()));
  CHK(ensure_dev(&p->d_aux, &p->aux_cap, 2 * (size_t)n));
  int* d_sl = p->d_flags + n;
  HIPCHK(hipMemcpyAsync(d_sl, slots.data(), n * sizeof(int), hipMemcpyHostToDevice,
                        p->stream));
  hipLaunchKernelGGL(k_refine_flags, dim3(n), dim3(kRefineThreads), 0, p->stream, d_sl, n, p->d_pd,
                     fields_of(p), p->d_scn, thr, minW, maxW, p->d_flags, p->d_aux,
                     static_cast<double*>(nullptr), 0ull);
  HIPCHK(hipGetLastError());
  if (flags)
    HIPCHK(hipMemcpyAsync(flags, p->d_flags, n * sizeof(int), hipMemcpyDeviceToHost,
                          p->stream));
  if (demanded)
    HIPCHK(hipMemcpyAsync(demanded, p->d_aux, 2 * n * sizeof(double), hipMemcpyDeviceToHost,
                          p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return PC_OK;
}

int pc_demanded_mesh_width(pc_pool* p, int slot, double thr, double minW, double maxW,
                           double out[2]) {
  if (!valid_slot(p, slot) || !out) return fail(PC_E_INVALID, "invalid slot");
  hipSetDevice(p->device);
  CHK(prepare(p));
  // one launch: the slot through the identity table, the result straight
  // into pinned host memory, the completion word after it
  const unsigned long long seq = ++p->call_seq;
  int* const flag_host = reinterpret_cast<int*>(p->d_result_host + 6);
  hipLaunchKernelGGL(k_refine_flags, dim3(1), dim3(kRefineThreads), 0, p->stream, p->d_iota + slot, 1,
                     p->d_pd, fields_of(p), p->d_scn, thr, minW, maxW, flag_host,
                     p->d_result_host, p->d_result_host, seq);
  HIPCHK(hipGetLastError());
  CHK(wait_call(p, seq));
  out[0] = p->h_result[0];
  out[1] = p->h_result[1];
  return PC_OK;
}

int pc_pool_slot_capacity(pc_pool* p, int* n) {
  if (!p || !n) return fail(PC_E_INVALID, "NULL argument");
  *n = (int)p->slots.size();
  return PC_OK;
}

int pc_refinement_flags(pc_pool* p, double thr, double minW, double maxW, int32_t* flags,
                        int cap) {
  if (!p || !flags) return fail(PC_E_INVALID, "NULL argument");
  if (cap < (int)p->slots.size())
    return fail(PC_E_INVALID, "flags capacity below pc_pool_slot_capacity");
  hipSetDevice(p->device);
  CHK(prepare(p));
  // flags are returned per slot index (dead slots get -1)
  std::vector<int32_t> f(p->h_live.size());
  CHK(refine_flags_slots(p, p->h_live, thr, minW, maxW, f.data(), nullptr));
  for (int i = 0; i < cap; ++i) flags[i] = -1;
  for (size_t k = 0; k < p->h_live.size(); ++k) flags[p->h_live[k]] = f[k];
  return PC_OK;
}

int pc_scenario_mass(pc_pool* p, int scenario, double* out) {
  if (!p || !out || scenario < 0) return fail(PC_E_INVALID, "invalid argument");
  hipSetDevice(p->device);
  CHK(prepare(p));
  std::vector<int> sl;
  for (int s : p->h_live)
    if (p->slots[s].desc.scenario == scenario) sl.push_back(s);
  const int n = (int)sl.size();
  *out = 0.0;
  if (!n) return PC_OK;
  CHK(ensure_dev(&p->d_flags, &p->flags_cap, (size_t)n));
  CHK(ensure_dev(&p->d_aux, &p->aux_cap, (size_t)n));
  HIPCHK(hipMemcpyAsync(p->d_flags, sl.data(), n * sizeof(int), hipMemcpyHostToDevice,
                        p->stream));
  hipLaunchKernelGGL(k_patch_sums, dim3(n), dim3(64), 0, p->stream, p->d_flags, n, p->d_pd,
                     fields_of(p), p->d_scn, p->d_aux);
  HIPCHK(hipGetLastError());
  std::vector<double> sums(n);
  HIPCHK(hipMemcpyAsync(sums.data(), p->d_aux, n * sizeof(double), hipMemcpyDeviceToHost,
                        p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  // Neumaier summation of the per-patch volumes
  double s = 0.0, c = 0.0;
  for (int k = 0; k < n; ++k) {
    const PatchDev d = make_patchdev(p->slots[sl[k]]);
    const double v = sums[k] * d.dx * d.dy;
    const double t = s + v;
    c += std::fabs(s) >= std::fabs(v) ? (s - t) + v : (v - t) + s;
    s = t;
  }
  *out = s + c;
  return PC_OK;
}

int pc_set_ghost_strips(pc_pool* p, const pc_ghost_strip* strips, int n) {
  if (!p || n < 0 || (n > 0 && !strips)) return fail(PC_E_INVALID, "invalid argument");
  hipSetDevice(p->device);
  std::vector<GhostStrip> v(n);
  for (int i = 0; i < n; ++i) {
    const pc_ghost_strip& g = strips[i];
    if (!valid_slot(p, g.dst_slot) || !valid_slot(p, g.src_slot) || g.length < 0 ||
        g.negate < 0 || g.negate > 3)
      return fail(PC_E_INVALID, "invalid ghost strip");
    const auto& dd = p->slots[g.dst_slot].desc;
    const auto& sd = p->slots[g.src_slot].desc;
    if (dd.scenario != sd.scenario)
      return fail(PC_E_INVALID, "ghost strips must stay within one scenario");
    const long long dcells = p->slots[g.dst_slot].cells, scells = p->slots[g.src_slot].cells;
    if (g.length > 0) {
      const long long d0 = g.dst_cell, d1 = g.dst_cell + (long long)(g.length - 1) * g.dst_stride;
      const long long s0 = g.src_cell, s1 = g.src_cell + (long long)(g.length - 1) * g.src_stride;
      if (std::min(d0, d1) < 0 || std::max(d0, d1) >= dcells || std::min(s0, s1) < 0 ||
          std::max(s0, s1) >= scells)
        return fail(PC_E_INVALID, "ghost strip leaves its tile");
    }
    v[i] = GhostStrip{g.dst_slot, g.dst_cell, g.dst_stride, g.src_slot, g.src_cell,
                      g.src_stride, g.length, g.negate};
  }
  std::stable_sort(v.begin(), v.end(),
                   [](const GhostStrip& a, const GhostStrip& b) { return a.dst_slot < b.dst_slot; });
  p->h_strips = v;
  p->strip_begin.assign(p->slots.size() + 1, 0);
  for (auto& g : v) p->strip_begin[g.dst_slot + 1]++;
  for (size_t i = 1; i < p->strip_begin.size(); ++i) p->strip_begin[i] += p->strip_begin[i - 1];
  CHK(ensure_dev(&p->d_strips, &p->strips_cap, std::max<size_t>(v.size(), 1)));
  if (n > 0)
    HIPCHK(hipMemcpyAsync(p->d_strips, v.data(), v.size() * sizeof(GhostStrip),
                          hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  p->nstrips = n;
  CHK(invalidate_graph(p));
  return PC_OK;
}

namespace {
// Symbolic source of one ghost cell: the value it holds after the
// reference's fill sequence, as (slot, cell x, cell y, negate mask).
struct GhostSrc {
  int slot = -1;  // -1: not written (keeps its old value)
  int x = 0, y = 0;
  int neg = 0;
  bool ghost = false;  // refers to a ghost cell of its own patch (stale read)
};
}  // namespace

int pc_grid_strips(pc_pool* p, int px, int py, const int* slot_of, pc_ghost_strip* out,
                   int cap, int* n_out) {
  if (!p || px <= 0 || py <= 0 || !slot_of || !n_out)
    return fail(PC_E_INVALID, "invalid argument");
  int nx = -1, ny = -1;
  for (int k = 0; k < px * py; ++k) {
    if (slot_of[k] < 0) continue;
    if (!valid_slot(p, slot_of[k])) return fail(PC_E_INVALID, "invalid slot in grid");
    const auto& d = p->slots[slot_of[k]].desc;
    if (nx < 0) {
      nx = d.subdivision_factor[0];
      ny = d.subdivision_factor[1];
    } else if (nx != d.subdivision_factor[0] || ny != d.subdivision_factor[1]) {
      return fail(PC_E_INVALID, "grid patches must share their subdivision factor");
    }
  }
  auto at = [&](int i, int j) -> int {
    if (i < 0 || j < 0 || i >= px || j >= py) return -1;
    return slot_of[i * py + j];
  };
  std::vector<pc_ghost_strip> res;
  const int sy = ny + 2;          // rows of the symbolic ghost ring below
  const int ts = tile_stride(ny);  // rows of the tiles (strip cells)
  auto cell = [&](int x, int y) { return (x + 1) * ts + (y + 1); };
  for (int i = 0; i < px; ++i)
    for (int j = 0; j < py; ++j) {
      const int me = at(i, j);
      if (me < 0) continue;
      // ghost ring indexed by tile coordinates
      std::vector<GhostSrc> g((size_t)(nx + 2) * (ny + 2));
      auto G = [&](int x, int y) -> GhostSrc& { return g[(size_t)(x + 1) * sy + (y + 1)]; };
      auto is_ghost = [&](int x, int y) { return x < 0 || y < 0 || x >= nx || y >= ny; };
      // phase 1: same-level transfers from face and diagonal neighbours
      for (int di = -1; di <= 1; ++di)
        for (int dj = -1; dj <= 1; ++dj) {
          if (!di && !dj) continue;
          const int nb = at(i + di, j + dj);
          if (nb < 0) continue;
          const int x0 = di < 0 ? -1 : (di > 0 ? nx : 0), x1 = di ? x0 : nx - 1;
          const int y0 = dj < 0 ? -1 : (dj > 0 ? ny : 0), y1 = dj ? y0 : ny - 1;
          for (int x = x0; x <= x1; ++x)
            for (int y = y0; y <= y1; ++y) {
              GhostSrc s;
              s.slot = nb;
              s.x = di < 0 ? nx - 1 : (di > 0 ? 0 : x);
              s.y = dj < 0 ? ny - 1 : (dj > 0 ? 0 : y);
              G(x, y) = s;
            }
        }
      // phase 2: walls left, top, right, bottom (SolveTimestep.cpp:94-131)
      const int walls[4][2] = {{0, 0}, {1, 1}, {0, 1}, {1, 0}};
      for (auto& w : walls) {
        const int dim = w[0], up = w[1];
        const bool wall = dim == 0 ? at(i + (up ? 1 : -1), j) < 0 : at(i, j + (up ? 1 : -1)) < 0;
        if (!wall) continue;
        const int len = (dim == 0 ? ny : nx) + 2;
        for (int k = -1; k < len - 1; ++k) {
          int dx, dy, sx, sy_;
          if (dim == 0) {
            dx = up ? nx : -1; sx = dx + (up ? -1 : 1); dy = sy_ = k;
          } else {
            dy = up ? ny : -1; sy_ = dy + (up ? -1 : 1); dx = sx = k;
          }
          GhostSrc s;
          if (!is_ghost(sx, sy_)) {
            s.slot = me; s.x = sx; s.y = sy_;
          } else if (G(sx, sy_).slot >= 0) {
            s = G(sx, sy_);
          } else {
            s.slot = me; s.x = sx; s.y = sy_; s.ghost = true;
          }
          s.neg ^= (dim == 0 ? 1 : 2);
          G(dx, dy) = s;
        }
      }
      // emit the ring in a fixed order: bottom row, top row, then columns
      std::vector<std::pair<int, int>> ring;
      for (int x = -1; x <= nx; ++x) ring.push_back({x, -1});
      for (int x = -1; x <= nx; ++x) ring.push_back({x, ny});
      for (int y = 0; y < ny; ++y) ring.push_back({-1, y});
      for (int y = 0; y < ny; ++y) ring.push_back({nx, y});
      for (auto& c : ring) {
        const GhostSrc& s = G(c.first, c.second);
        if (s.slot < 0 || s.ghost) continue;  // unwritten, or a stale-only read
        pc_ghost_strip e;
        e.dst_slot = me;
        e.dst_cell = cell(c.first, c.second);
        e.src_slot = s.slot;
        e.src_cell = cell(s.x, s.y);
        e.length = 1;
        e.dst_stride = 0;
        e.src_stride = 0;
        e.negate = s.neg;
        // merge with the previous strip when strides continue
        if (!res.empty()) {
          pc_ghost_strip& b = res.back();
          if (b.dst_slot == e.dst_slot && b.src_slot == e.src_slot && b.negate == e.negate) {
            const int ds = e.dst_cell - (b.dst_cell + (b.length - 1) * b.dst_stride);
            const int ss = e.src_cell - (b.src_cell + (b.length - 1) * b.src_stride);
            if (b.length == 1 && ds != 0) {
              b.dst_stride = ds;
              b.src_stride = ss;
              b.length = 2;
              continue;
            }
            if (b.length > 1 && ds == b.dst_stride && ss == b.src_stride) {
              b.length += 1;
              continue;
            }
          }
        }
        res.push_back(e);
      }
    }
  *n_out = (int)res.size();
  if (out) {
    if (cap < (int)res.size()) return fail(PC_E_INVALID, "output capacity too small");
    std::copy(res.begin(), res.end(), out);
  }
  return PC_OK;
}

int pc_scenario_set(pc_pool* p, int scenario, double time, double dt, double end_time) {
  if (!p || scenario < 0) return fail(PC_E_INVALID, "invalid scenario");
  CHK(sync_scn_host(p));
  if ((int)p->scn.size() <= scenario) {
    size_t old = p->scn.size();
    p->scn.resize(scenario + 1);
    for (size_t k = old; k < p->scn.size(); ++k) {
      ScnDev z{};
      z.end_time = INFINITY;
      z.minbits = 0x7FEFFFFFFFFFFFFFull;  // DBL_MAX
      p->scn[k] = z;
    }
  }
  p->scn[scenario].time = time;
  p->scn[scenario].dt = dt;
  p->scn[scenario].end_time = end_time;
  p->scn_dev_dirty = true;
  return PC_OK;
}

int pc_scenario_get(pc_pool* p, int scenario, double* time, double* dt, double* est,
                    long long* steps) {
  if (!p || scenario < 0) return fail(PC_E_INVALID, "invalid scenario");
  CHK(sync_scn_host(p));
  if (scenario >= (int)p->scn.size()) return fail(PC_E_INVALID, "unknown scenario");
  const ScnDev& S = p->scn[scenario];
  if (time) *time = S.time;
  if (dt) *dt = S.dt;
  if (est) *est = S.est;
  if (steps) *steps = S.steps;
  return PC_OK;
}

int pc_set_prune_slack(pc_pool* p, double slack) {
  if (!p || !(slack > 0.0) || !(slack < 1e3)) return fail(PC_E_INVALID, "slack must be in (0, 1e3)");
  p->prune_slack = slack;
  CHK(invalidate_graph(p));
  return PC_OK;
}

int pc_set_wave_propagation_solver(pc_pool* p, int solver) {
  if (!p) return fail(PC_E_INVALID, "pool is NULL");
  if (solver != 1 && solver != 2 && solver != 4)
    return fail(PC_E_INVALID, "solver must be 1 (f-wave), 2 (augmented Riemann) or 4 (f-wave)");
  const int s = solver == 2 ? 1 : 0;
  if (s != p->solver) {
    hipSetDevice(p->device);
    CHK(sync_scn_host(p));
    p->solver = s;
    p->fused_ready = false;
    for (auto& S : p->scn) S.est = 0.0;  // pruning thresholds came from the other solver
    p->scn_dev_dirty = true;
    CHK(invalidate_graph(p));
  }
  return PC_OK;
}

int pc_set_launch_config(pc_pool* p, int persistent_wgs, int fused, int cap) {
  if (!p || persistent_wgs < 0 || cap < 0) return fail(PC_E_INVALID, "invalid config");
  p->persistent_wgs = persistent_wgs;
  if (fused > 2) return fail(PC_E_INVALID, "fused must be 0, 1 or 2");
  if ((fused != 0) != (p->fused != 0)) p->fused_ready = false;
  p->fused = fused;
  if (cap != p->task_cap) p->topo_dirty = true;
  p->task_cap = cap;
  CHK(invalidate_graph(p));
  return PC_OK;
}

int pc_set_row_segments(pc_pool* p, int segments) {
  if (!p) return fail(PC_E_INVALID, "pool is NULL");
  if (segments != 0 && segments != 1 && segments != 2 && segments != 4)
    return fail(PC_E_INVALID, "row segments must be 0 (automatic), 1, 2 or 4");
  if (segments != p->row_segments) p->topo_dirty = true;
  p->row_segments = segments;
  CHK(invalidate_graph(p));
  return PC_OK;
}

static void mark_batched_clock(pc_pool* p) {
  for (int s : p->h_live) p->slots[s].clock_from_scenario = true;
}

int pc_step_all(pc_pool* p, int k) {
  if (!p || k < 0) return fail(PC_E_INVALID, "invalid argument");
  hipSetDevice(p->device);
  CHK(prepare_batched(p));
  // one event pair around every update launch: the update kernel's mean
  // duration over exactly these k steps
  while (p->upd_ev.size() < 2 * (size_t)k) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    p->upd_ev.push_back(e);
  }
  HIPCHK(hipEventRecord(p->ev0, p->stream));
  for (int i = 0; i < k; ++i) CHK(launch_step(p, true, p->upd_ev[2 * i], p->upd_ev[2 * i + 1]));
  HIPCHK(hipEventRecord(p->ev1, p->stream));
  HIPCHK(hipEventSynchronize(p->ev1));
  HIPCHK(hipEventElapsedTime(&p->last_total_ms, p->ev0, p->ev1));
  p->last_update_ms = 0.f;
  if (k > 0 && p->ntasks > 0) {
    double sum = 0.0;
    for (int i = 0; i < k; ++i) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, p->upd_ev[2 * i], p->upd_ev[2 * i + 1]));
      sum += ms;
    }
    p->last_update_ms = (float)(sum / k);
  }
  p->scn_host_stale = true;
  mark_batched_clock(p);
  return PC_OK;
}

int pc_step_all_per_patch(pc_pool* p, int k) {
  if (!p || k < 0) return fail(PC_E_INVALID, "invalid argument");
  hipSetDevice(p->device);
  const int fused = p->fused;
  p->fused = 0;  // the plain form computes every patch's CFL step separately
  int rc = prepare_batched(p);
  p->fused = fused;
  if (rc != PC_OK) return rc;
  // per-patch tasks (one patch per launch)
  std::vector<Task> single;
  for (int s : p->h_live) {
    const auto& d = p->slots[s].desc;
    single.push_back(Task{0, 1, d.subdivision_factor[0], d.subdivision_factor[1]});
  }
  const int n = (int)single.size();
  CHK(ensure_dev(&p->d_single, &p->single_cap, std::max<size_t>(n, 1)));
  CHK(ensure_dev(&p->d_single_slot, &p->single_slot_cap, std::max<size_t>(n, 1)));
  Task* const d_single = p->d_single;
  int* const d_sl = p->d_single_slot;
  for (int i = 0; i < n; ++i) single[i].first = i;
  if (n) {
    HIPCHK(hipMemcpyAsync(d_single, single.data(), n * sizeof(Task), hipMemcpyHostToDevice,
                          p->stream));
    HIPCHK(hipMemcpyAsync(d_sl, p->h_live.data(), n * sizeof(int), hipMemcpyHostToDevice,
                          p->stream));
  }
  Fields F = fields_of(p);
  EdgeBufs E = edges_of(p);
  const SideDev* sides = p->sides_any ? p->d_sides : nullptr;
  const int nscn = (int)p->scn.size();
  HIPCHK(hipEventRecord(p->ev0, p->stream));
  for (int it = 0; it < k; ++it) {
    // ghost fill of tile-side patches, one launch per destination patch
    for (int i = 0; i < n; ++i) {
      const int s = p->h_live[i];
      if ((size_t)s + 1 >= p->strip_begin.size()) continue;
      const int b = p->strip_begin[s], e = p->strip_begin[s + 1];
      if (e > b)
        hipLaunchKernelGGL(k_gather, dim3((e - b + 3) / 4), dim3(256), 0, p->stream,
                           p->d_strips + b, e - b, p->d_pd, F, p->d_scn);
    }
    // per-patch CFL estimate, one launch per patch
    for (int i = 0; i < n; ++i)
      launch_speed(p, dim3(1), d_single + i, 1, d_sl, F, sides, E, p->d_smax, nullptr);
    if (p->nentries > 0)
      hipLaunchKernelGGL(k_dt_partial, dim3((p->nentries + 1023) / 1024), dim3(256), 0,
                         p->stream, p->d_scn, p->d_plist_scn, p->d_scn_of, p->nentries,
                         p->d_pd, p->d_smax, p->d_est, nullptr, nullptr);
    if (nscn > 0)
      hipLaunchKernelGGL(k_dt_final, dim3((nscn + 63) / 64), dim3(64), 0, p->stream,
                         p->d_scn, nscn, p->d_counter, 0, p->d_redo, 1.0);
    // per-patch update, one launch per patch
    for (int i = 0; i < n; ++i)
      launch_update_plain(p, dim3(1), d_single + i, 1, d_sl, F, sides, E, 0);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipEventRecord(p->ev1, p->stream));
  HIPCHK(hipEventSynchronize(p->ev1));
  HIPCHK(hipEventElapsedTime(&p->last_total_ms, p->ev0, p->ev1));
  p->last_update_ms = 0.f;
  p->scn_host_stale = true;
  p->fused_ready = false;
  mark_batched_clock(p);
  return PC_OK;
}

int pc_set_grid_sides(pc_pool* p, int px, int py, const int* slot_of) {
  if (!p || px <= 0 || py <= 0 || !slot_of) return fail(PC_E_INVALID, "invalid argument");
  auto at = [&](int i, int j) -> int {
    if (i < 0 || j < 0 || i >= px || j >= py) return -1;
    return slot_of[i * py + j];
  };
  for (int k = 0; k < px * py; ++k)
    if (slot_of[k] >= 0 && !valid_slot(p, slot_of[k]))
      return fail(PC_E_INVALID, "invalid slot in grid");
  if (p->h_sides.size() < p->slots.size()) p->h_sides.resize(p->slots.size());
  for (int i = 0; i < px; ++i)
    for (int j = 0; j < py; ++j) {
      const int me = at(i, j);
      if (me < 0) continue;
      const auto& d = p->slots[me].desc;
      const int nb[4] = {at(i - 1, j), at(i + 1, j), at(i, j - 1), at(i, j + 1)};
      SideDev sd{};
      for (int k = 0; k < 4; ++k) {
        if (nb[k] >= 0) {
          const auto& e = p->slots[nb[k]].desc;
          if (e.subdivision_factor[0] != d.subdivision_factor[0] ||
              e.subdivision_factor[1] != d.subdivision_factor[1] || e.scenario != d.scenario)
            return fail(PC_E_INVALID, "grid neighbours must share size and scenario");
          sd.kind[k] = kSideNeighbour;
          sd.src[k] = nb[k];
        } else {
          sd.kind[k] = kSideWall;
          sd.src[k] = me;
        }
      }
      p->h_sides[me] = sd;
    }
  p->topo_dirty = true;
  return PC_OK;
}

int pc_clear_sides(pc_pool* p) {
  if (!p) return fail(PC_E_INVALID, "pool is NULL");
  p->h_sides.assign(p->slots.size(), SideDev{});
  p->topo_dirty = true;
  return PC_OK;
}

int pc_graph_capture(pc_pool* p) {
  if (!p) return fail(PC_E_INVALID, "pool is NULL");
  hipSetDevice(p->device);
  CHK(prepare_batched(p));
  CHK(invalidate_graph(p));
  HIPCHK(hipStreamSynchronize(p->stream));
  HIPCHK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
  int rc = launch_step(p, false);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(p->stream, &g);
  if (rc != PC_OK) return rc;
  if (e != hipSuccess) return fail(PC_E_HIP, std::string("capture: ") + hipGetErrorString(e));
  p->graph = g;
  HIPCHK(hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0));
  p->graph_valid = true;
  return PC_OK;
}

int pc_graph_replay(pc_pool* p, int k) {
  if (!p || k < 0) return fail(PC_E_INVALID, "invalid argument");
  hipSetDevice(p->device);
  if (p->topo_dirty || !p->graph_valid || p->scn_dev_dirty ||
      (p->sides_any && p->edges_dirty) || (fused_mode(p) && p->sides_any && !p->fused_ready))
    return fail(PC_E_INVALID, "no valid captured graph (grid, scenarios or state changed)");
  CHK(sync_scn_host(p));
  HIPCHK(hipEventRecord(p->ev0, p->stream));
  for (int i = 0; i < k; ++i) HIPCHK(hipGraphLaunch(p->gexec, p->stream));
  HIPCHK(hipEventRecord(p->ev1, p->stream));
  HIPCHK(hipEventSynchronize(p->ev1));
  HIPCHK(hipEventElapsedTime(&p->last_total_ms, p->ev0, p->ev1));
  p->scn_host_stale = true;
  mark_batched_clock(p);
  return PC_OK;
}

int pc_last_timing(pc_pool* p, double* total_ms, double* update_ms) {
  if (!p) return fail(PC_E_INVALID, "pool is NULL");
  if (total_ms) *total_ms = p->last_total_ms;
  if (update_ms) *update_ms = p->last_update_ms;
  return PC_OK;
}

int pc_edge_batch(pc_pool* p, long long n, const double* in, double* out) {
  if (!p || n < 0 || (n > 0 && (!in || !out))) return fail(PC_E_INVALID, "invalid argument");
  if (n == 0) return PC_OK;
  hipSetDevice(p->device);
  CHK(ensure_dev(&p->d_edge_io, &p->edge_io_cap, (size_t)n * 11));
  double* const din = p->d_edge_io;
  double* const dout = din + (size_t)n * 6;
  HIPCHK(hipMemcpyAsync(din, in, (size_t)n * 6 * sizeof(double), hipMemcpyHostToDevice,
                        p->stream));
  const long long blocks = std::min<long long>((n + 255) / 256, 65536);
  if (p->solver)
    hipLaunchKernelGGL(k_edge_batch<1>, dim3((unsigned)blocks), dim3(256), 0, p->stream, n,
                       din, dout);
  else
    hipLaunchKernelGGL(k_edge_batch<0>, dim3((unsigned)blocks), dim3(256), 0, p->stream, n,
                       din, dout);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, dout, (size_t)n * 5 * sizeof(double), hipMemcpyDeviceToHost,
                        p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return PC_OK;
}

int pc_edge_batch_static(pc_pool* p, long long n, const double* in, double* out) {
  if (!p || n < 0 || (n > 0 && (!in || !out))) return fail(PC_E_INVALID, "invalid argument");
  if (n > kStaticEdgeMax)
    return fail(PC_E_INVALID, "pc_edge_batch_static: at most 65536 edges per call "
                              "(the size of its static staging array)");
  if (n == 0) return PC_OK;
  hipSetDevice(p->device);
  double* const base = g_scratch_ptr[p->device];
  if (!base) return fail(PC_E_HIP, "pc_edge_batch_static: staging not set up on this device");
  double* const din = base;
  double* const dout = base + kStaticEdgeMax * 6;
  HIPCHK(hipMemcpyAsync(din, in, (size_t)n * 6 * sizeof(double), hipMemcpyHostToDevice,
                        p->stream));
  const long long blocks = std::min<long long>((n + 255) / 256, 65536);
  if (p->solver)
    hipLaunchKernelGGL(k_edge_batch<1>, dim3((unsigned)blocks), dim3(256), 0, p->stream, n,
                       din, dout);
  else
    hipLaunchKernelGGL(k_edge_batch<0>, dim3((unsigned)blocks), dim3(256), 0, p->stream, n,
                       din, dout);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, dout, (size_t)n * 5 * sizeof(double), hipMemcpyDeviceToHost,
                        p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return PC_OK;
}

int pc_copy_probe(pc_pool* p, long long bytes, int reps, double* best_gbs) {
  if (!p || bytes < 16 || reps < 1 || !best_gbs) return fail(PC_E_INVALID, "invalid argument");
  hipSetDevice(p->device);
  const long long n2 = bytes / 16;
  double2 *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, (size_t)n2 * 16));
  if (hipMalloc(&b, (size_t)n2 * 16) != hipSuccess) {
    hipFree(a);
    return fail(PC_E_OOM, "copy probe: hipMalloc failed");
  }
  const long long blocks = (n2 + 255) / 256;  // one element per thread
  float best = 1e30f;
  hipError_t e = hipSuccess;
  for (int r = 0; r <= reps && e == hipSuccess; ++r) {  // r = 0 is an untimed warm-up
    if ((e = hipEventRecord(p->ev0, p->stream)) != hipSuccess) break;
    hipLaunchKernelGGL(k_copy_probe, dim3((unsigned)blocks), dim3(256), 0, p->stream, n2,
                       (const double2*)(r & 1 ? b : a), r & 1 ? a : b);
    if ((e = hipGetLastError()) != hipSuccess) break;
    if ((e = hipEventRecord(p->ev1, p->stream)) != hipSuccess) break;
    if ((e = hipEventSynchronize(p->ev1)) != hipSuccess) break;
    float ms = 0.f;
    if ((e = hipEventElapsedTime(&ms, p->ev0, p->ev1)) != hipSuccess) break;
    if (r > 0) best = std::min(best, ms);
  }
  hipFree(a);
  hipFree(b);
  if (e != hipSuccess) return fail(PC_E_HIP, std::string("copy probe: ") + hipGetErrorString(e));
  *best_gbs = 2.0 * (double)n2 * 16 / (best * 1e-3) / 1e9;
  return PC_OK;
}

int pc_stream_probe(pc_pool* p, long long cells, int reps, double* best_gbs) {
  if (!p || cells < 1 || reps < 1 || !best_gbs) return fail(PC_E_INVALID, "invalid argument");
  hipSetDevice(p->device);
  double* a[7] = {};
  for (int k = 0; k < 7; ++k) {
    if (hipMalloc(&a[k], (size_t)cells * sizeof(double)) != hipSuccess) {
      for (int j = 0; j < k; ++j) hipFree(a[j]);
      return fail(PC_E_OOM, "stream probe: hipMalloc failed");
    }
  }
  hipError_t e = hipSuccess;
  for (int k = 0; k < 7 && e == hipSuccess; ++k)
    e = hipMemsetAsync(a[k], 0, (size_t)cells * sizeof(double), p->stream);
  const long long blocks = (cells / 2 + 1 + 255) / 256;  // one pair per thread
  float best = 1e30f;
  for (int r = 0; r <= reps && e == hipSuccess; ++r) {  // r = 0 is an untimed warm-up
    if ((e = hipEventRecord(p->ev0, p->stream)) != hipSuccess) break;
    hipLaunchKernelGGL(k_stream_probe, dim3((unsigned)blocks), dim3(256), 0, p->stream, cells,
                       a[0], a[1], a[2], a[3], a[4], a[5], a[6]);
    if ((e = hipGetLastError()) != hipSuccess) break;
    if ((e = hipEventRecord(p->ev1, p->stream)) != hipSuccess) break;
    if ((e = hipEventSynchronize(p->ev1)) != hipSuccess) break;
    float ms = 0.f;
    if ((e = hipEventElapsedTime(&ms, p->ev0, p->ev1)) != hipSuccess) break;
    if (r > 0) best = std::min(best, ms);
  }
  hipStreamSynchronize(p->stream);
  for (int k = 0; k < 7; ++k) hipFree(a[k]);
  if (e != hipSuccess) return fail(PC_E_HIP, std::string("stream probe: ") + hipGetErrorString(e));
  *best_gbs = 56.0 * (double)cells / (best * 1e-3) / 1e9;
  return PC_OK;
}

int pc_march_probe(pc_pool* p, int reps, double* best_gbs) {
  if (!p || reps < 1 || !best_gbs) return fail(PC_E_INVALID, "invalid argument");
  hipSetDevice(p->device);
  CHK(prepare_batched(p));
  if (p->ntasks == 0) return fail(PC_E_INVALID, "march probe: no live patches");
  const int nscn = (int)p->scn.size();
  const int par = nscn > 0 ? (p->scn[0].stepc & 1) : 0;
  const size_t n = (size_t)p->cap_cells;
  double* o[3] = {};
  int* ctr = nullptr;
  for (int k = 0; k < 3; ++k) {
    if (hipMalloc(&o[k], n * sizeof(double)) != hipSuccess) {
      for (int j = 0; j < k; ++j) hipFree(o[j]);
      return fail(PC_E_OOM, "march probe: hipMalloc failed");
    }
  }
  if (hipMalloc(&ctr, sizeof(int)) != hipSuccess) {
    for (int k = 0; k < 3; ++k) hipFree(o[k]);
    return fail(PC_E_OOM, "march probe: hipMalloc failed");
  }
  const Fields F = fields_of(p);
  // the update's launch shape (pc_step_all): a persistent grid only when the
  // list has more tasks than it has workgroups
  const int pers = (p->persistent_wgs > 0 && p->ntasks > p->persistent_wgs) ? 1 : 0;
  const int grid = pers ? p->persistent_wgs : p->ntasks;
  long long cells = 0;
  for (int s : p->h_live)
    cells += (long long)p->slots[s].desc.subdivision_factor[0] * p->slots[s].desc.subdivision_factor[1];
  float best = 1e30f;
  hipError_t e = hipSuccess;
  for (int r = 0; r <= reps && e == hipSuccess; ++r) {  // r = 0 is an untimed warm-up
    if ((e = hipMemsetAsync(ctr, 0, sizeof(int), p->stream)) != hipSuccess) break;
    if ((e = hipEventRecord(p->ev0, p->stream)) != hipSuccess) break;
    hipLaunchKernelGGL(k_march_probe, dim3(grid), dim3(64), 0, p->stream, p->d_tasks, p->ntasks,
                       p->d_plist, p->d_pd, F.h[par], F.hu[par], F.hv[par], F.b, o[0], o[1], o[2],
                       ctr, pers);
    if ((e = hipGetLastError()) != hipSuccess) break;
    if ((e = hipEventRecord(p->ev1, p->stream)) != hipSuccess) break;
    if ((e = hipEventSynchronize(p->ev1)) != hipSuccess) break;
    float ms = 0.f;
    if ((e = hipEventElapsedTime(&ms, p->ev0, p->ev1)) != hipSuccess) break;
    if (r > 0) best = std::min(best, ms);
  }
  hipStreamSynchronize(p->stream);
  for (int k = 0; k < 3; ++k) hipFree(o[k]);
  hipFree(ctr);
  if (e != hipSuccess) return fail(PC_E_HIP, std::string("march probe: ") + hipGetErrorString(e));
  *best_gbs = 56.0 * (double)cells / (best * 1e-3) / 1e9;
  return PC_OK;
}

int pc_pool_device_fields(pc_pool* p, int state, double** h, double** hu, double** hv,
                          double** b) {
  if (!p || (state != 0 && state != 1)) return fail(PC_E_INVALID, "invalid argument");
  if (h) *h = p->f[0 + state];
  if (hu) *hu = p->f[2 + state];
  if (hv) *hv = p->f[4 + state];
  if (b) *b = p->f[6];
  return PC_OK;
}

}  // extern "C"

#include "pc_server.hpp"
#include "pc_amr.hpp"
#include "pc_vtu.hpp"
