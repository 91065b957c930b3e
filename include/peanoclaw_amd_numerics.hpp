/*
 * peanoclaw_amd — header-only C++ adapter from PeanoClaw's Numerics plugin
 * contract to the C ABI of peanoclaw_amd.h.
 *
 * peanoclaw::Numerics (src/peanoclaw/Numerics.h:30-204) is called by Peano's
 * traversal with `Patch&` facades over heap-owned arrays.  SWENumerics<PatchT>
 * implements every method of that contract for the SWE kernel
 * (src/peanoclaw/native/SWEKernel.cpp) on the device pool:
 *
 *  - a patch's data moves into the pool at initializePatch (registration +
 *    upload of the heap vector); the heap vector is refreshed from the device
 *    by addPatchToSolution (the plotters read it) and can be pushed again
 *    with uploadFromHeap after host-side edits;
 *  - device slots are keyed by the cell's CellDescription index, which stays
 *    fixed for the cell's lifetime.  The heap vector behind it (uIndex) does
 *    not: Patch::switchToNonVirtual deletes it and Patch::switchToVirtual
 *    creates a fresh one (Patch.cpp:732-790), once per grid iteration for a
 *    refined cell (GridLevelTransfer.cpp:108-135, 338-485).  When a call
 *    sees a new uIndex the slot is refilled from the new heap vector; a
 *    virtual patch first seen in a call (a cell created refined, never
 *    initialised as a leaf here) is registered then;
 *  - the CellDescription fields the kernels read (geometry, level, virtual
 *    flag, TimeIntervals, demanded mesh width, restriction bounds) are pushed
 *    to the pool before every call that reads them, so the host's
 *    TimeIntervals stay the one source of truth;
 *  - after solveTimestep the host's TimeIntervals receive the clock the pool
 *    advanced, as SWEKernel::advanceBlockInTime does (SWEKernel.cpp:83-87).
 *
 * PatchT is reached only through an access policy (default:
 * PeanoPatchAccess, the member names of peanoclaw::Patch, Patch.h:249-514,
 * its SubgridAccessor, grid/SubgridAccessor.h:230-236, and TimeIntervals,
 * grid/TimeIntervals.h:77-130).  The heap scalar type may be double or
 * float (the reference's SWE builds store float); values are staged through
 * contiguous double buffers in the reference layout
 * (grid/LinearizationZYXQ.h:24-82).
 *
 * Errors: a failing C call throws peanoclaw_amd::Error with the status code
 * and pc_last_error_string(); a Numerics subclass forwards to this class and
 * turns the exception into the reference's logError + assertion.
 *
 * A Peano build would add (the only file that includes both Peano headers and
 * this one):
 *
 *   class HipSWENumerics : public peanoclaw::Numerics {
 *     peanoclaw_amd::SWENumerics<peanoclaw::Patch> _impl;
 *    public:
 *     void solveTimestep(peanoclaw::Patch& p, double maxDt, bool split,
 *                        tarch::la::Vector<4, bool> flags) override {
 *       _impl.solveTimestep(p, maxDt, split, flags);
 *     }
 *     ...  // one line per Numerics method
 *   };
 */
#ifndef PEANOCLAW_AMD_NUMERICS_HPP
#define PEANOCLAW_AMD_NUMERICS_HPP

#include <cstddef>
#include <functional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "peanoclaw_amd.h"

namespace peanoclaw_amd {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc, const char* call) {
  if (rc != PC_OK) throw Error(rc, std::string(call) + ": " + pc_last_error_string());
}

/* Access to a peanoclaw::Patch by the reference's member names. */
template <class PatchT>
struct PeanoPatchAccess {
  /* the cell's identity for its lifetime (CellDescription index) */
  static int key(PatchT& p) { return p.getCellDescriptionIndex(); }
  /* the heap vector currently behind it; -1 when it holds no data */
  static int dataIndex(PatchT& p) { return p.getUIndex(); }
  static bool isVirtual(PatchT& p) { return p.isVirtual(); }
  static auto& timeIntervals(PatchT& p) { return p.getTimeIntervals(); }
  static auto* uNew(PatchT& p) { return p.getAccessor().getUNewArray(); }
  static auto* uOld(PatchT& p, int unknown) {
    return p.getAccessor().getUOldWithGhostLayerArray(unknown);
  }
  /* SWE's b: the one parameter field with ghost layer (SWEKernel.h:124-133) */
  static auto* bathymetry(PatchT& p) { return p.getAccessor().getParameterWithGhostLayerArray(0); }

  static void describe(PatchT& p, pc_cell_description& d) {
    const auto pos = p.getPosition();
    const auto size = p.getSize();
    const auto sub = p.getSubdivisionFactor();
    const auto demanded = p.getDemandedMeshWidth();
    for (int k = 0; k < 2; ++k) {
      d.position[k] = pos[k];
      d.size[k] = size[k];
      d.subdivision_factor[k] = sub[k];
      d.demanded_mesh_width[k] = demanded[k];
    }
    d.ghostlayer_width = p.getGhostlayerWidth();
    d.level = p.getLevel();
    d.is_virtual = p.isVirtual() ? 1 : 0;
    auto& t = p.getTimeIntervals();
    d.time = t.getCurrentTime();
    d.timestep_size = t.getTimestepSize();
    d.estimated_next_timestep_size = t.getEstimatedNextTimestepSize();
    d.minimal_neighbor_time_constraint = t.getMinimalNeighborTimeConstraint();
    if (p.isVirtual()) {
      // Patch.cpp:916-941: the region overlapped by coarse neighbours' ghosts
      const auto lo = p.getLowerNeighboringGhostlayerBounds();
      const auto hi = p.getUpperNeighboringGhostlayerBounds();
      for (int k = 0; k < 2; ++k) {
        d.restriction_lower_bounds[k] = lo[k];
        d.restriction_upper_bounds[k] = hi[k];
      }
    }
  }
};

struct SWENumericsOptions {
  int device = 0;
  size_t hbm_budget_bytes = 0;  // 0: 90% of the device memory
  int solver = 1;               // WAVE_PROPAGATION_SOLVER: 1/4 f-wave, 2 augmented Riemann
  // BreakingDamSWEScenario::computeDemandedMeshWidth (BreakingDam.cpp:85-139)
  double refinement_threshold = 0.1;
  double minimal_mesh_width = 1.0 / 81.0;
  double maximal_mesh_width = 1.0 / 9.0;
};

template <class PatchT, class Access = PeanoPatchAccess<PatchT>>
class SWENumerics {
 public:
  /* `scenario` fills a new patch's heap vector (uNew, b) before it moves to
   * the device: SWEKernel::initializePatch -> SWEScenario::initializePatch
   * (SWEKernel.cpp:111-119). */
  explicit SWENumerics(const SWENumericsOptions& o = SWENumericsOptions(),
                       std::function<void(PatchT&)> scenario = {})
      : opt_(o), scenario_(std::move(scenario)) {
    check(pc_pool_create(o.device, o.hbm_budget_bytes, &pool_), "pc_pool_create");
    const int rc = pc_set_wave_propagation_solver(pool_, o.solver);
    if (rc != PC_OK) {
      pc_pool_destroy(pool_);
      check(rc, "pc_set_wave_propagation_solver");
    }
  }
  ~SWENumerics() {
    if (pool_) pc_pool_destroy(pool_);
  }
  SWENumerics(const SWENumerics&) = delete;
  SWENumerics& operator=(const SWENumerics&) = delete;

  pc_pool* pool() const { return pool_; }
  bool registered(PatchT& p) const { return slots_.count(Access::key(p)) != 0; }
  /* The patch's slot, with the device following the host's heap vector:
   * a new uIndex (a fresh vector from Patch::switchToVirtual) is uploaded
   * before the slot is used; a virtual patch seen for the first time is
   * registered and uploaded. */
  int slot(PatchT& p) {
    const int data = Access::dataIndex(p);
    auto it = slots_.find(Access::key(p));
    if (it == slots_.end()) {
      if (!Access::isVirtual(p) || data == -1)
        throw Error(PC_E_INVALID, "patch was not initialised on the device");
      const int s = registerPatch(p);
      uploadFromHeap(p);
      return s;
    }
    if (it->second.data != data) {
      if (data == -1)
        throw Error(PC_E_INVALID, "patch holds no data (uIndex -1: switched to non-virtual)");
      it->second.data = data;
      uploadFromHeap(p);
    }
    return it->second.slot;
  }

  /* ---- Numerics (Numerics.h:52-183) ---- */

  void initializePatch(PatchT& p) {
    if (scenario_) scenario_(p);
    registerPatch(p);
    uploadFromHeap(p);
  }

  void addPatchToSolution(PatchT& p) { downloadToHeap(p); }

  /* SWEKernel::update -> SWEScenario::update; the dam break changes nothing
   * (BreakingDam.h:44). */
  void update(PatchT&) {}

  template <class Flags>
  void solveTimestep(PatchT& p, double maximumTimestepSize, bool /*useDimensionalSplitting*/,
                     const Flags& domainBoundaryFlags) {
    const int s = push(p);
    int f[4];
    for (int k = 0; k < 4; ++k) f[k] = domainBoundaryFlags[k] ? 1 : 0;
    double dt = 0.0, est = 0.0;
    check(pc_solve_timestep(pool_, s, maximumTimestepSize, f, &dt, &est), "pc_solve_timestep");
    // SWEKernel.cpp:83-87, as the pool applied it to its copy of the clock
    pc_cell_description d;
    check(pc_patch_get_description(pool_, s, &d), "pc_patch_get_description");
    auto& t = Access::timeIntervals(p);
    t.setCurrentTime(d.time);
    t.setTimestepSize(d.timestep_size);
    t.setEstimatedNextTimestepSize(d.estimated_next_timestep_size);
    cache_[s] = d;
  }

  void fillBoundaryLayer(PatchT& p, int dimension, bool setUpper) {
    check(pc_fill_boundary_layer(pool_, slot(p), dimension, setUpper ? 1 : 0),
          "pc_fill_boundary_layer");
  }

  template <class IVec>
  void transferGhostlayer(const IVec& size, const IVec& sourceOffset,
                          const IVec& destinationOffset, PatchT& source, PatchT& destination) {
    const int src = push(source), dst = push(destination);
    int sz[2] = {size[0], size[1]}, so[2] = {sourceOffset[0], sourceOffset[1]},
        dof[2] = {destinationOffset[0], destinationOffset[1]};
    check(pc_transfer_ghostlayer(pool_, sz, so, dof, src, dst), "pc_transfer_ghostlayer");
  }

  template <class IVec>
  void interpolateSolution(const IVec& destinationSize, const IVec& destinationOffset,
                           PatchT& source, PatchT& destination, bool interpolateToUOld,
                           bool interpolateToCurrentTime, bool useTimeUNewOrTimeUOld) {
    const int src = push(source), dst = push(destination);
    int sz[2] = {destinationSize[0], destinationSize[1]},
        dof[2] = {destinationOffset[0], destinationOffset[1]};
    check(pc_interpolate_solution(pool_, sz, dof, src, dst, interpolateToUOld ? 1 : 0,
                                  interpolateToCurrentTime ? 1 : 0,
                                  useTimeUNewOrTimeUOld ? 1 : 0),
          "pc_interpolate_solution");
  }

  void restrictSolution(PatchT& source, PatchT& destination, bool restrictOnlyOverlappedRegions) {
    const int src = push(source), dst = push(destination);
    check(pc_restrict_solution(pool_, src, dst, restrictOnlyOverlappedRegions ? 1 : 0),
          "pc_restrict_solution");
  }

  void postProcessRestriction(PatchT& destination, bool restrictOnlyOverlappedRegions) {
    check(pc_post_process_restriction(pool_, push(destination),
                                      restrictOnlyOverlappedRegions ? 1 : 0),
          "pc_post_process_restriction");
  }

  /* Numerics::computeFluxes / applyFluxCorrection (Numerics.cpp:80-98), called
   * by SolveTimestep when flux correction is enabled (SolveTimestep.cpp:673-679). */
  void computeFluxes(PatchT& subgrid) {
    check(pc_compute_fluxes(pool_, push(subgrid)), "pc_compute_fluxes");
  }
  void applyFluxCorrection(PatchT& sourceSubgrid, PatchT& destinationSubgrid, int dimension,
                           int direction) {
    const int src = push(sourceSubgrid), dst = push(destinationSubgrid);
    check(pc_apply_flux_correction(pool_, src, dst, dimension, direction),
          "pc_apply_flux_correction");
  }

  /* BreakingDamSWEScenario::computeDemandedMeshWidth on the current uNew */
  std::pair<double, double> getDemandedMeshWidth(PatchT& p, bool /*isInitializing*/) {
    double w[2];
    check(pc_demanded_mesh_width(pool_, slot(p), opt_.refinement_threshold,
                                 opt_.minimal_mesh_width, opt_.maximal_mesh_width, w),
          "pc_demanded_mesh_width");
    return {w[0], w[1]};
  }

  /* The scenario's refinement parameters (BreakingDamSWEScenario's
   * threshold and minimal / maximal mesh widths) for getDemandedMeshWidth. */
  void setRefinementCriterion(double threshold, double minimalMeshWidth,
                              double maximalMeshWidth) {
    opt_.refinement_threshold = threshold;
    opt_.minimal_mesh_width = minimalMeshWidth;
    opt_.maximal_mesh_width = maximalMeshWidth;
  }

  int getNumberOfUnknownsPerCell() const { return pc_get_number_of_unknowns_per_cell(); }
  int getNumberOfParameterFieldsWithoutGhostlayer() const {
    return pc_get_number_of_parameter_fields_without_ghostlayer();
  }
  int getNumberOfParameterFieldsWithGhostlayer() const {
    return pc_get_number_of_parameter_fields_with_ghostlayer();
  }
  int getGhostlayerWidth() const { return pc_get_ghostlayer_width(); }

  /* ---- DefaultTransfer (DefaultTransfer.cpph:103-127) ---- */
  void copyUNewToUOld(PatchT& p) { check(pc_copy_unew_to_uold(pool_, slot(p)), "pc_copy_unew_to_uold"); }
  void swapUNewAndUOld(PatchT& p) {
    check(pc_swap_unew_and_uold(pool_, slot(p)), "pc_swap_unew_and_uold");
  }

  /* The step solveTimestep would take with an unbounded maximum (the CFL
   * part of SWEKernel.cpp:60-68), for a caller reducing a global dt. */
  double cfl(PatchT& p) {
    double dt = 0.0;
    check(pc_patch_cfl(pool_, push(p), &dt), "pc_patch_cfl");
    return dt;
  }

  /* ---- heap <-> pool mirroring ---- */

  /* A new patch gets a slot; data start zeroed as a fresh heap vector. */
  int registerPatch(PatchT& p) {
    if (registered(p)) throw Error(PC_E_INVALID, "patch is already registered");
    pc_cell_description d{};
    Access::describe(p, d);
    d.scenario = 0;
    int s = -1;
    check(pc_patch_register(pool_, &d, &s), "pc_patch_register");
    slots_[Access::key(p)] = Entry{s, Access::dataIndex(p)};
    cache_[s] = d;
    return s;
  }

  /* Remesh::destroyCell: the slot's tile returns to the pool. */
  void destroyPatch(PatchT& p) {
    auto it = slots_.find(Access::key(p));
    if (it == slots_.end()) throw Error(PC_E_INVALID, "patch was not initialised on the device");
    const int s = it->second.slot;
    check(pc_patch_release(pool_, s), "pc_patch_release");
    slots_.erase(it);
    cache_.erase(s);
  }

  /* Heap vector (uNew, uOld with ghost layer, b with ghost layer) -> pool. */
  void uploadFromHeap(PatchT& p) {
    const int s = push(p);
    const Dims g = dims(s);
    stage_.assign(g.unew + 3 * g.tile + g.tile, 0.0);
    double* un = stage_.data();
    double* uo = un + g.unew;
    double* b = uo + 3 * g.tile;
    gather(Access::uNew(p), un, g.unew);
    for (int q = 0; q < 3; ++q) gather(Access::uOld(p, q), uo + q * g.tile, g.tile);
    gather(Access::bathymetry(p), b, g.tile);
    check(pc_patch_upload(pool_, s, un, uo, b), "pc_patch_upload");
  }

  /* Pool -> heap vector (the plotters and checkpoints read it). */
  void downloadToHeap(PatchT& p) {
    const int s = slot(p);
    const Dims g = dims(s);
    stage_.assign(g.unew + 3 * g.tile + g.tile, 0.0);
    double* un = stage_.data();
    double* uo = un + g.unew;
    double* b = uo + 3 * g.tile;
    check(pc_patch_download(pool_, s, un, uo, b), "pc_patch_download");
    scatter(un, Access::uNew(p), g.unew);
    for (int q = 0; q < 3; ++q) scatter(uo + q * g.tile, Access::uOld(p, q), g.tile);
    scatter(b, Access::bathymetry(p), g.tile);
  }

  /* Pushes the host's CellDescription fields (geometry after a virtual/leaf
   * switch, level, TimeIntervals) to the pool; returns the slot. */
  int push(PatchT& p) {
    const int s = slot(p);
    pc_cell_description d = cache_[s];
    Access::describe(p, d);
    if (!same(d, cache_[s])) {
      check(pc_patch_set_description(pool_, s, &d), "pc_patch_set_description");
      cache_[s] = d;
    }
    return s;
  }

 private:
  struct Dims {
    size_t unew, tile;
  };
  Dims dims(int s) const {
    const pc_cell_description& d = cache_.at(s);
    const size_t nx = d.subdivision_factor[0], ny = d.subdivision_factor[1];
    return {3 * nx * ny, (nx + 2) * (ny + 2)};
  }
  template <class T>
  static void gather(const T* src, double* dst, size_t n) {
    for (size_t i = 0; i < n; ++i) dst[i] = static_cast<double>(src[i]);
  }
  template <class T>
  static void scatter(const double* src, T* dst, size_t n) {
    for (size_t i = 0; i < n; ++i) dst[i] = static_cast<T>(src[i]);
  }
  static bool same(const pc_cell_description& a, const pc_cell_description& b) {
    for (int k = 0; k < 2; ++k)
      if (a.position[k] != b.position[k] || a.size[k] != b.size[k] ||
          a.demanded_mesh_width[k] != b.demanded_mesh_width[k] ||
          a.restriction_lower_bounds[k] != b.restriction_lower_bounds[k] ||
          a.restriction_upper_bounds[k] != b.restriction_upper_bounds[k])
        return false;
    return a.level == b.level && a.is_virtual == b.is_virtual && a.time == b.time &&
           a.timestep_size == b.timestep_size &&
           a.estimated_next_timestep_size == b.estimated_next_timestep_size &&
           a.minimal_neighbor_time_constraint == b.minimal_neighbor_time_constraint;
  }

  SWENumericsOptions opt_;
  std::function<void(PatchT&)> scenario_;
  pc_pool* pool_ = nullptr;
  struct Entry {
    int slot;  // pool slot
    int data;  // uIndex of the heap vector the slot mirrors
  };
  std::unordered_map<int, Entry> slots_;                // CellDescription index -> slot
  std::unordered_map<int, pc_cell_description> cache_;  // last description the pool holds
  std::vector<double> stage_;
};

}  // namespace peanoclaw_amd

#endif
