/*
 * peanoclaw_amd — C ABI of the MI355X (gfx950) shallow-water patch path.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.
 * Every call runs on the HIP stream owned by the pool; calls on one pool are
 * thread-compatible (not thread-safe), different pools are independent.
 *
 * The per-patch entry points mirror peanoclaw::Numerics
 * (src/peanoclaw/Numerics.h:30-204) as implemented by the SWE plugin
 * (src/peanoclaw/native/SWEKernel.{h,cpp}) with the default transfer /
 * interpolation / restriction strategies wired by
 * NumericsFactory::createSWENumerics (src/peanoclaw/NumericsFactory.cpp:27-51).
 * A `Patch&` of the reference becomes a slot id returned by
 * pc_patch_register.  The batched entry points (pc_*_all, pc_graph_*) have
 * no single counterpart: they run the whole per-step sequence of
 * mappings/SolveTimestep.cpp:591-751 for every registered patch at once.
 *
 * Data layout of the host arrays exchanged with a slot is the reference's
 * own (grid/LinearizationZYXQ.h:24-82): unknown q slowest, then x, then y
 * fastest; uNew holds 3*nx*ny values, uOld and b hold (nx+2)*(ny+2) per
 * field including the ghost layer of width 1.
 */
#ifndef PEANOCLAW_AMD_H
#define PEANOCLAW_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define PC_OK 0
#define PC_E_INVALID 1   /* bad handle, slot, size or argument            */
#define PC_E_OOM 2       /* request exceeds the pool's HBM budget          */
#define PC_E_HIP 3       /* a HIP runtime call failed                      */
#define PC_E_NONFINITE 4 /* a non-finite value was found in a patch        */
#define PC_E_NODEVICE 5  /* no usable gfx950 device                        */

typedef struct pc_pool pc_pool;

/* Mirrors the fields of dastgen/CellDescription.def:10-201 that the SWE
 * path reads or writes. */
typedef struct {
  double position[2];
  double size[2];
  int32_t subdivision_factor[2];     /* nx, ny (each 2..64)                */
  int32_t ghostlayer_width;          /* must be 1 (SWEKernel.h:133)        */
  int32_t level;
  int32_t is_virtual;
  int32_t scenario;                  /* independent problem id (>= 0)      */
  double time;
  double timestep_size;
  double estimated_next_timestep_size;
  double minimal_neighbor_time_constraint;
  double demanded_mesh_width[2];
  /* neighbouring ghost-layer bounds of a virtual patch
   * (CellDescription restrictionLowerBounds / restrictionUpperBounds,
   * Patch.cpp:916-941), used by restrictSolution(onlyOverlapped). */
  double restriction_lower_bounds[2];
  double restriction_upper_bounds[2];
} pc_cell_description;

/* Error text of the last failing call on this thread. */
const char *pc_last_error_string(void);

/* ---- pool lifecycle (replaces Peano's DataHeap/CellDescriptionHeap) ---- */
int pc_pool_create(int device, size_t hbm_budget_bytes, pc_pool **out);
int pc_pool_destroy(pc_pool *pool);
int pc_pool_bytes_in_use(pc_pool *pool, size_t *out);
/* Pre-allocate field storage for `cells` tile cells (sum of
 * (nx+2)*pc_tile_stride(ny), plus up to 15 alignment cells per tile), 56
 * bytes each; fails with PC_E_OOM beyond the budget. */
int pc_pool_reserve(pc_pool *pool, long long cells);
int pc_pool_synchronize(pc_pool *pool);

/* Register a patch; returns its slot (replaces CellDescription.uIndex,
 * Patch.cpp:742-769).  Data starts zeroed like the reference's heap vector. */
int pc_patch_register(pc_pool *pool, const pc_cell_description *desc,
                      int *out_slot);
/* Release a slot (DataHeap::deleteData of the patch's uIndex).  Grid sides
 * of other patches that read their ghosts from it become reflective walls
 * (the grid then has a hole there); its slot id may be reused. */
int pc_patch_release(pc_pool *pool, int slot);
/* The slot's description.  While the batched step drives the slot's clock,
 * time and timestep size are its scenario's, and estimated_next_timestep_size
 * is the patch's own CFL estimate (SWEKernel.cpp:88) - except in the pruned
 * launch mode (pc_set_launch_config fused = 2) on a grid with side sources,
 * where only the scenario minimum is computed exactly and is reported for
 * every patch (a lower bound of the patch's own estimate); fused = 0 or 1
 * gives exact per-patch estimates at the same dt. */
int pc_patch_get_description(pc_pool *pool, int slot, pc_cell_description *out);
int pc_patch_set_description(pc_pool *pool, int slot,
                             const pc_cell_description *desc);
/* Host <-> device copies of a slot's arrays (any pointer may be NULL). */
int pc_patch_upload(pc_pool *pool, int slot, const double *unew,
                    const double *uold_with_ghost, const double *b_with_ghost);
int pc_patch_download(pc_pool *pool, int slot, double *unew,
                      double *uold_with_ghost, double *b_with_ghost);

/* Bulk access for large pools: the tile offset of a slot in the field
 * arrays and its row stride (row x of the tile, ghost rows included, starts
 * at offset + x*stride; stride = pc_tile_stride(ny): ny + 2, padded to a
 * multiple of 16 for ny >= 8 so that every row's interior starts on a
 * 128-byte line), the state (0/1) holding a scenario's current values, and a
 * copy of cells [begin, begin+count) of state `state` (and of b when
 * b != NULL) between host arrays and the pool (to_device = 1 uploads).  No
 * counterpart in the reference (its patches live in Peano's heap); these
 * serve hosts that fill or read a whole pool at once. */
int pc_patch_offset(pc_pool *pool, int slot, long long *offset);
int pc_patch_stride(pc_pool *pool, int slot, int *stride);
int pc_tile_stride(int ny); /* -1 outside [2, 64] */
int pc_scenario_current_state(pc_pool *pool, int scenario, int *state);
int pc_pool_copy_fields(pc_pool *pool, int to_device, int state,
                        long long begin, long long count, double *h,
                        double *hu, double *hv, double *b);

/* ---- Numerics-mirroring per-patch calls ---- */
/* Numerics::solveTimestep (Numerics.h:147-152; SWEKernel.cpp:125-167).
 * boundary_flags[4] is accepted for interface parity and unused, as in the
 * reference.  Outputs the applied dt and estimatedNextTimestepSize.  Returns
 * as soon as both are known; the update itself finishes on the pool's
 * stream, before any later call's work on the pool (an error of it is
 * reported by a later call that synchronises). */
int pc_solve_timestep(pc_pool *pool, int slot, double maximum_timestep_size,
                      const int boundary_flags[4], double *out_dt,
                      double *out_estimated_next_dt);
/* The CFL step solveTimestep would take on this patch with an unbounded
 * maximumTimestepSize (SWEKernel.cpp:60-68, computeNumericalFluxes' rule
 * SWE_WaveAccumulationBlock.cpp:225-239), without updating anything: the
 * minimum over patches gives a global time step. */
int pc_patch_cfl(pc_pool *pool, int slot, double *out_dt);
/* Numerics::fillBoundaryLayer (Numerics.h:131-135; SWEKernel.cpp:176-237) */
int pc_fill_boundary_layer(pc_pool *pool, int slot, int dimension,
                           int set_upper);
/* Numerics::transferGhostlayer (Numerics.h:64-70; DefaultTransfer.cpph:15-101) */
int pc_transfer_ghostlayer(pc_pool *pool, const int size[2],
                           const int source_offset[2],
                           const int destination_offset[2], int source_slot,
                           int destination_slot);
/* Numerics::interpolateSolution (Numerics.h:77-85; SWEKernel.cpp:239-286 over
 * DefaultInterpolation.cpph:29-270) */
int pc_interpolate_solution(pc_pool *pool, const int destination_size[2],
                            const int destination_offset[2], int source_slot,
                            int destination_slot, int interpolate_to_uold,
                            int interpolate_to_current_time,
                            int use_time_unew_or_uold);
/* Numerics::restrictSolution (Numerics.h:92-96; SWEKernel.cpp:288-307 over
 * DefaultRestriction.cpph:108-150).  The destination must be virtual. */
int pc_restrict_solution(pc_pool *pool, int source_slot, int destination_slot,
                         int restrict_only_overlapped_regions);
/* Numerics::postProcessRestriction (Numerics.h:103-106; SWEKernel.cpp:309-316) */
int pc_post_process_restriction(pc_pool *pool, int destination_slot,
                                int restrict_only_overlapped_regions);

/* Batched level transfers for regridding (Remesh::createCell /
 * destroyCell's interpolateSolution and restrictSolution sequences,
 * Remesh.cpp:331-540).  A batch has exactly the effect of its calls made one
 * after another in the order given: calls that share no patch run together,
 * one launch per step for all of them.  Every call is checked before any
 * runs; an invalid one fails the batch with nothing changed.
 *
 * pc_interpolate_batch: `count` pc_interpolate_solution calls; call i has
 * destination_size {dsize[2i], dsize[2i+1]}, destination_offset
 * {doff[2i], doff[2i+1]}, source src[i], destination dst[i], and
 * opts[3i..3i+2] = interpolate_to_uold, interpolate_to_current_time,
 * use_time_unew_or_uold. */
int pc_interpolate_batch(pc_pool *pool, int count, const int *dsize, const int *doff,
                         const int *src, const int *dst, const int *opts);
/* `count` pc_restrict_solution(src[i], dst[i], restrict_only_overlapped_regions)
 * calls (the restrictions of several fine patches into one virtual patch
 * keep their order). */
int pc_restrict_batch(pc_pool *pool, int count, const int *src, const int *dst,
                      int restrict_only_overlapped_regions);
/* Patch::switchToLeaf (Patch.cpp:793-830) of `count` distinct patches: the
 * patch becomes a leaf with interval [time[i], time[i] + dt[i]], and its
 * restricted values, which hold the [0, 1] time basis, are evaluated at the
 * interval's ends (uOld at time[i], uNew at time[i] + dt[i]). */
int pc_switch_to_leaf(pc_pool *pool, int count, const int *slots, const double *time,
                      const double *dt);
/* Zero uNew and uOld (ghost ring included) of `count` distinct patches, as
 * a virtual patch is cleared before the restrictions into it; b is kept. */
int pc_patches_clear(pc_pool *pool, int count, const int *slots);
/* Set b (ghost ring included, the layout of pc_patch_upload's b) of `count`
 * distinct patches from one array: patch i takes the next
 * (nx_i + 2)(ny_i + 2) values. */
int pc_patches_set_b(pc_pool *pool, int count, const int *slots, const double *b);
/* Flux correction, off by default in the reference (SolveTimestep's
 * _correctFluxes; called after solveTimestep, SolveTimestep.cpp:673-679).
 * Each slot holds the reference's flux array: faces x-lower, x-upper,
 * y-lower, y-upper, each 3 unknowns x the cells along the face (unknown
 * slowest), 6 * (nx + ny) doubles.  The present reference code is
 * reproduced as it is: the upwind term of computeFluxes is dropped by its
 * ';' (DefaultFluxCorrection.cpph:71-72), Numerics::computeFlux uses g = 1
 * and scales the momenta by h (Numerics.h:185-203), a source cell reaches
 * only destination cells overlapping the destination's first tangential
 * subcell (HyperplaneRegion.cpp:76), and the destination flux is read from
 * the start of its face (SubgridFaceIterator::restart).  Square patches
 * only (the face iterator places every upper face at subdivisionFactor[1]).
 *
 * Numerics::computeFluxes (Numerics.cpp:93-98; DefaultFluxCorrection.cpph:44-86):
 * the slot's fluxes from uOld (interior cells next to each face and their
 * ghost neighbours). */
int pc_compute_fluxes(pc_pool *pool, int slot);
/* Numerics::applyFluxCorrection (Numerics.cpp:80-91; DefaultFluxCorrection.cpph:
 * 88-164): corrects the destination's uNew next to its face (dimension,
 * -direction) by (F_source - F_destination) * time overlap * overlap area /
 * subcell area.  PC_E_INVALID when the two patches share their position. */
int pc_apply_flux_correction(pc_pool *pool, int source_slot, int destination_slot,
                             int dimension, int direction);
/* Host <-> device copy of a slot's flux array (to_device = 1 uploads). */
int pc_patch_fluxes(pc_pool *pool, int slot, int to_device, double *fluxes);

/* DefaultTransfer::copyUNewToUOld / swapUNewAndUOld (DefaultTransfer.cpph:103-127) */
int pc_copy_unew_to_uold(pc_pool *pool, int slot);
int pc_swap_unew_and_uold(pc_pool *pool, int slot);
/* Numerics::getDemandedMeshWidth -> BreakingDamSWEScenario::
 * computeDemandedMeshWidth (BreakingDam.cpp:85-139) with the gradient
 * threshold and minimal/maximal widths of the scenario. */
int pc_demanded_mesh_width(pc_pool *pool, int slot, double threshold,
                           double minimal_width, double maximal_width,
                           double out[2]);
/* Numerics constants (SWEKernel.h:124-133) */
int pc_get_number_of_unknowns_per_cell(void);
int pc_get_number_of_parameter_fields_without_ghostlayer(void);
int pc_get_number_of_parameter_fields_with_ghostlayer(void);
int pc_get_ghostlayer_width(void);

/* ---- batched step (all patches, one scenario clock per scenario) ---- */
/* One ghost strip: `length` cells copied from source cells
 * (src_slot, src_cell + k*src_stride) to (dst_slot, dst_cell + k*dst_stride),
 * cell indices into the slot's tile (cell (x, y) of the tile, ghost ring
 * included, is x*pc_patch_stride + y); negate is a bit mask:
 * bit 0 negates hu, bit 1 negates hv (reflective walls; a corner reflected
 * by two walls has both).  Strips must be independent (no strip reads a
 * ghost cell written by another).                                         */
typedef struct {
  int32_t dst_slot, dst_cell, dst_stride;
  int32_t src_slot, src_cell, src_stride;
  int32_t length, negate;
} pc_ghost_strip;
/* Replaces the device neighbour table; called only when the grid changes. */
int pc_set_ghost_strips(pc_pool *pool, const pc_ghost_strip *strips, int n);
/* Ghost strips of a px x py grid of equal-size patches (slot_of[i*py + j]
 * is patch (i, j), i along x; -1 marks a hole).  Sides facing a hole or the
 * border are reflective walls.  The strips reproduce the reference's fill
 * order: same-level transfers from the 8 neighbours (GhostLayerCompositor,
 * Vertex.cpp:61-145), then walls left, top, right, bottom including the
 * corner cells (SolveTimestep.cpp:87-142).  Writes up to `cap` strips to
 * `out` and their number to *n_out (call with out = NULL to count). */
int pc_grid_strips(pc_pool *pool, int px, int py, const int *slot_of,
                   pc_ghost_strip *out, int cap, int *n_out);
/* Fused ghost sources of the batched step for a px x py grid of equal-size
 * patches (slot_of as for pc_grid_strips): every side of every patch reads
 * its ghost values directly from the face neighbour (rows from the
 * neighbour's boundary row, columns from its edge buffer) or, facing a hole
 * or the border, reflects its own boundary (negated normal momentum).  The
 * values equal those the reference's ghost fill writes into the ghost
 * layer (DefaultTransfer at time factor 1, SWEKernel::fillBoundaryLayer);
 * corner ghosts are not needed by the update (its x-edges use rows 1..ny,
 * its y-edges columns 1..nx) and are not produced.  Patches without sides
 * read their tile's ghost layer, filled by the ghost strips. */
int pc_set_grid_sides(pc_pool *pool, int px, int py, const int *slot_of);
int pc_clear_sides(pc_pool *pool);
/* Per scenario clock and global end time; every step uses
 * maximumTimestepSize = end_time - (time + timestep_size)
 * (SolveTimestep.cpp:623-625); a scenario whose bound is <= 0 does not
 * advance. */
int pc_scenario_set(pc_pool *pool, int scenario, double time,
                    double timestep_size, double end_time);
int pc_scenario_get(pc_pool *pool, int scenario, double *time,
                    double *timestep_size, double *estimated_next_dt,
                    long long *steps);
/* Execution options: tasks per launch (0 = one task per workgroup,
 * > 0 = persistent grid of that many workgroups when the step has more
 * tasks than that, else one task per workgroup), and how the CFL edge
 * speeds of the next step are obtained (grids with side sources only):
 *   fused = 0: a separate exact speed pass before every dt reduction;
 *   fused = 1: exact speeds of every interior edge inside the update kernel;
 *   fused = 2: pruned - the update bounds every edge speed cheaply and
 *              computes the exact speed only where it may reach the previous
 *              CFL limit; an exact fallback pass runs whenever the pruned
 *              minimum is not proven exact.  dt and the scenario estimate are
 *              identical in all three modes; in mode 2 a patch's
 *              estimated_next_timestep_size reports the scenario estimate.
 * patches_per_task_cap limits patches per wavefront task (0 = no cap). */
int pc_set_launch_config(pc_pool *pool, int persistent_workgroups, int fused,
                         int patches_per_task_cap);
/* Edge solver of every update of the pool, the build-time choice of the
 * reference (SConstruct:291-303, WAVE_PROPAGATION_SOLVER): 1 = f-wave,
 * 2 = augmented Riemann (George 2008), 4 = f-wave (the reference's
 * vectorised variant; same arithmetic here).  Applies to the batched step,
 * the per-patch entry points and pc_edge_batch; with the augmented solver
 * the batched step runs a separate exact speed pass (fused settings are
 * kept but not used).  Default 1. */
int pc_set_wave_propagation_solver(pc_pool *pool, int solver);
/* Tuning of fused = 2 (no reference counterpart): exact speeds are computed
 * for edges whose bound reaches the speed of a CFL step est * slack
 * (default 1.03125); the pruned minimum is accepted when it is at most
 * est * slack * (1 - 1/256), else the exact fallback pass runs.  Results do
 * not depend on slack; slack < 1 makes every step take the fallback. */
int pc_set_prune_slack(pc_pool *pool, double slack);
/* Row segments of the batched update (no reference counterpart): 0 =
 * automatic (a step with fewer tasks than the GPU holds splits the rows of
 * each patch into 2 or 4 segments of at least 8 rows, each marched by its
 * own wavefront), 1 = whole patches, 2 or 4 = that many segments where the
 * patch has at least 8 rows per segment.  Results do not depend on it. */
int pc_set_row_segments(pc_pool *pool, int segments);
/* Run k global-dt steps of every scenario that has patches. */
int pc_step_all(pc_pool *pool, int k);
/* Plain per-patch-launch form of the same step (reference call pattern:
 * ghost fill, per-patch CFL estimate, per-patch update, one launch each). */
int pc_step_all_per_patch(pc_pool *pool, int k);
/* Capture one step as a hipGraph and replay it k times. */
int pc_graph_capture(pc_pool *pool);
int pc_graph_replay(pc_pool *pool, int k);
/* Refinement flags of all patches (1 refine, 0 coarsening allowed, 2 keep),
 * SolveTimestep::leaveCell (SolveTimestep.cpp:753-797), indexed by slot id
 * (-1 for unused ids); flags_out holds cap >= pc_pool_slot_capacity
 * entries, else PC_E_INVALID. */
int pc_refinement_flags(pc_pool *pool, double threshold, double minimal_width,
                        double maximal_width, int32_t *flags_out, int cap);
/* Number of slot ids in use so far (live or released): the size of
 * per-slot outputs. */
int pc_pool_slot_capacity(pc_pool *pool, int *n);
/* Total water volume sum(h)*dx*dy of one scenario (compensated sum). */
int pc_scenario_mass(pc_pool *pool, int scenario, double *out);
/* Device time (ms) of the last pc_step_all / pc_graph_replay call measured
 * with HIP events on the pool stream, and (pc_step_all only) the mean
 * duration of one update-kernel launch over that call's steps, from an
 * event pair recorded around every update launch. */
int pc_last_timing(pc_pool *pool, double *total_ms, double *update_ms);

/* Diagnostic: run the pool's device edge solver on n host edges
 * in[6n] = (hL, hR, huL, huR, bL, bR) -> out[5n] = (dhL, dhR, dqL, dqR,
 * speed); the per-edge computeNetUpdates call of
 * tools/SWE_WaveAccumulationBlock.cpp:154-159. */
int pc_edge_batch(pc_pool *pool, long long n, const double *in, double *out);

/* pc_edge_batch without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The edges are staged through a static device array of at most 65,536
 * edges (5.8 MB per GPU, held until the process ends); larger n fails with
 * PC_E_INVALID.  Every call on one GPU shares that array: calls that overlap
 * on the same GPU (other threads, other pools) race. */
int pc_edge_batch_static(pc_pool *pool, long long n, const double *in, double *out);

/* Diagnostic (no reference counterpart): device-to-device copy of `bytes`
 * (rounded down to 16) between two scratch buffers, `reps` timed launches of
 * one plain copy kernel (k_copy_probe, one 16-byte element per thread) after
 * a warm-up; *best_gbs = (read +
 * written bytes) / fastest launch.  Gives the achievable HBM copy rate and
 * the byte scale of hardware traffic counters. */
int pc_copy_probe(pc_pool *pool, long long bytes, int reps, double *best_gbs);

/* Diagnostic (no reference counterpart): the streaming ceiling of the
 * update's traffic mix.  Seven scratch arrays of `cells` doubles; one flat
 * kernel (k_stream_probe, a pair of elements per thread in 16-byte lanes)
 * reads four of them and writes three per
 * element, as the update reads h, hu, hv, b and writes h, hu, hv of a cell,
 * with no neighbour access.  `reps` timed launches after a warm-up;
 * *best_gbs = 56 * cells / fastest launch. */
int pc_stream_probe(pc_pool *pool, long long cells, int reps, double *best_gbs);

/* Diagnostic (no reference counterpart): the memory-pattern ceiling of the
 * batched update in the pool's own layout.  One launch of k_march_probe
 * walks the batched step's task list as the update does (its launch shape,
 * a wavefront per task, rows marched in order, four rows prefetched), reading
 * h, hu, hv, b of every tile row of the current state and writing h, hu, hv
 * of the interior rows (and, as the update does, of the ghost columns next
 * to them) to scratch arrays, without the edge arithmetic; the pool's state
 * is not changed.  `reps` timed launches after a warm-up;
 * *best_gbs = 56 * interior cells / fastest launch. */
int pc_march_probe(pc_pool *pool, int reps, double *best_gbs);

/* ---- output (grid/plotter/VTUSubgridPlotter.cpp:9-241) ----
 * One quadrilateral per subcell in a VTK unstructured grid (XML, appended raw
 * binary, Float64 points and cell arrays), vertices shared between subcells
 * and patches; cell arrays q0..q2 (uNew), pWG0 (b), SubdivisionFactor,
 * GhostlayerWidth, timeOld, timeNew, demandedMeshWidth, age, as the
 * reference's plotter writes them (plotSubcell, :97-150). */
typedef struct {
  int32_t subdivision_factor[2];
  int32_t ghostlayer_width;
  int32_t level;
  double position[2];
  double size[2];
  double time, timestep_size;  /* timeOld = time, timeNew = time + size */
  double demanded_mesh_width;  /* norm1(demanded width) / 2 */
  double age;                  /* grid iterations since creation */
  const double *q[3];          /* uNew h, hu, hv: nx*ny, x rows, y fastest */
  const double *b;             /* b at the nx*ny interior cells, same order */
} pc_vtu_patch;
/* Write host data of `npatches` patches to `path` (no device involved). */
int pc_vtu_write(const char *path, int npatches, const pc_vtu_patch *patches);
/* Download the listed slots (uNew of the current state, b, descriptor) and
 * write them to `path`; age is written as 0. */
int pc_pool_write_vtu(pc_pool *pool, const char *path, const int32_t *slots, int nslots);
/* The same cells as legacy ASCII VTK (grid/plotter/VTKSubgridPlotter.cpp:
 * 9-241): POINTS, CELLS, CELL_TYPES and one SCALARS block per array, in the
 * VTK plotter's geometry (each patch shrunk by 1 + 0.015 * level around its
 * centre, VTKSubgridPlotter.cpp:22, 104-106). */
int pc_vtk_write(const char *path, int npatches, const pc_vtu_patch *patches);
int pc_pool_write_vtk(pc_pool *pool, const char *path, const int32_t *slots, int nslots);

/* ---- serving: a stream of independent dam-break scenarios ----
 *
 * No reference counterpart: PeanoClaw runs one simulation per process
 * (native/sweMain.cpp, peanoclaw/runners).  A server keeps every admitted
 * scenario's patches and state resident in the pool and advances all of
 * them in shared batched steps (pc_step_all), each with its own CFL dt and
 * end time.  Calls on a server are thread-compatible; destroy the server
 * before its pool. */
#define PC_MAX_HILLS 8
typedef struct {
  int32_t cells_per_side;   /* N: N x N cells on [0,1]^2                    */
  int32_t patch_cells;      /* n: patches of n x n cells, N % n == 0, n<=64 */
  double end_time;          /* simulated time to reach                      */
  double column_x, column_y, column_radius, column_height;
  double depth;             /* water depth outside the column               */
  int32_t num_hills;        /* Gaussian hills of the bathymetry (<= 8)      */
  int32_t reserved;
  double hill_x[PC_MAX_HILLS], hill_y[PC_MAX_HILLS];
  double hill_sigma[PC_MAX_HILLS], hill_height[PC_MAX_HILLS];
} pc_request;
/* Initial state: h = column_height where |x - (column_x, column_y)| <=
 * column_radius, depth elsewhere; hu = hv = 0; b = sum of
 * hill_height * exp(-|x - hill|^2 / hill_sigma^2) at every cell centre
 * (ghost ring included); reflective walls at the box. */

typedef struct {
  int32_t id;
  int32_t scenario;          /* pool scenario id it ran under              */
  long long steps;           /* time steps taken                           */
  long long cell_updates;    /* steps * N^2                                */
  double time;               /* simulated time reached                     */
  double mass;               /* final water volume sum(h) dx dy            */
  uint64_t state_hash;       /* order-independent digest of the bits of   */
                             /* h, hu, hv of every interior cell           */
  long long admitted_round, completed_round;
} pc_request_result;

typedef struct pc_server pc_server;
/* keep_completed != 0: a finished request keeps its patches (readable via
 * pc_server_request_slots / pc_patch_download) until pc_server_release. */
int pc_server_create(pc_pool *pool, int max_active, int keep_completed,
                     pc_server **out);
int pc_server_destroy(pc_server *srv);
/* Queue a request (arrival order); PC_E_OOM if it can never fit. */
int pc_server_submit(pc_server *srv, const pc_request *req, int *out_id);
/* One scheduling round: admit waiting requests that fit the pool's HBM
 * budget and the active limit (first fit in arrival order), run `steps`
 * batched steps of all active requests, retire those that reached their end
 * time.  Reports the active and queued counts after the round. */
int pc_server_round(pc_server *srv, int steps, int *active, int *queued);
/* Results of retired requests, in completion order (each reported once). */
int pc_server_poll(pc_server *srv, pc_request_result *out, int cap, int *n);
/* Slots of an active or kept request in grid order (k = i * (N/n) + j);
 * with slots == NULL only *n is set. */
int pc_server_request_slots(pc_server *srv, int id, int *slots, int cap, int *n);
int pc_server_release(pc_server *srv, int id);

/* ---- batched step of an adaptive grid ----
 * The per-call Numerics sequence of one global AMR step (restriction of
 * every leaf into each virtual patch containing it, copyUNewToUOld, the
 * ghost layer of every leaf face from the same-level patch or by
 * interpolation from the coarser leaf, domain walls, global CFL minimum,
 * solveTimestep of every leaf; GridLevelTransfer.cpp:140-198,
 * GhostlayerCompositorFunctors.cpp:59-69, SolveTimestep.cpp:87-142 and
 * 591-751) handed over once as a plan and run as seven batched launches per
 * step.  Level transfers read h + b instead of transforming the source in
 * place and back (the reference's round trip can move a depth by an ulp;
 * results agree with the per-call form to ~1e-14 relative).  No reference
 * counterpart as one entry point: it replaces the traversal's calls for a
 * grid whose leaves share one clock.  The plan stays valid while its
 * patches stay registered; the patches must belong to one scenario, which
 * pc_step_all must not advance meanwhile. */
typedef struct pc_amr pc_amr;
/* kind 0: transferGhostlayer of size at src_offset of src to dst_offset of
 * dst (src a leaf or virtual patch of the same level); kind 1:
 * interpolateSolution(toUOld, not toCurrentTime, useTimeUNewOrUOld) from
 * the coarser leaf src into the region (size, dst_offset) of dst.  The
 * destination region must be a strip of dst's ghost layer. */
typedef struct {
  int32_t kind, src, dst;
  int32_t size[2], src_offset[2], dst_offset[2];
} pc_amr_ghost;
/* fillBoundaryLayer(dimension, upper) of a leaf; entries of one leaf are
 * consecutive, in call order. */
typedef struct {
  int32_t slot, dimension, upper;
} pc_amr_wall;
/* The face, corner and wall entries of every leaf of a patch tree with
 * refinement factor 3, as pc_amr_create takes them (host only, no pool):
 * patch k has level[k] in 1..12 and index (ij[2k], ij[2k+1]) inside its
 * level's grid of px 3^(l-1) x py 3^(l-1) patches, each index below 2^25
 * (level 1: a px x py grid of n x n cells); leaves[] lists the leaves in
 * plan order (indices into the patch arrays); entries name patches by
 * handle[k].  Per leaf: faces x-, x+, y-, y+ (outside the domain: a wall),
 * corners (-,-), (-,+), (+,-), (+,+) (skipped outside the domain), each a
 * transfer from the patch of the leaf's level there, else an interpolation
 * from the nearest coarser patch covering it (the compositor's choice,
 * GhostLayerCompositor.cpp:280-320, GhostlayerCompositorFunctors.cpp:87-141);
 * then the walls left, top, right, bottom (SolveTimestep.cpp:94-131).  With
 * NULL outputs only the counts are returned. */
int pc_amr_face_plan(int px, int py, int n, int npatch, const int32_t *level,
                     const int32_t *ij, const int32_t *handle, const int32_t *leaves,
                     int nleaves, pc_amr_ghost *ghosts, int cap_ghosts, int *nghosts,
                     pc_amr_wall *walls, int cap_walls, int *nwalls);
/* restrictions: n pairs (leaf, virtual) in call order. */
int pc_amr_create(pc_pool *pool, const int32_t *leaves, int nleaves,
                  const int32_t *virtuals, int nvirtuals, const int32_t *restrictions,
                  int nrestrictions, const pc_amr_ghost *ghosts, int nghosts,
                  const pc_amr_wall *walls, int nwalls, pc_amr **out);
/* k global steps, each bounded by end_time - (time + timestep_size) and
 * skipped once that is <= 1e-8; *last_dt = the last step's dt (0 if
 * skipped).  The leaves' descriptors carry the clock afterwards. */
int pc_amr_step(pc_amr *amr, int k, double end_time, double *last_dt);
int pc_amr_clock(pc_amr *amr, double *time, double *timestep_size, long long *steps);
/* Local time stepping over the same plan (SolveTimestep::enterCell,
 * SolveTimestep.cpp:591-751, with TimeIntervals::isAllowedToAdvanceInTime,
 * TimeIntervals.cpp:24-84): every leaf keeps its own clock.  `nptr` (nleaves
 * + 1 offsets, leaves in plan order) and `nbr` list for each leaf the plan
 * patches sharing a face or a corner with it (a leaf, or the virtual patch
 * covering that region); a leaf advances when end_time - (time + dt) > 1e-8
 * and time + dt - (minimal end time of its neighbours; of a virtual patch:
 * of the leaves below it) <= 1e-8.  Initialises the leaves' clocks from
 * their descriptors. */
int pc_amr_lts_init(pc_amr *amr, const int32_t *nptr, const int32_t *nbr);
/* k grid iterations: the advance flags from the clocks at the start of each,
 * restriction of every leaf into the virtual patches on its own interval,
 * ghost faces of the advancing leaves (time factors from the clocks), uNew
 * -> uOld, walls, and per advancing leaf solveTimestep with
 * maximumTimestepSize = end_time - (time + dt); *leaf_steps = leaf steps
 * taken (0: every leaf has reached end_time).  The leaves' descriptors
 * carry their clocks afterwards; the plan no longer runs pc_amr_step. */
int pc_amr_lts_step(pc_amr *amr, int k, double end_time, long long *leaf_steps);
int pc_amr_destroy(pc_amr *amr);

/* Raw device pointers of the pool (for zero-copy interop); state = 0/1.
 * The library's own work runs on the pool's stream and some calls return
 * before it has finished (pc_solve_timestep returns once dt is known, while
 * the update still runs; the batched steps queue): call
 * pc_pool_synchronize before touching these arrays from another stream or
 * from the host. */
int pc_pool_device_fields(pc_pool *pool, int state, double **h, double **hu,
                          double **hv, double **b);

#ifdef __cplusplus
}
#endif
#endif
