/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64 CPU restatement of PeanoClaw's patch shallow-water path,
 * written from the reference's source text.  Only tests/, the smoke() entry
 * point and bench.py's CPU-baseline leg may load it, and only as the checker.
 * The product (peanoclaw_amd) never links, imports or calls anything here.
 *
 * Reference files followed (paths relative to the reference repository):
 *   tools/SWE_WaveAccumulationBlock.cpp      edge loops, accumulation, CFL, update
 *   src/peanoclaw/native/SWEKernel.cpp       solveTimestep, fillBoundaryLayer,
 *                                            h <-> h+b transforms
 *   src/peanoclaw/interSubgridCommunication/DefaultTransfer.cpph
 *   src/peanoclaw/interSubgridCommunication/DefaultInterpolation.cpph
 *   src/peanoclaw/interSubgridCommunication/DefaultRestriction.cpph (+ .h)
 *   src/peanoclaw/interSubgridCommunication/GhostLayerCompositor.cpp
 *   src/peanoclaw/interSubgridCommunication/GhostlayerCompositorFunctors.cpp
 *   src/peanoclaw/interSubgridCommunication/aspects/{Face,Edge}AdjacentPatchTraversal.cpph
 *   src/peanoclaw/geometry/Region.cpp
 *   src/peanoclaw/grid/LinearizationZYXQ.h, grid/TimeIntervals.cpp, Patch.cpp
 *   src/peanoclaw/native/scenarios/BreakingDam.cpp
 *   src/peanoclaw/mappings/SolveTimestep.cpp (fillBoundaryLayers, leaveCell)
 *
 * Parity status per function (what pins it) is listed in the header comment
 * of swe_oracle.c.  The edge solver (or_fwave) restates the published f-wave
 * method because the reference's Riemann solver lives in the absent SWE
 * library: it is pinned by invariants, not by reference vectors.
 */
#ifndef PEANOCLAW_SWE_ORACLE_H
#define PEANOCLAW_SWE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Fixed solver constants of this build (see swe_oracle.c). */
#define OR_GRAVITY 9.81
#define OR_DRY_TOL 0.01
#define OR_ZERO_TOL 1e-7
#define OR_CFL 0.4
#define OR_NUMERICAL_ZERO 1e-12

/* A patch ("subgrid") in the reference's own data layout
 * (grid/LinearizationZYXQ.h:24-82): q slowest, then x, y fastest.
 *   uNew [q][x][y]               q < nq, x < nx, y < ny
 *   uOld [q][x+gw][y+gw]         x, y in [-gw, n+gw)
 *   pwg  [p][x+gw][y+gw]         parameters with ghost layer (SWE: b)   */
typedef struct {
  int nx, ny, gw, nq, npw;
  double pos[2], size[2];
  double subcell[2];            /* size[d] / n[d]  (Patch.cpp:38) */
  int level;
  int is_virtual, is_leaf, is_valid;
  double time, dt, est_next;
  double min_neighbor_time_constraint;
  double min_leaf_neighbor_time_constraint;
  double min_neighbor_time, max_neighbor_dt;   /* virtual time interval */
  double lower_ghost_bounds[2], upper_ghost_bounds[2];
  int cell_description_index;
  double *unew, *uold, *pwg;
} or_patch;

/* ---- layout ---- */
int or_unew_index(const or_patch *p, int q, int x, int y);
int or_uold_index(const or_patch *p, int q, int x, int y);
void or_patch_init(or_patch *p, int nq, int npw, int nx, int ny, int gw,
                   double px, double py, double sx, double sy, int level,
                   double time, double dt, double min_neighbor_time,
                   double *unew, double *uold, double *pwg);

/* ---- edge solver (f-wave) ---- */
void or_fwave(double hL, double hR, double huL, double huR, double bL, double bR,
              double out[5]); /* dhL, dhR, dqL, dqR, speed */
double or_fwave_speed(double hL, double hR, double huL, double huR);
void or_fwave_batch(long long n, const double *in, double *out);
/* augmented Riemann solver (George 2008); see swe_oracle.c */
void or_aug_middle_state(double hL, double hR, double uL, double uR, int maxiter,
                         double out[5]);
int or_aug_waves(double hL, double hR, double huL, double huR, double bL, double bR,
                 double sw[3], double fh[3], double fq[3]);
void or_augrie(double hL, double hR, double huL, double huR, double bL, double bR,
               double out[5]);
/* edge solver of the block: 1, 4 = f-wave, 2 = augmented Riemann */
int or_set_wave_propagation_solver(int solver);
void or_edge(double hL, double hR, double huL, double huR, double bL, double bR,
             double out[5]);
double or_edge_speed(double hL, double hR, double huL, double huR, double bL, double bR);
void or_edge_batch(long long n, const double *in, double *out);

/* ---- SWE block on one (nx+2)x(ny+2) tile, ghost width 1 ---- */
double or_block_fluxes(int nx, int ny, double dx, double dy,
                       const double *h, const double *hu, const double *hv,
                       const double *b, double *netH, double *netHu,
                       double *netHv);
double or_block_max_timestep(double maxWaveSpeed, double dx, double dy);
/* number of cells whose depth the update clamped from below 0 to 0 (the
   only place mass changes on a single level); test diagnostics */
extern long long or_clamp_events;
void or_block_update(int nx, int ny, double dt, double *h, double *hu,
                     double *hv, double *netH, double *netHu, double *netHv);

/* ---- Numerics (SWEKernel) on or_patch ---- */
double or_solve_timestep(or_patch *p, double maximumTimestepSize);
double or_patch_cfl(or_patch *p);
void or_fill_boundary_layer(or_patch *p, int dimension, int setUpper);
void or_copy_unew_to_uold(or_patch *p);
void or_swap_unew_uold(or_patch *p);
void or_transfer_ghostlayer(const int size[2], const int srcOff[2],
                            const int dstOff[2], or_patch *src, or_patch *dst);
void or_interpolate_default(const int dstSize[2], const int dstOff[2],
                            or_patch *src, or_patch *dst, int toUOld,
                            int toCurrentTime, int useTimeUNewOrUOld);
/* 0, or -1 (nothing changed) when the destination region or the grown
 * source region it maps to is outside its patch */
int or_interpolate_swe(const int dstSize[2], const int dstOff[2],
                       or_patch *src, or_patch *dst, int toUOld,
                       int toCurrentTime, int useTimeUNewOrUOld);
void or_restrict_default(or_patch *src, or_patch *dst, int onlyOverlapped);
void or_restrict_swe(or_patch *src, or_patch *dst, int onlyOverlapped);
/* the SWE transfers between levels with the h <-> h+b transform applied to
   the values read, the source left unchanged (batched AMR semantics) */
void or_restrict_swe_pure(or_patch *src, or_patch *dst);
int or_interpolate_swe_pure(const int dstSize[2], const int dstOff[2], or_patch *src,
                            or_patch *dst, int toUOld, int toCurrentTime,
                            int useTimeUNewOrUOld);
void or_post_process_restriction_swe(or_patch *dst, int onlyOverlapped);
/* Patch::switchToLeaf: restricted [0, 1]-basis values -> the patch's own
   interval [time, time + dt] (set p->time, p->dt first) */
void or_switch_to_leaf(or_patch *p);

/* ---- flux correction (DefaultFluxCorrection; off by default) ---- */
int or_flux_face_offset(const or_patch *p, int dim, int direction);
int or_flux_array_size(const or_patch *p);
void or_compute_fluxes(or_patch *p, double *flux);
int or_apply_flux_correction(or_patch *src, const double *srcFlux, or_patch *dst,
                             const double *dstFlux, int dimension, int direction);
void or_transform_water_height(or_patch *p, const int off[2], const int size[2],
                               int modifyUOld, int absoluteToAboveSeaFloor);

/* ---- geometry::Region ---- */
double or_overlapping_region(const double p1[2], const double s1[2],
                             const double p2[2], const double s2[2]);
void or_region_map_to_patch(const int off[2], const int size[2],
                            const or_patch *src, const or_patch *dst,
                            double epsilon, int outOff[2], int outSize[2]);
int or_regions_overlapped_by_ghostlayers(const double lower[2],
                                         const double upper[2],
                                         const double srcPos[2],
                                         const double srcSize[2],
                                         const double srcSubcell[2],
                                         const int srcN[2], int outOff[8],
                                         int outSize[8]);

/* ---- GhostLayerCompositor over the 4 patches around a vertex ---- */
typedef struct { int src, dst, dir[2]; } or_traversal_call;
int or_face_traversal_order(or_traversal_call out[8]);
int or_edge_traversal_order(or_traversal_call out[4]);
int or_should_transfer(const or_patch *src, const or_patch *dst);
/* numerics_kind: 0 = default strategies (NumericsTestStump), 1 = SWE */
void or_fill_ghostlayers_and_update_neighbor_times(or_patch *patches[4],
                                                   int level, int dest,
                                                   int numerics_kind);

/* ---- TimeIntervals ---- */
double or_time_uold(const or_patch *p);
double or_time_unew(const or_patch *p);
int or_is_blocked_by_neighbors(const or_patch *p);
void or_reset_min_neighbor_time_constraint(or_patch *p);
void or_update_neighbor_time(or_patch *updated, const or_patch *neighbor);
void or_minimal_fine_grid_interval(const double *t, const double *dt, int n,
                                   double out[2]);

/* ---- BreakingDam scenario + refinement flags ---- */
void or_breaking_dam_init(or_patch *p, double cx, double cy, double radius,
                          double hIn, double hOut);
void or_demanded_mesh_width(const or_patch *p, double threshold,
                            double minimalWidth, double maximalWidth,
                            double out[2]);
/* returns 1 = refine, 0 = coarsening allowed, 2 = keep (erase veto) */
int or_refinement_decision(const double subcell[2], const double demanded[2]);

/* ---- uniform grid of equal patches, global dt (configs 0 and parity cases) */
typedef struct {
  int px, py;          /* patches per dimension */
  int n;               /* cells per patch per dimension */
  double dx, dy;
  double *h[2], *hu[2], *hv[2], *b;   /* [patch][(n+2)*(n+2)], two states */
  int cur;
} or_grid;
void or_grid_fill_ghosts(or_grid *gr);
void or_grid_fill(int px, int py, int n, const int *slot_of,
                  const long long *off, double *h, double *hu, double *hv);
void or_grid_fill_rect(int px, int py, int nx, int ny, const int *slot_of,
                       const long long *off, double *h, double *hu, double *hv);
double or_grid_step(or_grid *gr, double maxDt, double *out_dtcfl);
double or_grid_max_speed(const or_grid *gr);

/* ---- generic patch list with a neighbour table, global dt --------------
 * ghost strips: each entry copies one source cell (q = 0..2) of a source
 * patch into a ghost cell of a destination patch, optionally negating q1 or
 * q2 (reflective wall, SWEKernel.cpp:176-237).                            */
typedef struct {
  int dst, dst_cell;   /* patch id, linear index into the (n+2)^2 tile     */
  int src, src_cell;   /* patch id, linear index into the (n+2)^2 tile     */
  int negate;          /* bit 0 negates hu, bit 1 negates hv               */
} or_ghost_entry;
typedef struct {
  int np;
  const int *nx, *ny;
  const long long *off;       /* start of each patch in the field arrays */
  const double *dx, *dy;
  double *h[2], *hu[2], *hv[2];
  const double *b;
  int cur;
} or_pool;
void or_pool_fill_ghosts(or_pool *pl, const or_ghost_entry *e, long long ne);
double or_pool_step(or_pool *pl, double maxDt, int nthreads);
void or_pool_max_speeds(const or_pool *pl, int state, double *smax);

#ifdef __cplusplus
}
#endif
#endif
