// HIP kernels of the batched SWE patch step on gfx950.
//
// Patch pool layout in HBM (structure of arrays): every field (h, hu, hv in
// two states for ping-pong, b once) is one device array; patch p owns the
// contiguous tile [off_p, off_p + (nx+2)·stride_p) in each of them, stored in
// the reference's order (x rows, y contiguous; grid/LinearizationZYXQ.h),
// row x of the tile (ghost rows included) at off_p + x·stride_p.  For ny >= 8
// the rows are padded to a multiple of 16 doubles and off_p = 15 mod 16, so
// the interior of every row (columns 1..ny) starts on a 128-byte cache line:
// a wavefront's load of a row's interior touches ny/16 lines, not one more.
// Narrower patches keep stride ny + 2.
//
// The update is a row march: one 64-lane wavefront owns a task of up to
// 64/W patches of the same size (W = ny rounded up to a power of two); lane
// (seg, col) walks down column `col` of patch `seg`, keeps the previous row
// in registers, solves the x-edge between the two rows and the y-edge to its
// left neighbour lane (cell terms handed over by DPP wavefront shifts), and
// finalises the previous row once both of its x-edges are known.  Each cell
// is read from HBM once and written once.  The two y-edges per row that
// touch the ghost columns are solved up front by a pre-pass in which the
// lanes take rows.
//
// Ghost values in the batched step come straight from where they live
// (the ghost fill of the reference, DefaultTransfer at time factor 1 and
// SWEKernel::fillBoundaryLayer, fused into the loads): a ghost row is read
// from the neighbour's boundary row, a ghost column from the neighbour's
// edge buffer (its first and last interior column, written contiguously by
// the previous update), a wall side from the patch's own boundary with the
// normal momentum negated.  Side kind 0 reads the tile's own ghost cells
// (filled by the gather kernel, e.g. across refinement levels).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "swe_device.hpp"

namespace pcamd {

constexpr int kERows = 64;  // edge-buffer rows per side (max nx)

// Row stride of a tile with ny interior columns, and the residue of its
// offset modulo kTileAlign (interior rows on 128-byte lines for ny >= 8).
constexpr int kTileAlign = 16;
__host__ __device__ constexpr int tile_stride(int ny) {
  return ny >= 8 ? (ny + 2 + kTileAlign - 1) / kTileAlign * kTileAlign : ny + 2;
}
__host__ __device__ constexpr int tile_phase(int ny) { return ny >= 8 ? kTileAlign - 1 : -1; }


// Global-memory accessors: loads and stores through address space 1, so the
// compiler emits global_* instead of flat_* instructions (flat accesses
// also count against the LDS wait counter, which would make every LDS
// hand-off wait for the prefetched rows).
template <class T>
__device__ __forceinline__ T gl(const T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return *(const __attribute__((address_space(1))) T*)p;
#else
  return *p;
#endif
}
template <class T>
__device__ __forceinline__ void gs(T* p, T v) {
#if defined(__HIP_DEVICE_COMPILE__)
  *(__attribute__((address_space(1))) T*)p = v;
#else
  *p = v;
#endif
}

struct PatchDev {
  long long off;  // tile base in the field arrays
  int nx, ny;
  int scn;        // scenario of the patch
  int stride;     // row stride of the tile (tile_stride(ny))
  double dx, dy;  // subcell size (Patch.cpp:38)
  double dxinv, dyinv;
};

struct Fields {
  double* h[2];
  double* hu[2];
  double* hv[2];
  double* b;
  // state s picked with a select: a run-time index into these arrays
  // would place the argument struct in scratch memory
  __device__ __forceinline__ double* H(int s) const { return s ? h[1] : h[0]; }
  __device__ __forceinline__ double* HU(int s) const { return s ? hu[1] : hu[0]; }
  __device__ __forceinline__ double* HV(int s) const { return s ? hv[1] : hv[0]; }
};

// Where the four ghost sides of a patch come from in the batched step:
// 0 = x-lower (row 0), 1 = x-upper (row nx+1), 2 = y-lower (col 0),
// 3 = y-upper (col ny+1).
enum SideKind { kSideTile = 0, kSideNeighbour = 1, kSideWall = 2 };
struct SideDev {
  int kind[4];
  int src[4];
};

// Ghost-edge sources of one live patch for the boundary speed pass, resolved
// on the host when the grid changes (one load chain instead of patch ->
// sides -> neighbour patch -> data):
//   xrow[s]: index of column 0 of the row across x side s (the neighbour's
//            boundary row, the own boundary row for a wall (hu negated), or
//            the own ghost row);
//   ebo[s]:  offset of the own edge-buffer column of y side s;
//   ebg[s]:  offset of the neighbour's edge-buffer column across y side s,
//            -1 for a wall (own column, hv negated), -2 for the own ghost
//            column of the tile;
//   shared[d]: the lower side in dimension d is a neighbour of the same
//            subcell size whose upper side points back here: the edges
//            between the two belong to that neighbour's upper side.
struct BoundaryDesc {
  long long off;
  long long xrow[2];
  long long ebo[2];
  long long ebg[2];
  int nx, ny, scn, slot;
  int xwall[2];
  int shared[2];
  int rseg;  // row segments of the update (seams: seg_first_row(k >= 1))
  int stride;  // row stride of the tile
  double dx, dy;
};

// Edge buffers: h and hv of the first (side 0) and last (side 1) interior
// column of every patch, per state: e[s][((p*2 + side)*2 + q)*kERows + row-1]
// (q 0 = h, 1 = hv); bcol[(p*4 + k)*kERows + row-1] = b at columns
// 0, 1, ny, ny+1 (static).
struct EdgeBufs {
  double* e[2];
  const double* bcol;
};

struct Task {
  int first;  // index into the task patch list
  int nseg;   // patches in this task (same nx, ny)
  int nx, ny;
  // row segment: the task finalises interior rows r0..r1 (r1 = 0: nx) of its
  // patches; segment sidx of the patch's row segments (see seg_first_row)
  int r0 = 1, r1 = 0, sidx = 0;
};

// Row segments of a patch: a launch with fewer tasks than the GPU holds
// splits the rows of each patch into rseg segments (1, 2 or 4), each marched
// by its own wavefront.  Segment k starts at row seg_first_row(k, nx, rseg);
// the update's next-step interior speed maximum of segment k of patch p is
// kept in smax_int[p * kMaxRowSeg + k].
constexpr int kMaxRowSeg = 4;
__host__ __device__ __forceinline__ int seg_first_row(int k, int nx, int rseg) {
  return 1 + (k * nx) / rseg;
}

struct ScnDev {
  double time;        // TimeIntervals "time" (start of the last interval)
  double dt;          // TimeIntervals timestep size
  double est;         // estimatedNextTimestepSize (min over patches)
  double end_time;    // global timestep end time
  double dt_applied;  // dt of the current step, 0 if it did not advance
  long long steps;
  unsigned long long minbits;  // running min of the patch CFL steps (bits)
  int stepc;          // parity counter: state (stepc & 1) is current
  int advance;        // current step advances this scenario
  int pbeg, pend;     // its patches in the scenario patch list
  int redo;           // pruned CFL minimum not proven exact: fallback pass
  int pad;
};

// Pruned CFL pass (launch mode fused = 2): the update computes exact
// next-state edge speeds only where a cheap bound (bound_terms) can reach
// kCfl * min(dx, dy) / (est * slack), est = the scenario's last CFL step.
// Any edge left out is slower than that, so its patch step exceeds
// est * slack * (1 - 1e-6) (bound margin 2^-10 minus FP32 and rounding
// error); the next step accepts the pruned minimum only if it is
// <= est * slack * kPruneCheckFrac, otherwise an exact speed pass runs
// first.  dt is therefore the exact minimum over all edges for any slack > 0.
constexpr double kPruneSlackDefault = 1.03125;
constexpr double kPruneCheckFrac = 1.0 - 1.0 / 256.0;

struct GhostStrip {
  int dst_slot, dst_cell, dst_stride;
  int src_slot, src_cell, src_stride;
  int length, negate;
};

// kUpdateFromOld: the reads and writes of kUpdateSwap without its uOld
// store, for callers that copied uNew -> uOld before the update (the store
// would write the values already there); with nothing written into the tile
// it reads, row segments of one patch may run in parallel.
enum MarchMode { kSpeed = 0, kUpdate = 1, kUpdateSwap = 2, kUpdateFromOld = 3 };

__device__ __forceinline__ int pow2ceil(int v) {
  int w = 1;
  while (w < v) w <<= 1;
  return w;
}

// Intra-wavefront LDS hand-off: LDS operations of one wavefront complete in
// order, so lanes only need the compiler not to move LDS accesses across
// this point (no s_barrier, no wait on outstanding global loads).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane exchange of the march's y-sweep through DPP wavefront shifts:
// from_left(v) is lane-1's v (lane 0 keeps its own), from_right(v) lane+1's
// (lane 63 keeps its own).  One v_mov_b32_dpp per dword, no LDS round trip.
__device__ __forceinline__ int dpp_shr1(int v) {
  return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false);  // wave_shr:1
}
__device__ __forceinline__ int dpp_shl1(int v) {
  return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false);  // wave_shl:1
}
__device__ __forceinline__ double from_left(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_shr1((int)(unsigned)(b & 0xffffffffLL));
  const int hi = dpp_shr1((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double from_right(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_shl1((int)(unsigned)(b & 0xffffffffLL));
  const int hi = dpp_shl1((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ float from_left(float v) {
  return __int_as_float(dpp_shr1(__float_as_int(v)));
}
// from_left for values whose lane-0 copy is never used: lane 0 reads 0
// (DPP bound_ctrl), so v need not be copied into the destination first.
__device__ __forceinline__ int dpp_shr1_z(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true);  // wave_shr:1, bound_ctrl
}
__device__ __forceinline__ double from_left_z(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_shr1_z((int)(unsigned)(b & 0xffffffffLL));
  const int hi = dpp_shr1_z((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ float from_left_z(float v) {
  return __int_as_float(dpp_shr1_z(__float_as_int(v)));
}

// LDS scratch of one wavefront (64 lanes): the pre-pass's boundary y-edges,
// [side*2 + k][seg*nx + row-1].
struct MarchLds {
  double pre[4][64];
};

struct MarchArgs {
  const int* plist;
  const PatchDev* pd;
  Fields F;
  const SideDev* sides;  // nullptr: ghost cells from the tile
  EdgeBufs E;
  int par;
  double dt;
  double* smax;       // kSpeed: full per-patch max; fused: next-state interior max
  const ScnDev* scn;  // pruned mode: the scenario's last CFL step (threshold)
  double slack;       // pruned mode: threshold slack
  // per-patch steps (PERSEG): the step of patch p is seg_dt[p], and only
  // patches with seg_adv[p] != 0 are updated (local time stepping)
  const double* seg_dt = nullptr;
  const int* seg_adv = nullptr;
  // slots per patch of `smax` (kMaxRowSeg for the update's interior maxima
  // of row segments; a whole-patch kSpeed pass fills slot 0 and clears the
  // others)
  int smax_stride = 1;
};

// Tile accesses of the march.  With O32 the field arrays hold fewer than
// 2^29 doubles, so an element index fits 32 bits with its byte offset: every
// access is then a uniform array base plus one 32-bit byte offset, which the
// loads and stores take as their saddr / voffset operands, instead of a
// 64-bit address computed per access.
template <bool O32>
struct TileMem {
  using I = long long;
  __device__ static __forceinline__ double ld(const double* b, I i) { return gl(b + i); }
  __device__ static __forceinline__ void st(double* b, I i, double v) { gs(b + i, v); }
};
template <>
struct TileMem<true> {
  using I = unsigned;
  __device__ static __forceinline__ double ld(const double* b, I i) {
    return gl(reinterpret_cast<const double*>(reinterpret_cast<const char*>(b) + (i << 3)));
  }
  __device__ static __forceinline__ void st(double* b, I i, double v) {
    gs(reinterpret_cast<double*>(reinterpret_cast<char*>(b) + (i << 3)), v);
  }
};

// FUSE: 0 = no next-state speeds, 1 = exact speeds of every interior edge of
// the new state, 2 = pruned (exact speeds only for bound candidates).
// PERSEG: every patch of the task has its own step and advance flag.
// O32: 32-bit tile offsets (TileMem).
template <int MODE, bool SIDES, int FUSE, int SOLVER, bool PERSEG = false, bool O32 = false>
__device__ __forceinline__ void march_task(const Task& t, const MarchArgs& A, MarchLds& L) {
  using M = TileMem<O32>;
  using I = typename M::I;
  constexpr bool FUSED = FUSE != 0;
  constexpr bool PRUNE = FUSE == 2;
  // the f-wave's speeds do not depend on b; the augmented solver's do (the
  // wall decision at a dry side), so its speed passes read b as well
  constexpr bool SPEED_B = SOLVER != 0;
  static_assert(SOLVER == 0 || FUSE == 0, "next-step speeds are fused for the f-wave only");
  using ES = EdgeSolver<SOLVER>;
  constexpr bool DUMMY_FREE = MODE != kSpeed && ES::kDummyLanesFree;
  const int lane = threadIdx.x & 63;
  const int nx = t.nx, ny = t.ny;
  const int W = pow2ceil(ny);
  const int seg = lane / W;
  const int col = lane - seg * W + 1;  // 1-based interior column
  const bool segok = seg < t.nseg;
  // the update's y-edges on col-1 lanes are dummies (see the y-sweep)
  const unsigned long long col1_lanes = __builtin_amdgcn_ballot_w64(col == 1);
  const int p = gl(A.plist + t.first + (segok ? seg : 0));
  const bool active = segok && (col <= ny) && (!PERSEG || gl(A.seg_adv + p) != 0);
  // lane masks for the pruned mode's candidate votes
  const unsigned long long act_lanes = __builtin_amdgcn_ballot_w64(active);
  const PatchDev P = gl(A.pd + p);
  const int stride = P.stride;  // the same for every patch of the task
  const int par = A.par;
  constexpr bool FROM_OLD = MODE == kUpdateSwap || MODE == kUpdateFromOld;
  const int src = FROM_OLD ? 1 - par : par;
  const int dst = FROM_OLD ? par : 1 - par;
  const Fields& F = A.F;
  // kUpdateSwap writes into the tile it reads (rows already consumed), so
  // its source pointers must not be declared non-aliasing.
  using SrcPtr = typename std::conditional<MODE == kUpdateSwap, const double*,
                                           const double* __restrict__>::type;
  // state pointers picked with selects (a run-time index into the argument
  // arrays would place them in scratch memory)
  double* const Hsrc = src ? F.h[1] : F.h[0];
  double* const HUsrc = src ? F.hu[1] : F.hu[0];
  double* const HVsrc = src ? F.hv[1] : F.hv[0];
  double* const Hdst = dst ? F.h[1] : F.h[0];
  double* const HUdst = dst ? F.hu[1] : F.hu[0];
  double* const HVdst = dst ? F.hv[1] : F.hv[0];
  SrcPtr Hs = Hsrc;
  SrcPtr HUs = HUsrc;
  SrcPtr HVs = HVsrc;
  const double* __restrict__ Bs = F.b;
  const double* Ecur = SIDES ? (par ? A.E.e[1] : A.E.e[0]) : nullptr;
  double* const Edst = SIDES ? (dst ? A.E.e[1] : A.E.e[0]) : nullptr;

  // rows finalised by this task (a row segment, or all rows)
  const int ra = t.r0, rb = t.r1 ? t.r1 : nx;
  const int nrows = rb - ra + 1;

  // ---- pre-pass: the y-edges (row, 0|1) and (row, ny|ny+1) ----
  // (after the previous task's last reads of L.pre on this wavefront)
  wave_sync();
  const int npre = t.nseg * nrows * 2;
  for (int e = lane; e < npre; e += 64) {
    const int s = e / (2 * nrows);
    const int rem = e - s * 2 * nrows;
    const int row = (rem >> 1) + ra;
    const int side = rem & 1;
    const int ps = gl(A.plist + t.first + s);
    const PatchDev Q = gl(A.pd + ps);
    double hL, hvL, bL, hR, hvR, bR;
    const long long iL = Q.off + (long long)row * stride + (side ? ny : 0);
    if (SIDES) {
      const int k = gl(&A.sides[ps].kind[2 + side]);
      const double* Eo = Ecur + (long long)((ps * 2 + side) * 2) * kERows + row - 1;
      const double* Bc = A.E.bcol + (long long)ps * 4 * kERows + row - 1;
      const double ho = gl(Eo), hvo = gl(Eo + kERows);  // own (row,1) or (row,ny)
      double hg, hvg;  // the ghost (row,0) or (row,ny+1)
      if (k == kSideNeighbour) {
        const int q = gl(&A.sides[ps].src[2 + side]);
        const double* Eq = Ecur + (long long)((q * 2 + (1 - side)) * 2) * kERows + row - 1;
        hg = gl(Eq);
        hvg = gl(Eq + kERows);
      } else if (k == kSideWall) {
        hg = ho;
        hvg = -hvo;
      } else {
        const long long ig = side ? iL + 1 : iL;
        hg = gl(Hs + ig);
        hvg = gl(HVs + ig);
      }
      if (side == 0) {
        hL = hg; hvL = hvg; bL = gl(Bc);
        hR = ho; hvR = hvo; bR = gl(Bc + kERows);
      } else {
        hL = ho; hvL = hvo; bL = gl(Bc + 2 * kERows);
        hR = hg; hvR = hvg; bR = gl(Bc + 3 * kERows);
      }
    } else {
      hL = gl(Hs + iL); hvL = gl(HVs + iL); bL = gl(Bs + iL);
      hR = gl(Hs + iL + 1); hvR = gl(HVs + iL + 1); bR = gl(Bs + iL + 1);
    }
    if (MODE == kSpeed && !SPEED_B) { bL = 0.0; bR = 0.0; }
    const CellQ cl = cell_dir(cell_base(hL, bL), hvL);
    const CellQ cr = cell_dir(cell_base(hR, bR), hvR);
    const int slot = s * nx + row - 1;
    if (MODE == kSpeed) {
      L.pre[side][slot] = ES::speed(cl, cr);
    } else {
      double dhL, dhR, dqL, dqR;
      ES::solve(cl, cr, dhL, dhR, dqL, dqR);
      // side 0: the right cell (row,1) takes the "upwards" part;
      // side 1: the left cell (row,ny) takes the "downwards" part.
      L.pre[side * 2 + 0][slot] = Q.dyinv * (side ? dhL : dhR);
      L.pre[side * 2 + 1][slot] = Q.dyinv * (side ? dqL : dqR);
    }
  }
  wave_sync();

  const I base = (I)(P.off + col);
  const double dxinv = P.dxinv, dyinv = P.dyinv;
  const double dt = PERSEG ? gl(A.seg_dt + p) : A.dt;
  double smaxl = 0.0;
  // pruned mode: speed threshold of this lane's patch (0 = every edge exact)
  float thr = 0.0f;
  if (PRUNE) {
    const double est = gl(&A.scn[P.scn].est);
    if (est > 0.0) thr = (float)(kCfl * dmin(P.dx, P.dy) / (est * A.slack));
  }
  // ghost-row sources (x sides), kept in scalars
  int kx0 = kSideTile, kx1 = kSideTile;
  I gx0 = 0, gx1 = 0;  // index of the first interior cell of the ghost row source
  bool neg0 = false, neg1 = false;
  if (SIDES) {
    kx0 = gl(&A.sides[p].kind[0]);
    kx1 = gl(&A.sides[p].kind[1]);
    if (kx0 == kSideNeighbour) {
      const PatchDev Q = gl(A.pd + gl(&A.sides[p].src[0]));
      gx0 = (I)(Q.off + (long long)Q.nx * stride + col);
    } else if (kx0 == kSideWall) {
      gx0 = base + (I)stride;
      neg0 = true;
    } else {
      gx0 = base;
    }
    if (kx1 == kSideNeighbour) {
      const PatchDev Q = gl(A.pd + gl(&A.sides[p].src[1]));
      gx1 = (I)(Q.off + stride + col);
    } else if (kx1 == kSideWall) {
      gx1 = base + (I)nx * (I)stride;
      neg1 = true;
    } else {
      gx1 = base + (I)(nx + 1) * (I)stride;
    }
  }

  // row loads (r = 0 and nx+1 are ghost rows; their hv is never used)
  auto load_row = [&](int r, double& h, double& hu, double& hv, double& b) {
    h = 1.0; hu = 0.0; hv = 0.0; b = 0.0;
    if (!active) return;
    const I i = base + (I)r * (I)stride;
    if (MODE != kSpeed || SPEED_B) b = M::ld(Bs, i);
    if (SIDES && (r == 0 || r == nx + 1)) {
      const I j = r == 0 ? gx0 : gx1;
      h = M::ld(Hs, j);
      hu = M::ld(HUs, j);
      if (r == 0 ? neg0 : neg1) hu = -hu;
      return;
    }
    h = M::ld(Hs, i);
    hu = M::ld(HUs, i);
    hv = M::ld(HVs, i);
  };

  // previous row (x-sweep terms + raw values)
  CellQ px = {};
  double phv = 0.0;
  // accumulators of the previous row
  double accH = 0.0, accHu = 0.0, yUpH = 0.0, yUpHv = 0.0, yDnH = 0.0, yDnHv = 0.0;
  // fused: x-sweep speed terms of the previous new row
  CellQ pnx = {};
  CellBase nbase = {};  // new row terms awaiting their y exchange
  double nhv = 0.0;
  bool have_new = false;
  // pruned: bound terms and raw h, hu of the latest new row
  BoundTerms pbt = {0.0f, 0.0f, 0.0f};
  double pnh = 1.0, pnhu = 0.0;

  double h, hu, hv, b;
  load_row(ra - 1, h, hu, hv, b);
  // wet lanes of the previous row (DUMMY_FREE: the update takes the wetness
  // votes of its edges from the rows' masks)
  unsigned long long wet_prev = 0;
  for (int r = ra - 1; r <= rb + 1; ++r) {
    const double ch = h, chu = hu, chv = hv, cbv = b;
    if (r < rb + 1) load_row(r + 1, h, hu, hv, b);  // prefetch the next row
    const CellBase cb = cell_base(ch, cbv);
    const CellQ cx = cell_dir(cb, chu);
    const unsigned long long wet_row = __builtin_amdgcn_ballot_w64(ch >= kDryTol);

    if (r >= ra) {
      // x-edge (r-1 | r)
      if (MODE == kSpeed) {
        smaxl = dmax(smaxl, ES::speed(px, cx));
      } else {
        double dhL, dhR, dqL, dqR;
        if (DUMMY_FREE)
          ES::solve_voted((wet_prev & wet_row) == __builtin_amdgcn_read_exec(), px, cx, 0,
                          dhL, dhR, dqL, dqR);
        else
          ES::solve(px, cx, dhL, dhR, dqL, dqR);
        if (r >= ra + 1) {
          // finalise row r-1: x-left (already in acc), x-right, y-up, y-down
          double nH = accH + dxinv * dhL;
          double nHu = accHu + dxinv * dqL;
          nH = nH + yUpH;
          nH = nH + yDnH;
          double nHv = 0.0 + yUpHv;
          nHv = nHv + yDnHv;
          double h1 = px.h - dt * nH;
          double hu1 = px.q - dt * nHu;
          double hv1 = phv - dt * nHv;
          if (h1 < kClampH) { hu1 = 0.0; hv1 = 0.0; }
          if (h1 < 0.0) h1 = 0.0;
          if (active) {
            const I o = base + (I)(r - 1) * (I)stride;
            if (MODE == kUpdateSwap) {
              // uOld interior <- old uNew, uNew <- new values
              M::st(Hsrc, o, M::ld(Hdst, o));
              M::st(HUsrc, o, M::ld(HUdst, o));
              M::st(HVsrc, o, M::ld(HVdst, o));
            }
            M::st(Hdst, o, h1);
            M::st(HUdst, o, hu1);
            M::st(HVdst, o, hv1);
            if (MODE == kUpdate && stride == ny + 2 && (col == 1 || col == ny)) {
              // Packed rows (ny < 8): also store into the row's ghost column
              // (0 or ny+1) of the new state, which no one reads before it is
              // refilled: the row's cache lines are then written completely
              // (no masked partial writebacks at the row boundaries).  Padded
              // rows are whole lines without them.
              const I og = col == 1 ? o - 1 : o + 1;
              M::st(Hdst, og, h1);
              M::st(HUdst, og, hu1);
              M::st(HVdst, og, hv1);
            }
            if (SIDES && (col == 1 || col == ny)) {
              double* Ed = Edst + (long long)((p * 2 + (col == 1 ? 0 : 1)) * 2) * kERows + r - 2;
              gs(Ed, h1);
              gs(Ed + kERows, hv1);
            }
          }
          if (PRUNE) {
            const BoundTerms bt = bound_terms(h1, hu1, hv1);
            if (r >= ra + 2) {
              const bool cand = bound_candidate(pbt.ax, pbt.c, bt.ax, bt.c, thr);
              if ((__builtin_amdgcn_ballot_w64(cand) & act_lanes) != 0) {
                const CellQ ea = cell_dir(cell_base(pnh, 0.0), pnhu);
                const CellQ eb = cell_dir(cell_base(h1, 0.0), hu1);
                const double sp = ES::speed(ea, eb);
                if (active) smaxl = dmax(smaxl, sp);
              }
            }
            pbt = bt;
            pnh = h1;
            pnhu = hu1;
            nhv = hv1;
            have_new = true;
          } else if (FUSED) {
            nbase = cell_base(h1, 0.0);
            const CellQ ncx = cell_dir(nbase, hu1);
            if (r >= ra + 2) {
              const double sp = ES::speed(pnx, ncx);
              if (active) smaxl = dmax(smaxl, sp);
            }
            pnx = ncx;
            nhv = hv1;
            have_new = true;
          }
        }
        accH = 0.0 + dxinv * dhR;
        accHu = 0.0 + dxinv * dqR;
      }
    }

    if (r >= ra && r <= rb) {
      // y-edges of row r: lane col solves (col-1 | col) for col >= 2
      const CellQ cy = cell_dir(cb, chv);
      // left neighbour lane; on a lane without one (col 1) the y-edge is a
      // dummy whose results are not used.  Where the solver takes any
      // operands on such lanes (DUMMY_FREE), they are left out of its choice
      // of path and lane 0 reads zeros.
      CellQ cl;
      if (DUMMY_FREE) {
        cl.h = from_left_z(cy.h); cl.q = from_left_z(cy.q); cl.b = from_left_z(cy.b);
        cl.u = from_left_z(cy.u); cl.sq = from_left_z(cy.sq); cl.c = from_left_z(cy.c);
        cl.f = from_left_z(cy.f);
      } else {
        cl.h = from_left(cy.h); cl.q = from_left(cy.q); cl.b = from_left(cy.b);
        cl.u = from_left(cy.u); cl.sq = from_left(cy.sq); cl.c = from_left(cy.c);
        cl.f = from_left(cy.f);
      }
      CellQ ncy = {};
      if (FUSED && !PRUNE && have_new) ncy = cell_dir(nbase, nhv);
      const int slot = seg * nx + r - 1;
      if (MODE == kSpeed) {
        if (!SPEED_B) cl.b = 0.0;
        const double sp = ES::speed(cl, cy);
        if (col >= 2 && active) smaxl = dmax(smaxl, sp);
        if (active && col == 1) smaxl = dmax(smaxl, L.pre[0][slot]);
        if (active && col == ny) smaxl = dmax(smaxl, L.pre[1][slot]);
      } else {
        double dhL, dhR, dqL, dqR;
        if (DUMMY_FREE)
          // lane k's left cell is lane k-1's: wet where the row mask shifted
          // by one lane is set
          ES::solve_voted((((wet_row << 1) & wet_row) | col1_lanes) == __builtin_amdgcn_read_exec(),
                          cl, cy,
                          col1_lanes, dhL, dhR, dqL, dqR);
        else
          ES::solve(cl, cy, dhL, dhR, dqL, dqR);
        double up_h, up_hv;
        // the lower neighbour's share, handed down one lane
        const double ydh = from_right(dyinv * dhL), ydq = from_right(dyinv * dqL);
        if (col >= 2) {
          up_h = dyinv * dhR;
          up_hv = dyinv * dqR;
        } else {
          up_h = segok ? L.pre[0][slot] : 0.0;
          up_hv = segok ? L.pre[1][slot] : 0.0;
        }
        if (PRUNE && have_new) {
          const float2 lb = make_float2(from_left_z(pbt.ay), from_left_z(pbt.c));
          const bool cand = bound_candidate(lb.x, lb.y, pbt.ay, pbt.c, thr);
          if ((__builtin_amdgcn_ballot_w64(cand) & act_lanes & ~col1_lanes) != 0) {
            // (a wave-uniform branch: every lane takes part in the shifts)
            const double lnh = from_left(pnh), lnhv = from_left(nhv);
            const CellQ nr = cell_dir(cell_base(pnh, 0.0), nhv);
            CellQ nl = cell_dir(cell_base(lnh, 0.0), lnhv);
            const double sp = ES::speed(nl, nr);
            if (col >= 2 && active) smaxl = dmax(smaxl, sp);
          }
        }
        if (FUSED && !PRUNE && have_new) {
          CellQ nl;
          nl.h = from_left(ncy.h); nl.q = from_left(ncy.q); nl.b = 0.0;
          nl.u = from_left(ncy.u); nl.sq = from_left(ncy.sq); nl.c = from_left(ncy.c);
          nl.f = 0.0;
          const double sp = ES::speed(nl, ncy);
          if (col >= 2 && active) smaxl = dmax(smaxl, sp);
        }
        double dn_h = 0.0, dn_hv = 0.0;
        if (col < ny) {
          dn_h = ydh;
          dn_hv = ydq;
        } else if (segok) {
          dn_h = L.pre[2][slot];
          dn_hv = L.pre[3][slot];
        }
        yUpH = up_h; yUpHv = up_hv; yDnH = dn_h; yDnHv = dn_hv;
      }
    }
    px = cx;
    phv = chv;
    wet_prev = wet_row;
  }

  if (PRUNE && MODE == kUpdate) {
    // the y-edges of the last new row (row nx) are still open
    const float2 lb = make_float2(from_left_z(pbt.ay), from_left_z(pbt.c));
    const bool cand = bound_candidate(lb.x, lb.y, pbt.ay, pbt.c, thr);
    if ((__builtin_amdgcn_ballot_w64(cand) & act_lanes & ~col1_lanes) != 0) {
      const double lnh = from_left(pnh), lnhv = from_left(nhv);
      const CellQ nr = cell_dir(cell_base(pnh, 0.0), nhv);
      CellQ nl = cell_dir(cell_base(lnh, 0.0), lnhv);
      const double sp = ES::speed(nl, nr);
      if (col >= 2 && active) smaxl = dmax(smaxl, sp);
    }
  }
  if (FUSED && !PRUNE && MODE == kUpdate) {
    // the y-edges of the last new row (row nx) are still open
    const CellQ ncy = cell_dir(nbase, nhv);
    CellQ nl;
    nl.h = from_left(ncy.h); nl.q = from_left(ncy.q); nl.b = 0.0;
    nl.u = from_left(ncy.u); nl.sq = from_left(ncy.sq); nl.c = from_left(ncy.c); nl.f = 0.0;
    const double sp = ES::speed(nl, ncy);
    if (col >= 2 && active) smaxl = dmax(smaxl, sp);
  }

  if (MODE == kSpeed || FUSED) {
    // segmented max over the W lanes of each patch
    double v = active ? smaxl : 0.0;
    for (int m = W >> 1; m >= 1; m >>= 1) v = dmax(v, __shfl_xor(v, m));
    double* const out = A.smax + (long long)p * A.smax_stride;
    if (segok && col == 1) {
      gs(out + t.sidx, v);
      // a whole-patch pass clears the slots of the other row segments
      if (nrows == nx)
        for (int k = 1; k < A.smax_stride; ++k) gs(out + k, 0.0);
    }
  }
}

// Exact CFL edge speeds of the current state, per patch.  The grid strides
// over the tasks; with `only_if` set the launch does nothing unless
// *only_if != 0, and then covers only the tasks of scenarios flagged for
// the exact fallback of the pruned mode (a task holds one scenario).
template <bool SIDES, int SOLVER>
__global__ void __launch_bounds__(64)
k_speed(const Task* __restrict__ tasks, int ntasks, const int* __restrict__ plist,
        const PatchDev* __restrict__ pd, Fields F, const ScnDev* __restrict__ scn,
        const SideDev* __restrict__ sides, EdgeBufs E, double* __restrict__ smax,
        const int* __restrict__ only_if, int smax_stride) {
  __shared__ MarchLds L;
  if (only_if && gl(only_if) == 0) return;
  for (int ti = blockIdx.x; ti < ntasks; ti += gridDim.x) {
    const Task t = gl(tasks + ti);
    const ScnDev* S = scn + gl(&pd[gl(plist + t.first)].scn);
    // fallback: only the scenarios whose pruned minimum was rejected
    if (only_if && gl(&S->redo) == 0) continue;
    const int par = gl(&S->stepc) & 1;
    MarchArgs A{plist, pd, F, sides, E, par, 0.0, smax, scn, 1.0};
    A.smax_stride = smax_stride;
    march_task<kSpeed, SIDES, 0, SOLVER>(t, A, L);
  }
}

// Occupancy target of the update kernel (waves per SIMD): the register
// budget the compiler may use is 512 / this.
#ifndef PC_UPDATE_WAVES_PER_EU
#define PC_UPDATE_WAVES_PER_EU 3
#endif

// One task per workgroup, or a persistent grid pulling tasks from a
// device-side counter (reset by k_dt_final before every update launch).
template <int FUSE, bool SIDES, int SOLVER, bool O32 = false>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(PC_UPDATE_WAVES_PER_EU, 8)))
k_update(const Task* __restrict__ tasks, int ntasks, const int* __restrict__ plist,
         const PatchDev* __restrict__ pd, Fields F, const ScnDev* __restrict__ scn,
         const SideDev* __restrict__ sides, EdgeBufs E, double* __restrict__ smax_next,
         int* __restrict__ counter, int persistent, double slack) {
  __shared__ MarchLds L;
  __shared__ int next;
  int ti = blockIdx.x;
  for (;;) {
    if (persistent) {
      if (threadIdx.x == 0) next = atomicAdd(counter, 1);
      __syncthreads();
      ti = next;
      __syncthreads();
    }
    if (ti >= ntasks) return;
    const Task t = gl(tasks + ti);
    const ScnDev* S = scn + gl(&pd[gl(plist + t.first)].scn);
    if (gl(&S->advance)) {
      // k_dt_final already moved the parity counter past this step
      MarchArgs A{plist, pd, F, sides, E, (gl(&S->stepc) - 1) & 1, gl(&S->dt_applied),
                  smax_next, scn, slack};
      A.smax_stride = kMaxRowSeg;
      march_task<kUpdate, SIDES, SIDES ? FUSE : 0, SOLVER, false, O32>(t, A, L);
    }
    if (!persistent) return;
  }
}

// Ghost-edge speeds of every patch folded with the interior maximum of the
// same state (left by the fused update): smax[p] = max(smax_int[p],
// ghost-edge maximum).  One wavefront per patch (4 per workgroup) walks the
// 2 (nx + ny) ghost edges of its patch as one list (x sides: ghost row 0|1
// and nx|nx+1 along the columns; y sides: columns 0|1 and ny|ny+1 along the
// rows, the own column from the edge buffer).  PRUNE: exact speeds only
// where the bound may reach the scenario's last CFL limit, as in the update.
template <bool PRUNE>
__global__ void __launch_bounds__(256)
k_boundary_speed(const ScnDev* __restrict__ scn, const BoundaryDesc* __restrict__ bd, int nslots,
                 Fields F, EdgeBufs E, const double* __restrict__ smax_int,
                 double* __restrict__ smax, double slack) {
  constexpr int kIt = 4;  // 2 (nx + ny) <= 4 * 64 edges (nx, ny <= kERows)
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= nslots) return;
  const BoundaryDesc D = gl(bd + k);
  const int par = gl(&scn[D.scn].stepc) & 1;
  float thr = 0.0f;
  if (PRUNE) {
    const double est = gl(&scn[D.scn].est);
    if (est > 0.0) thr = (float)(kCfl * dmin(D.dx, D.dy) / (est * slack));
  }
  const int nx = D.nx, ny = D.ny, stride = D.stride;
  // edge list: x-lower (ny), x-upper (ny), y-lower (nx), y-upper (nx).  In
  // pruned mode only the scenario minimum of the CFL steps is used, so an
  // edge shared with a same-size neighbour is evaluated once, by the patch
  // below it; each patch keeps its upper sides, walls and tile ghosts.
  const int l0 = (PRUNE && D.shared[0]) ? 0 : ny;
  const int l2 = (PRUNE && D.shared[1]) ? 0 : nx;
  const int b1 = l0, b2 = l0 + ny, b3 = b2 + l2;
  const int ne = b3 + nx;
  const double* H = F.H(par);
  const double* HU = F.HU(par);
  const double* HV = F.HV(par);
  const double* Ecur = par ? E.e[1] : E.e[0];
  // all loads first: own boundary cell (ho, qo) and the cell across the
  // edge (hg, qg) of every edge this lane owns
  double ho[kIt], qo[kIt], hg[kIt], qg[kIt];
  bool up[kIt];
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    const int e = it * 64 + lane;
    ho[it] = 1.0; qo[it] = 0.0; hg[it] = 1.0; qg[it] = 0.0; up[it] = false;
    if (e < b2) {
      // x side: edge (row 0 | 1) or (nx | nx+1) at column c
      const bool upper = e >= b1;
      const int c = (upper ? e - b1 : e) + 1;
      const long long io = D.off + (long long)(upper ? nx : 1) * stride + c;
      const long long ig = (upper ? D.xrow[1] : D.xrow[0]) + c;
      ho[it] = gl(H + io);
      qo[it] = gl(HU + io);
      hg[it] = gl(H + ig);
      qg[it] = gl(HU + ig);
      up[it] = upper;
    } else if (e < ne) {
      // y side: edge (col 0 | 1) or (ny | ny+1) at row r
      const bool upper = e >= b3;
      const int r = (upper ? e - b3 : e - b2) + 1;
      const long long eo = (upper ? D.ebo[1] : D.ebo[0]) + r - 1;
      const long long eg = upper ? D.ebg[1] : D.ebg[0];
      ho[it] = gl(Ecur + eo);
      qo[it] = gl(Ecur + eo + kERows);
      if (eg >= 0) {
        hg[it] = gl(Ecur + eg + r - 1);
        qg[it] = gl(Ecur + eg + r - 1 + kERows);
      } else if (eg == -2) {
        const long long j = D.off + (long long)r * stride + (upper ? ny + 1 : 0);
        hg[it] = gl(H + j);
        qg[it] = gl(HV + j);
      }
      up[it] = upper;
    }
  }
  double sm = 0.0;
#pragma unroll
  for (int it = 0; it < kIt; ++it) {
    if (it * 64 >= ne) break;  // wave-uniform
    const int e = it * 64 + lane;
    const bool valid = e < ne;
    const bool upper = up[it];
    // walls mirror the own boundary cell with the normal momentum negated
    bool wall;
    if (e < b2)
      wall = upper ? D.xwall[1] != 0 : D.xwall[0] != 0;
    else
      wall = (upper ? D.ebg[1] : D.ebg[0]) == -1;
    double hgv = hg[it], qgv = qg[it];
    if (e < b2) {
      if (wall) qgv = -qgv;  // the own boundary row, read as the ghost source
    } else if (wall) {
      hgv = ho[it];
      qgv = -qo[it];
    }
    // orientation: the lower side's ghost is the left cell of its edge
    const double hL = upper ? ho[it] : hgv, qL = upper ? qo[it] : qgv;
    const double hR = upper ? hgv : ho[it], qR = upper ? qgv : qo[it];
    bool need = valid;
    if (PRUNE) {
      const BoundTerms bl = bound_terms(hL, qL, 0.0), br = bound_terms(hR, qR, 0.0);
      need = valid && bound_candidate(bl.ax, bl.c, br.ax, br.c, thr);
    }
    if (wave_any(need)) {
      const CellQ cl = cell_dir(cell_base(hL, 0.0), qL);
      const CellQ cr = cell_dir(cell_base(hR, 0.0), qR);
      const double sp = edge_speed_any(cl, cr);
      if (valid) sm = dmax(sm, sp);
    }
  }
  // seams of the update's row segments: the x-edges (a-1 | a) of the new
  // state at the first row a of every segment after the first (each segment
  // of the update saw only one side of them)
  if (D.rseg > 1) {
    const int nse = (D.rseg - 1) * ny;
    for (int e = lane; e - lane < nse; e += 64) {  // wave-uniform trip count
      const bool valid = e < nse;
      double hL = 1.0, qL = 0.0, hR = 1.0, qR = 0.0;
      if (valid) {
        const int k = e / ny + 1;
        const int c = e - (k - 1) * ny + 1;
        const long long i = D.off + (long long)seg_first_row(k, nx, D.rseg) * stride + c;
        hL = gl(H + i - stride);
        qL = gl(HU + i - stride);
        hR = gl(H + i);
        qR = gl(HU + i);
      }
      bool need = valid;
      if (PRUNE) {
        const BoundTerms bl = bound_terms(hL, qL, 0.0), br = bound_terms(hR, qR, 0.0);
        need = valid && bound_candidate(bl.ax, bl.c, br.ax, br.c, thr);
      }
      if (wave_any(need)) {
        const CellQ cl = cell_dir(cell_base(hL, 0.0), qL);
        const CellQ cr = cell_dir(cell_base(hR, 0.0), qR);
        const double sp = edge_speed_any(cl, cr);
        if (valid) sm = dmax(sm, sp);
      }
    }
  }
  for (int o = 32; o >= 1; o >>= 1) sm = dmax(sm, __shfl_xor(sm, o));
  if (lane == 0) {
    const double* si = smax_int + (long long)D.slot * kMaxRowSeg;
    double m = gl(si);
    for (int k = 1; k < D.rseg; ++k) m = dmax(m, gl(si + k));
    smax[D.slot] = dmax(m, sm);
  }
}

// Memory-pattern probe of the row march: the update kernel's traversal
// (same tasks and launch shape, one wavefront per task, lanes on columns,
// rows marched in order) reading h, hu, hv, b of every row of the tiles and
// writing h, hu, hv of the interior rows to scratch arrays, on packed rows
// the edge lanes also into the row's ghost columns as the update does,
// with no edge arithmetic.  Without arithmetic to hide behind, the loads run
// four rows ahead (rows in blocks of four).  The rate it reaches is the
// memory rate of this traversal of the layout.
__global__ void __launch_bounds__(64)
k_march_probe(const Task* __restrict__ tasks, int ntasks, const int* __restrict__ plist,
              const PatchDev* __restrict__ pd, const double* __restrict__ h,
              const double* __restrict__ hu, const double* __restrict__ hv,
              const double* __restrict__ b, double* __restrict__ oh, double* __restrict__ ohu,
              double* __restrict__ ohv, int* __restrict__ counter, int persistent) {
  constexpr int K = 4;
  __shared__ int next;
  const int lane = threadIdx.x & 63;
  int ti = blockIdx.x;
  for (;;) {
    if (persistent) {
      if (threadIdx.x == 0) next = atomicAdd(counter, 1);
      __syncthreads();
      ti = next;
      __syncthreads();
    }
    if (ti >= ntasks) return;
    const Task t = gl(tasks + ti);
    const int W = pow2ceil(t.ny);
    const int seg = lane / W;
    const int col = lane - seg * W + 1;
    const bool segok = seg < t.nseg;
    const bool active = segok && col <= t.ny;
    const PatchDev P = gl(pd + gl(plist + t.first + (segok ? seg : 0)));
    const long long stride = P.stride;
    const long long base = P.off + (active ? col : 1);
    const int last = t.nx + 1;  // rows 0 .. nx+1
    double ch[K], chu[K], chv[K], cb[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const long long i = base + (long long)(k <= last ? k : last) * stride;
      ch[k] = gl(h + i); chu[k] = gl(hu + i); chv[k] = gl(hv + i); cb[k] = gl(b + i);
    }
    for (int r0 = 0; r0 <= last; r0 += K) {
      double nh[K], nhu[K], nhv[K], nbv[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int rr = r0 + K + k;
        const long long i = base + (long long)(rr <= last ? rr : last) * stride;
        nh[k] = gl(h + i); nhu[k] = gl(hu + i); nhv[k] = gl(hv + i); nbv[k] = gl(b + i);
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int r = r0 + k;
        if (active && r >= 1 && r <= t.nx) {
          const long long o = base + (long long)r * stride;
          const double x = ch[k] + 0.5 * cb[k];
          gs(oh + o, x);
          gs(ohu + o, chu[k] * x);
          gs(ohv + o, chv[k] * x);
          if (stride == t.ny + 2 && (col == 1 || col == t.ny)) {
            const long long og = o + (col == 1 ? -1 : 1);
            gs(oh + og, x);
            gs(ohu + og, chu[k] * x);
            gs(ohv + og, chv[k] * x);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        ch[k] = nh[k]; chu[k] = nhu[k]; chv[k] = nhv[k]; cb[k] = nbv[k];
      }
    }
    if (!persistent) return;
  }
}

// Edge buffers of the current state from the tiles (after uploads or
// per-patch calls that changed the interior).
__global__ void k_edges_init(const int* __restrict__ slots, int nslots,
                             const PatchDev* __restrict__ pd, Fields F,
                             const ScnDev* __restrict__ scn, EdgeBufs E, double* bcol) {
  const int k = blockIdx.x;
  if (k >= nslots) return;
  const int p = slots[k];
  const PatchDev P = pd[p];
  const int par = scn[P.scn].stepc & 1;
  const int stride = P.stride;
  for (int r = threadIdx.x + 1; r <= P.nx; r += blockDim.x) {
    const long long row = P.off + (long long)r * stride;
    for (int side = 0; side < 2; ++side) {
      const long long i = row + (side ? P.ny : 1);
      double* Ed = (par ? E.e[1] : E.e[0]) + (long long)((p * 2 + side) * 2) * kERows + r - 1;
      Ed[0] = F.H(par)[i];
      Ed[kERows] = F.HV(par)[i];
    }
    double* Bc = bcol + (long long)p * 4 * kERows + r - 1;
    Bc[0] = F.b[row];
    Bc[kERows] = F.b[row + 1];
    Bc[2 * kERows] = F.b[row + P.ny];
    Bc[3 * kERows] = F.b[row + P.ny + 1];
  }
}

// Per-patch CFL steps and their minimum per scenario (first stage): the
// scenario patch list is cut into chunks; a chunk inside one scenario is
// reduced in LDS and folded with one atomic min on the bit pattern (all
// values are positive), otherwise each entry folds its own value.
// `redo_clear` (pruned mode) is reset here, ahead of the check in
// k_dt_final; with `only_if` the launch does nothing unless *only_if != 0.
__global__ void __launch_bounds__(256)
k_dt_partial(ScnDev* __restrict__ scn, const int* __restrict__ plist_scn,
             const int* __restrict__ scn_of_entry, int nentries,
             const PatchDev* __restrict__ pd, const double* __restrict__ smax,
             double* __restrict__ est_patch, int* __restrict__ redo_clear,
             const int* __restrict__ only_if) {
  __shared__ double red[256];
  if (only_if && gl(only_if) == 0) return;
  if (redo_clear && blockIdx.x == 0 && threadIdx.x == 0) *redo_clear = 0;
  const int beg = blockIdx.x * 1024;
  const int end = min(beg + 1024, nentries);
  if (beg >= end) return;
  const int s0 = scn_of_entry[beg], s1 = scn_of_entry[end - 1];
  double m = 1.7976931348623157e308;
  for (int k = beg + threadIdx.x; k < end; k += 256) {
    const int p = plist_scn[k];
    const double v = max_timestep(smax[p], pd[p].dx, pd[p].dy);
    est_patch[p] = v;
    if (s0 == s1)
      m = dmin(m, v);
    else
      atomicMin(&scn[scn_of_entry[k]].minbits, (unsigned long long)__double_as_longlong(v));
  }
  if (s0 != s1) return;
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = dmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    atomicMin(&scn[s0].minbits, (unsigned long long)__double_as_longlong(red[0]));
}

// Second stage, one thread per scenario: dt = min(CFL, end - (t + dt_prev)),
// the time bookkeeping of SWEKernel.cpp:83-87 and the parity flip.
//   mode 0: commit every scenario with its minimum;
//   mode 1 (pruned speeds): commit a scenario that advances only if its
//           minimum is proven exact (<= est * slack * kPruneCheckFrac), else flag it
//           and set *redo for the exact fallback pass;
//   mode 2 (after the fallback pass; nothing unless *redo): commit the
//           flagged scenarios, reset the running minimum of all.
// The decision for one scenario given its CFL minimum (the running minimum
// has been read and reset by the caller).
__device__ __forceinline__ void dt_final_one(ScnDev& S, double cfl, int mode, int* redo,
                                             double slack) {
  const double maxDt = S.end_time - (S.time + S.dt);
  const bool can_advance = S.pend > S.pbeg && maxDt - 0.0 > kLaEps;
  if (mode == 1 && can_advance && !(cfl <= S.est * (slack * kPruneCheckFrac))) {
    S.redo = 1;
    *redo = 1;
    return;
  }
  if (mode == 2) {
    if (!S.redo) return;
    S.redo = 0;
  }
  int adv = 0;
  double dt = 0.0;
  if (can_advance) {
    dt = fmin(cfl, maxDt);
    adv = 1;
    if (dt - 0.0 > kLaEps) {
      S.time = S.time + S.dt;
      S.dt = dt;
    }
    S.stepc = S.stepc + 1;
    S.steps = S.steps + 1;
  }
  S.est = cfl;
  S.dt_applied = dt;
  S.advance = adv;
}

__global__ void k_dt_final(ScnDev* __restrict__ scn, int nscn, int* __restrict__ counter,
                           int mode, int* __restrict__ redo, double slack) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (mode == 2 && gl(redo) == 0) return;
  if (s == 0 && counter) *counter = 0;
  if (s >= nscn) return;
  ScnDev& S = scn[s];
  const double cfl = __longlong_as_double((long long)S.minbits);
  S.minbits = (unsigned long long)__double_as_longlong(1.7976931348623157e308);
  dt_final_one(S, cfl, mode, redo, slack);
}

// Both stages in one workgroup, for steps with at most kDtBlockEntries
// patch entries: the per-patch CFL steps and the per-scenario minima
// (wavefront minima where a wavefront's entries share a scenario, then
// atomicMin), a workgroup barrier, then the decision per scenario with the
// minimum read and reset by one atomic exchange.  The same arguments and
// no-op conditions as k_dt_partial followed by k_dt_final.
constexpr int kDtBlockEntries = 4096;
__global__ void __launch_bounds__(1024)
k_dt_block(ScnDev* __restrict__ scn, int nscn, const int* __restrict__ plist_scn,
           const int* __restrict__ scn_of_entry, int nentries, const PatchDev* __restrict__ pd,
           const double* __restrict__ smax, double* __restrict__ est_patch,
           int* __restrict__ counter, int mode, int* __restrict__ redo, double slack) {
  if (mode == 2 && gl(redo) == 0) return;
  if (mode == 1 && threadIdx.x == 0) *redo = 0;
  const int lane = threadIdx.x & 63;
  for (int k0 = threadIdx.x - lane; k0 < nentries; k0 += blockDim.x) {
    const int k = k0 + lane;
    const bool ok = k < nentries;
    double v = 1.7976931348623157e308;
    int sc = -1;
    if (ok) {
      const int p = plist_scn[k];
      v = max_timestep(smax[p], pd[p].dx, pd[p].dy);
      est_patch[p] = v;
      sc = scn_of_entry[k];
    }
    const int s0 = __shfl(sc, 0);
    const int last = min(63, nentries - 1 - k0);
    if (__shfl(sc, last) == s0) {
      // entries of one scenario (they are grouped by scenario)
      for (int o = 32; o >= 1; o >>= 1) v = dmin(v, __shfl_xor(v, o));
      if (lane == 0) atomicMin(&scn[s0].minbits, (unsigned long long)__double_as_longlong(v));
    } else if (ok) {
      atomicMin(&scn[sc].minbits, (unsigned long long)__double_as_longlong(v));
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && counter) *counter = 0;
  for (int s = threadIdx.x; s < nscn; s += blockDim.x) {
    ScnDev& S = scn[s];
    const unsigned long long bits = atomicExch(
        &S.minbits, (unsigned long long)__double_as_longlong(1.7976931348623157e308));
    dt_final_one(S, __longlong_as_double((long long)bits), mode, redo, slack);
  }
}

// Ghost fill of the current state's tile ghost cells: one wavefront per strip.
__global__ void __launch_bounds__(256)
k_gather(const GhostStrip* __restrict__ strips, int nstrips,
         const PatchDev* __restrict__ pd, Fields F, const ScnDev* __restrict__ scn) {
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int k = threadIdx.x & 63;
  if (s >= nstrips) return;
  const GhostStrip g = strips[s];
  const PatchDev D = pd[g.dst_slot];
  const PatchDev S = pd[g.src_slot];
  const int par = scn[D.scn].stepc & 1;
  for (int c = k; c < g.length; c += 64) {
    const long long di = D.off + g.dst_cell + (long long)c * g.dst_stride;
    const long long si = S.off + g.src_cell + (long long)c * g.src_stride;
    const double h = F.H(par)[si];
    const double hu = F.HU(par)[si];
    const double hv = F.HV(par)[si];
    F.H(par)[di] = h;
    F.HU(par)[di] = (g.negate & 1) ? -hu : hu;
    F.HV(par)[di] = (g.negate & 2) ? -hv : hv;
  }
}

}  // namespace pcamd
