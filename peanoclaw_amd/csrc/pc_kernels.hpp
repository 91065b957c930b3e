// HIP kernels of the batched SWE patch step on gfx950.
//
// Patch pool layout in HBM (structure of arrays): every field (h, hu, hv in
// two states for ping-pong, b once) is one device array; patch p owns the
// contiguous tile [off_p, off_p + (nx+2)(ny+2)) in each of them, stored in
// the reference's order (x rows, y contiguous; grid/LinearizationZYXQ.h).
//
// The update is a row march: one 64-lane wavefront owns a task of up to
// 64/W patches of the same size (W = ny rounded up to a power of two); lane
// (seg, col) walks down column `col` of patch `seg`, keeps the previous row
// in registers, solves the x-edge between the two rows and the y-edge to its
// left neighbour lane (exchanged through LDS), and finalises the previous
// row once both of its x-edges are known.  Each cell is read from HBM once
// and written once.  The two y-edges per row that touch the ghost columns
// are solved up front by a pre-pass in which the lanes take rows.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "swe_device.hpp"

namespace pcamd {

struct PatchDev {
  long long off;  // tile base in the field arrays
  int nx, ny;
  int scn;        // scenario of the patch
  int pad;
  double dx, dy;  // subcell size (Patch.cpp:38)
  double dxinv, dyinv;
};

struct Fields {
  double* h[2];
  double* hu[2];
  double* hv[2];
  double* b;
};

struct Task {
  int first;  // index into the task patch list
  int nseg;   // patches in this task (same nx, ny)
  int nx, ny;
};

struct ScnDev {
  double time;        // TimeIntervals "time" (start of the last interval)
  double dt;          // TimeIntervals timestep size
  double est;         // estimatedNextTimestepSize (min over patches)
  double end_time;    // global timestep end time
  double dt_applied;  // dt of the current step, 0 if it did not advance
  long long steps;
  int stepc;          // parity counter: state (stepc & 1) is current
  int advance;        // current step advances this scenario
  int pbeg, pend;     // its patches in the scenario patch list
};

struct GhostStrip {
  int dst_slot, dst_cell, dst_stride;
  int src_slot, src_cell, src_stride;
  int length, negate;
};

enum MarchMode { kSpeed = 0, kUpdate = 1, kUpdateSwap = 2 };

__device__ __forceinline__ int pow2ceil(int v) {
  int w = 1;
  while (w < v) w <<= 1;
  return w;
}

// LDS scratch of one wavefront (64 lanes).
struct MarchLds {
  double yc[7][64];      // y-sweep cell terms of the current row
  double ydow[2][64];    // y-edge results handed to the lower neighbour
  double pre[4][128];    // boundary y-edges: [side*2 + k][seg*nx + row-1]
};

// One task.  kSpeed: per-patch maximum edge speed of state `par` written
// to smax[p].  kUpdate: new interior into state 1-par with step dt.
// kUpdateSwap: the reference's in-place update of uOld followed by the
// swap (SWEKernel.cpp:57-73): source = uOld tile (state 1-cur), new values
// to uNew (state cur), the old uNew into uOld's interior.
template <int MODE>
__device__ void march_task(const Task& t, const int* __restrict__ plist,
                           const PatchDev* __restrict__ pd, const Fields& F,
                           int par, double dt, double* __restrict__ smax,
                           MarchLds& L) {
  const int lane = threadIdx.x & 63;
  const int nx = t.nx, ny = t.ny;
  const int W = pow2ceil(ny);
  const int seg = lane / W;
  const int col = lane - seg * W + 1;  // 1-based interior column
  const bool active = (seg < t.nseg) && (col <= ny);
  const int p = plist[t.first + (seg < t.nseg ? seg : 0)];
  const PatchDev P = pd[p];
  const int stride = ny + 2;
  const int src = (MODE == kUpdateSwap) ? 1 - par : par;
  // kUpdateSwap writes into the tile it reads (rows already consumed), so
  // its source pointers must not be declared non-aliasing.
  using SrcPtr = typename std::conditional<MODE == kUpdateSwap, const double*,
                                           const double* __restrict__>::type;
  SrcPtr Hs = F.h[src];
  SrcPtr HUs = F.hu[src];
  SrcPtr HVs = F.hv[src];
  const double* __restrict__ Bs = F.b;

  // ---- pre-pass: the y-edges (row, 0|1) and (row, ny|ny+1) ----
  const int npre = t.nseg * nx * 2;
  for (int e = lane; e < npre; e += 64) {
    const int s = e / (2 * nx);
    const int rem = e - s * 2 * nx;
    const int row = (rem >> 1) + 1;
    const int side = rem & 1;
    const PatchDev Q = pd[plist[t.first + s]];
    const long long iL = Q.off + (long long)row * stride + (side ? ny : 0);
    const long long iR = iL + 1;
    const CellBase bl = cell_base(Hs[iL], MODE == kSpeed ? 0.0 : Bs[iL]);
    const CellBase br = cell_base(Hs[iR], MODE == kSpeed ? 0.0 : Bs[iR]);
    const CellQ cl = cell_dir(bl, HVs[iL]);
    const CellQ cr = cell_dir(br, HVs[iR]);
    const int slot = s * nx + row - 1;
    if (MODE == kSpeed) {
      L.pre[side][slot] = edge_speed(cl, cr);
    } else {
      double dhL, dhR, dqL, dqR;
      edge_solve(cl, cr, dhL, dhR, dqL, dqR);
      // side 0: the right cell (row,1) takes the "upwards" part;
      // side 1: the left cell (row,ny) takes the "downwards" part.
      L.pre[side * 2 + 0][slot] = Q.dyinv * (side ? dhL : dhR);
      L.pre[side * 2 + 1][slot] = Q.dyinv * (side ? dqL : dqR);
    }
  }
  __syncthreads();

  const long long base = P.off + col;
  const double dxinv = P.dxinv, dyinv = P.dyinv;
  double smaxl = 0.0;

  // previous row (x-sweep terms + raw values)
  CellQ px = {};
  double phv = 0.0;
  // accumulators of the previous row
  double accH = 0.0, accHu = 0.0, yUpH = 0.0, yUpHv = 0.0, yDnH = 0.0, yDnHv = 0.0;

  for (int r = 0; r <= nx + 1; ++r) {
    const long long i = base + (long long)r * stride;
    double h = 1.0, hu = 0.0, hv = 0.0, b = 0.0;
    if (active) {
      h = Hs[i];
      hu = HUs[i];
      hv = HVs[i];
      if (MODE != kSpeed) b = Bs[i];
    }
    const CellBase cb = cell_base(h, b);
    const CellQ cx = cell_dir(cb, hu);

    if (r >= 1) {
      // x-edge (r-1 | r)
      if (MODE == kSpeed) {
        smaxl = dmax(smaxl, edge_speed(px, cx));
      } else {
        double dhL, dhR, dqL, dqR;
        edge_solve(px, cx, dhL, dhR, dqL, dqR);
        if (r >= 2) {
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
            const long long o = i - stride;
            const int dst = (MODE == kUpdateSwap) ? par : 1 - par;
            if (MODE == kUpdateSwap) {
              // uOld interior <- old uNew, uNew <- new values
              F.h[src][o] = F.h[dst][o];
              F.hu[src][o] = F.hu[dst][o];
              F.hv[src][o] = F.hv[dst][o];
            }
            F.h[dst][o] = h1;
            F.hu[dst][o] = hu1;
            F.hv[dst][o] = hv1;
          }
        }
        accH = 0.0 + dxinv * dhR;
        accHu = 0.0 + dxinv * dqR;
      }
    }

    if (r >= 1 && r <= nx) {
      // y-edges of row r: lane col solves (col-1 | col) for col >= 2
      const CellQ cy = cell_dir(cb, hv);
      L.yc[0][lane] = cy.h;
      L.yc[1][lane] = cy.q;
      L.yc[2][lane] = cy.b;
      L.yc[3][lane] = cy.u;
      L.yc[4][lane] = cy.sq;
      L.yc[5][lane] = cy.c;
      L.yc[6][lane] = cy.f;
      __syncthreads();
      const int slot = seg * nx + r - 1;
      if (MODE == kSpeed) {
        if (col >= 2 && active) {
          CellQ cl;
          cl.h = L.yc[0][lane - 1]; cl.q = L.yc[1][lane - 1]; cl.b = 0.0;
          cl.u = L.yc[3][lane - 1]; cl.sq = L.yc[4][lane - 1];
          cl.c = L.yc[5][lane - 1]; cl.f = L.yc[6][lane - 1];
          smaxl = dmax(smaxl, edge_speed(cl, cy));
        }
        if (active && col == 1) smaxl = dmax(smaxl, L.pre[0][slot]);
        if (active && col == ny) smaxl = dmax(smaxl, L.pre[1][slot]);
        __syncthreads();
      } else {
        double up_h = 0.0, up_hv = 0.0;
        if (col >= 2) {
          CellQ cl;
          cl.h = L.yc[0][lane - 1]; cl.q = L.yc[1][lane - 1]; cl.b = L.yc[2][lane - 1];
          cl.u = L.yc[3][lane - 1]; cl.sq = L.yc[4][lane - 1];
          cl.c = L.yc[5][lane - 1]; cl.f = L.yc[6][lane - 1];
          double dhL, dhR, dqL, dqR;
          edge_solve(cl, cy, dhL, dhR, dqL, dqR);
          up_h = dyinv * dhR;
          up_hv = dyinv * dqR;
          if (active) {
            L.ydow[0][lane - 1] = dyinv * dhL;
            L.ydow[1][lane - 1] = dyinv * dqL;
          }
        } else if (seg < t.nseg) {
          up_h = L.pre[0][slot];
          up_hv = L.pre[1][slot];
        }
        __syncthreads();
        double dn_h = 0.0, dn_hv = 0.0;
        if (col < ny) {
          dn_h = L.ydow[0][lane];
          dn_hv = L.ydow[1][lane];
        } else if (seg < t.nseg) {
          dn_h = L.pre[2][slot];
          dn_hv = L.pre[3][slot];
        }
        yUpH = up_h; yUpHv = up_hv; yDnH = dn_h; yDnHv = dn_hv;
        __syncthreads();
      }
    }
    px = cx;
    phv = hv;
  }

  if (MODE == kSpeed) {
    // segmented max over the W lanes of each patch
    double v = active ? smaxl : 0.0;
    for (int m = W >> 1; m >= 1; m >>= 1) v = dmax(v, __shfl_xor(v, m));
    if (seg < t.nseg && col == 1) smax[p] = v;
  }
}

__global__ void __launch_bounds__(64)
k_speed(const Task* __restrict__ tasks, int ntasks, const int* __restrict__ plist,
        const PatchDev* __restrict__ pd, Fields F, const ScnDev* __restrict__ scn,
        double* __restrict__ smax) {
  __shared__ MarchLds L;
  const int ti = blockIdx.x;
  if (ti >= ntasks) return;
  const Task t = tasks[ti];
  const int par = scn[pd[plist[t.first]].scn].stepc & 1;
  march_task<kSpeed>(t, plist, pd, F, par, 0.0, smax, L);
}

// One task per workgroup, or a persistent grid pulling tasks from a
// device-side counter (reset by k_dt before every update launch).
__global__ void __launch_bounds__(64)
k_update(const Task* __restrict__ tasks, int ntasks, const int* __restrict__ plist,
         const PatchDev* __restrict__ pd, Fields F, const ScnDev* __restrict__ scn,
         int* __restrict__ counter, int persistent) {
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
    const Task t = tasks[ti];
    const ScnDev S = scn[pd[plist[t.first]].scn];
    if (S.advance) {
      // k_dt already moved the parity counter past this step
      const int par = (S.stepc - 1) & 1;
      march_task<kUpdate>(t, plist, pd, F, par, S.dt_applied, nullptr, L);
    }
    if (!persistent) return;
  }
}

// Per scenario: dt = min(min_p CFL(smax_p), end - (t + dt_prev)), the time
// bookkeeping of SWEKernel.cpp:83-87 and the parity flip.
__global__ void __launch_bounds__(256)
k_dt(ScnDev* __restrict__ scn, const int* __restrict__ plist_scn,
     const PatchDev* __restrict__ pd, const double* __restrict__ smax,
     double* __restrict__ est_patch, int* __restrict__ counter) {
  __shared__ double red[256];
  ScnDev& S = scn[blockIdx.x];
  double m = 1.7976931348623157e308;
  for (int k = S.pbeg + threadIdx.x; k < S.pend; k += blockDim.x) {
    const int p = plist_scn[k];
    const double v = max_timestep(smax[p], pd[p].dx, pd[p].dy);
    est_patch[p] = v;
    m = dmin(m, v);
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = dmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (blockIdx.x == 0 && counter) *counter = 0;
    const double cfl = red[0];
    const double maxDt = S.end_time - (S.time + S.dt);
    int adv = 0;
    double dt = 0.0;
    if (S.pend > S.pbeg && maxDt - 0.0 > kLaEps) {
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
}

// Ghost fill of the current state: one wavefront per strip.
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
    const double h = F.h[par][si];
    const double hu = F.hu[par][si];
    const double hv = F.hv[par][si];
    F.h[par][di] = h;
    F.hu[par][di] = (g.negate & 1) ? -hu : hu;
    F.hv[par][di] = (g.negate & 2) ? -hv : hv;
  }
}

}  // namespace pcamd
