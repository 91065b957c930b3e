// Batched step of an adaptive patch grid (included at the end of
// pc_abi.hip; uses the pool internals).
//
// The per-call form of a global AMR step is a long sequence of Numerics
// calls: restriction of every leaf into each virtual patch containing it
// (GridLevelTransfer::restrictToOverlappingVirtualSubgrids,
// GridLevelTransfer.cpp:140-198, with SWEKernel::restrictSolution and
// postProcessRestriction, SWEKernel.cpp:288-316), copyUNewToUOld of every
// leaf, a ghost-layer fill of every leaf face by transferGhostlayer from
// the same-level patch or interpolateSolution from the coarser leaf
// (GhostlayerCompositorFunctors.cpp:59-69; SWEKernel.cpp:239-286), walls
// (SolveTimestep.cpp:87-142), the global CFL minimum and solveTimestep of
// every leaf.  Here the host hands over that plan once (a device-side
// table rebuilt only when the grid changes) and every step is seven
// launches, each phase over all patches at once, replayed from a captured
// graph:
//
//   restrict one thread per virtual-patch cell: the contributions of the
//            leaves covering it (a per-cell list, in plan order), then
//            h+b -> h
//   copy     uNew -> uOld of every leaf
//   ghosts   one wavefront per face: copy with time factor, or d-linear
//            interpolation from the coarser leaf (h+b, back to h with the
//            destination's b)
//   walls    one workgroup per leaf with domain walls, in call order
//   speed    exact edge speeds of every leaf (row march over uOld)
//   dt       global CFL minimum, bounded by the end time
//   solve    update of every leaf with that step (row march, uOld -> uNew),
//            then the leaves' clock
//
// Every phase first checks the step bound end_time - (time + dt); a step
// past the end time does nothing.
//
// The level transfers read h + b instead of transforming the source in
// place and back: the reference's round trip (h + b) - b can move a depth
// by one ulp, which the per-call entry points reproduce and this form does
// not (results agree to ~1e-14 relative; both are checked against the
// oracle's matching drivers).  Every other operation is the per-call
// arithmetic in the same order.
#pragma once

#include <unordered_map>

namespace {

struct AmrClock {
  double time, dt;  // the leaves' clock (time, timestep_size)
  double end_time;
  double bound;     // dt of the step being run
  int advance;      // the step runs (set by the dt phase for the solve)
  int pad;
  long long steps;
};

// end_time - (time + dt) > tolerance (AMR driver's step condition)
__device__ __forceinline__ bool amr_advance(const AmrClock& c) {
  const double maxdt = c.end_time - (c.time + c.dt);
  return maxdt - 0.0 > kLaEps;
}

// Local time stepping: per-slot clocks of the leaves (indexed by slot id).
struct AmrLts {
  double* t;        // time (start of the current interval)
  double* dt;       // timestep size
  int* adv;         // the leaf advances in this iteration
  double* stepdt;   // the step an advancing leaf takes
  int* iter;        // leaves advancing in this iteration
  unsigned long long* steps;  // leaf steps taken
  int* itn;         // iteration number
  int* need;        // virtual patch: iteration in which an advancing leaf reads it
};

// Gate of a phase: global steps run while the shared clock may advance;
// local time stepping while some leaf advances in this iteration.
template <bool LTS>
__device__ __forceinline__ bool amr_phase_runs(const AmrClock* clock, const AmrLts& L) {
  if (LTS) return *L.iter > 0;
  return amr_advance(*clock);
}

// 64 consecutive cells of one virtual patch
// the restriction kernel's lanes per virtual-patch cell, and cells per
// wavefront (one AmrVBlock)
constexpr int kRestrictLanes = 8;
constexpr int kRestrictCells = 64 / kRestrictLanes;

struct AmrVBlock {
  int slot;   // the virtual patch
  int cell0;  // first cell (x * ny + y)
  int base;   // index of cell 0 of this patch in the per-cell lists
  int pad;
};

// Restriction of the whole leaf `src` into one virtual patch
// (DefaultRestriction.cpph:16-106 over region (0, n)).
struct AmrRestrict {
  int src;
  int rSize[2];
  int dOff[2], dSize[2];
  int pad;
  double srcSub[2], dstSub[2], dstRegion;
  double spos[2], dpos[2];
};

enum AmrGhostKind { kAmrTransferLeaf = 0, kAmrTransferVirtual = 1, kAmrInterpolate = 2 };

struct AmrGhost {
  int kind, src, dst;
  int sx, sy, sox, soy, dox, doy;
  int sn[2];
  int pad;
  double spos[2], ssub[2], sinv[2], dpos[2], dsub[2];
};

struct AmrWalls {
  int slot, first, count, pad;  // walls[first .. first+count) as (dim, upper) pairs
};

// Restriction time factors of a leaf source on the [0, 1] basis
// (DefaultRestriction.cpph:46-52).
__device__ __forceinline__ void amr_restrict_factors(const AmrClock& c, double& tfOld,
                                                     double& tfNew) {
  tfOld = 1.0;
  tfNew = 1.0;
  if (c.dt - 0.0 > kLaEps) {
    const double sOld = c.time, sNew = c.time + c.dt;
    tfOld = (0.0 - sOld) / (sNew - sOld);
    tfNew = (1.0 - sOld) / (sNew - sOld);
  }
}

// One virtual-patch cell per group of kRestrictLanes lanes (kRestrictCells
// cells per wavefront): the contributions of its fine cells are computed
// kRestrictLanes at a time in parallel and summed by the group's first lane
// in the order of the per-call sequence (pairs in plan order, fine cells x
// outer, y inner), so the sums equal the sequential ones bit for bit.
// `cptr`/`cpairs` list, per cell, the restriction pairs whose destination
// region holds it, in plan order.
template <bool LTS>
__global__ void __launch_bounds__(64)
k_amr_restrict(const AmrVBlock* __restrict__ blocks, const int* __restrict__ cptr,
               const int* __restrict__ cpairs, const AmrRestrict* __restrict__ pairs,
               const PatchDev* __restrict__ pd, Fields F, int cur, const AmrClock* clock,
               AmrLts L) {
  constexpr int K = kRestrictLanes;
  __shared__ double con[kRestrictCells][6][K];
  const AmrClock c = *clock;
  if (!amr_phase_runs<LTS>(clock, L)) return;
  const AmrVBlock B = blocks[blockIdx.x];
  // local time stepping: only the virtual patches an advancing leaf reads
  if (LTS && L.need[B.slot] != *L.itn) return;
  const PatchDev D = pd[B.slot];
  const int g = (int)threadIdx.x / K, sub = (int)threadIdx.x - g * K;
  const int k = B.cell0 + g;
  const bool cellok = k < D.nx * D.ny;
  double tfOld, tfNew;
  amr_restrict_factors(c, tfOld, tfNew);
  const double eps = 1e-12;
  const int oldSt = 1 - cur;
  const int di0 = cellok ? k / D.ny : 0, di1 = cellok ? k - di0 * D.ny : 0;
  double accO[3] = {0.0, 0.0, 0.0};
  double accN[3] = {0.0, 0.0, 0.0};
  const int jb = cellok ? cptr[B.base + k] : 0, je = cellok ? cptr[B.base + k + 1] : 0;
  for (int j = jb; j < je; ++j) {  // uniform within the group
    const AmrRestrict a = pairs[cpairs[j]];
    const PatchDev S = pd[a.src];
    if (LTS) {
      // each leaf restricts on its own interval
      AmrClock cs = c;
      cs.time = L.t[a.src];
      cs.dt = L.dt[a.src];
      amr_restrict_factors(cs, tfOld, tfNew);
    }
    const double dp[2] = {a.dpos[0] + (double)di0 * a.dstSub[0],
                          a.dpos[1] + (double)di1 * a.dstSub[1]};
    int cOff[2], cSize[2];
    for (int d = 0; d < 2; ++d) {
      double ot = (dp[d] - a.spos[d] + eps);
      ot /= a.srcSub[d];
      const int o = (int)ot;
      double st = dp[d] + a.dstSub[d] - a.spos[d] - eps;
      st /= a.srcSub[d];
      st -= ot;
      st += 1.0;
      const int s = (int)st;
      const int upper = o + s, rUpper = a.rSize[d];
      cOff[d] = o > 0 ? o : 0;
      cSize[d] = (upper < rUpper ? upper : rUpper) - cOff[d];
    }
    const int nfine = cSize[0] > 0 && cSize[1] > 0 ? cSize[0] * cSize[1] : 0;
    for (int r0 = 0; r0 < nfine; r0 += K) {  // uniform within the group
      const int m = r0 + sub;
      double dO[3] = {0.0, 0.0, 0.0}, dN[3] = {0.0, 0.0, 0.0};
      if (m < nfine) {
        const int sx = m / cSize[1];
        const int s0 = cOff[0] + sx, s1 = cOff[1] + (m - sx * cSize[1]);
        const double sp[2] = {a.spos[0] + (double)s0 * a.srcSub[0],
                              a.spos[1] + (double)s1 * a.srcSub[1]};
        double region = 1.0;
        for (int d = 0; d < 2; ++d) {
          const double hi = dmin(dp[d] + a.dstSub[d], sp[d] + a.srcSub[d]);
          const double lo = dmax(dp[d], sp[d]);
          region *= hi - lo;
          region = dmax(region, 0.0);
        }
        const long long si = tile_index(S, s0, s1);
        const double bv = F.b[si];
        const double vo[3] = {F.H(oldSt)[si] + 1.0 * bv, F.HU(oldSt)[si], F.HV(oldSt)[si]};
        const double vn[3] = {F.H(cur)[si] + 1.0 * bv, F.HU(cur)[si], F.HV(cur)[si]};
        for (int q = 0; q < 3; ++q) {
          dO[q] = (vo[q] * (1.0 - tfOld) + vn[q] * tfOld) * region / a.dstRegion;
          dN[q] = (vo[q] * (1.0 - tfNew) + vn[q] * tfNew) * region / a.dstRegion;
        }
      }
      for (int q = 0; q < 3; ++q) {
        con[g][q][sub] = dO[q];
        con[g][3 + q][sub] = dN[q];
      }
      wave_sync();
      if (sub == 0) {
        const int cnt = nfine - r0 < K ? nfine - r0 : K;
        for (int i = 0; i < cnt; ++i)
          for (int q = 0; q < 3; ++q) {
            accO[q] = accO[q] + con[g][q][i];
            accN[q] = accN[q] + con[g][3 + q][i];
          }
      }
      wave_sync();
    }
  }
  if (cellok && sub == 0) {
    // SWEKernel::postProcessRestriction: h + b -> h with the virtual's b
    const long long di = tile_index(D, di0, di1);
    const double bd = F.b[di];
    F.H(oldSt)[di] = accO[0] + (-1.0) * bd;
    F.HU(oldSt)[di] = accO[1];
    F.HV(oldSt)[di] = accO[2];
    F.H(cur)[di] = accN[0] + (-1.0) * bd;
    F.HU(cur)[di] = accN[1];
    F.HV(cur)[di] = accN[2];
  }
}

__global__ void __launch_bounds__(256)
k_amr_copy(const int* __restrict__ leaves, const PatchDev* __restrict__ pd, Fields F, int cur,
           const AmrClock* clock) {
  if (!amr_advance(*clock)) return;
  const PatchDev P = pd[leaves[blockIdx.x]];
  for (int k = threadIdx.x; k < P.nx * P.ny; k += blockDim.x) {
    const long long i = tile_index(P, k / P.ny, k - (k / P.ny) * P.ny);
    F.H(1 - cur)[i] = F.H(cur)[i];
    F.HU(1 - cur)[i] = F.HU(cur)[i];
    F.HV(1 - cur)[i] = F.HV(cur)[i];
  }
}

// One wavefront per face entry (4 per workgroup).
template <bool LTS>
__global__ void __launch_bounds__(256)
k_amr_ghosts(const AmrGhost* __restrict__ ghosts, int nghosts, const PatchDev* __restrict__ pd,
             Fields F, int cur, const AmrClock* clock, AmrLts L) {
  const AmrClock c = *clock;
  if (!amr_phase_runs<LTS>(clock, L)) return;
  const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (e >= nghosts) return;
  const int lane = threadIdx.x & 63;
  const AmrGhost g = ghosts[e];
  if (LTS && !L.adv[g.dst]) return;
  const PatchDev S = pd[g.src], D = pd[g.dst];
  const int oldSt = 1 - cur;
  // global steps: every leaf runs on the clock (time, dt); local time
  // stepping: each leaf on its own
  const double dTime = LTS ? L.t[g.dst] : c.time, dDt = LTS ? L.dt[g.dst] : c.dt;
  const double sTime = (LTS && g.kind != kAmrTransferVirtual) ? L.t[g.src] : c.time;
  const double sDt = (LTS && g.kind != kAmrTransferVirtual) ? L.dt[g.src] : c.dt;
  const double dstTimeUNew = dTime + dDt;
  if (g.kind != kAmrInterpolate) {
    // DefaultTransfer.cpph:27-51
    double tf;
    if (g.kind == kAmrTransferVirtual) {
      tf = (dstTimeUNew - 0.0) / 1.0;
    } else {
      const double sNew = sTime + sDt, sOld = sTime;
      tf = (sNew - sOld - 0.0 > kLaEps) ? (dstTimeUNew - sOld) / (sNew - sOld) : 1.0;
    }
    const double inv = 1.0 - tf;
    for (int k = lane; k < g.sx * g.sy; k += 64) {
      const int x = k / g.sy, y = k - (k / g.sy) * g.sy;
      const long long si = tile_index(S, g.sox + x, g.soy + y);
      const long long di = tile_index(D, g.dox + x, g.doy + y);
      F.H(oldSt)[di] = F.H(cur)[si] * tf + F.H(oldSt)[si] * inv;
      F.HU(oldSt)[di] = F.HU(cur)[si] * tf + F.HU(oldSt)[si] * inv;
      F.HV(oldSt)[di] = F.HV(cur)[si] * tf + F.HV(oldSt)[si] * inv;
    }
    return;
  }
  // DefaultInterpolation.cpph:45-67 (toCurrentTime = 0, useTimeUNewOrUOld = 1)
  double tf;
  if (fabs(sDt - 0.0) <= kLaEps) {
    tf = 1.0;
  } else {
    const double sOld = sTime, sNew = sTime + sDt;
    tf = (dstTimeUNew - sOld) / (sNew - sOld);
  }
  tf = dmax(0.0, dmin(1.0, tf));
  for (int k = lane; k < g.sx * g.sy; k += 64) {
    const int x = g.dox + k / g.sy, y = g.doy + (k - (k / g.sy) * g.sy);
    const long long di = tile_index(D, x, y);
    double acc[3] = {0.0, 0.0, 0.0};
    double cc[2];
    cc[0] = g.dpos[0] + g.dsub[0] * (double)x + g.dsub[0] * 0.5;
    cc[1] = g.dpos[1] + g.dsub[1] * (double)y + g.dsub[1] * 0.5;
    int si[2];
    for (int d = 0; d < 2; ++d) {
      const double base = g.spos[d] + g.ssub[d] * 0.5;
      const double loc = cc[d] - base;
      int v = (int)floor(loc * g.sinv[d]);
      v = v < g.sn[d] - 2 ? v : g.sn[d] - 2;
      v = v > 0 ? v : 0;
      si[d] = v;
    }
    for (int o = 0; o < 4; ++o) {
      const int off[2] = {o & 1, (o >> 1) & 1};
      const int nb[2] = {si[0] + off[0], si[1] + off[1]};
      double np[2];
      np[0] = g.spos[0] + g.ssub[0] * (double)nb[0] + g.ssub[0] * 0.5;
      np[1] = g.spos[1] + g.ssub[1] * (double)nb[1] + g.ssub[1] * 0.5;
      double sf = 1.0;
      for (int d = 0; d < 2; ++d) {
        const double sign = (double)(1 - 2 * off[d]);
        sf *= 1.0 - (sign * (cc[d] - np[d])) * g.sinv[d];
      }
      const long long sidx = tile_index(S, nb[0], nb[1]);
      const double bv = F.b[sidx];
      const double vo[3] = {F.H(oldSt)[sidx] + 1.0 * bv, F.HU(oldSt)[sidx], F.HV(oldSt)[sidx]};
      const double vn[3] = {F.H(cur)[sidx] + 1.0 * bv, F.HU(cur)[sidx], F.HV(cur)[sidx]};
      for (int q = 0; q < 3; ++q) acc[q] = acc[q] + sf * (vo[q] * (1.0 - tf) + vn[q] * tf);
    }
    F.H(oldSt)[di] = acc[0] + (-1.0) * F.b[di];
    F.HU(oldSt)[di] = acc[1];
    F.HV(oldSt)[di] = acc[2];
  }
}

// SWEKernel::fillBoundaryLayer (SWEKernel.cpp:176-237) of one leaf, its
// walls in call order (corner cells read what the previous wall wrote).
__global__ void __launch_bounds__(64)
k_amr_walls(const AmrWalls* __restrict__ walls, const int* __restrict__ dims,
            const PatchDev* __restrict__ pd, Fields F, int cur, const AmrClock* clock) {
  if (!amr_advance(*clock)) return;
  const AmrWalls W = walls[blockIdx.x];
  const PatchDev P = pd[W.slot];
  const int st = 1 - cur;
  for (int w = 0; w < W.count; ++w) {
    const int dim = dims[2 * (W.first + w)], upper = dims[2 * (W.first + w) + 1];
    const int n_other = dim == 0 ? P.ny : P.nx;
    for (int k = -1 + (int)threadIdx.x; k < n_other + 1; k += blockDim.x) {
      int dx, dy, sx, sy;
      if (dim == 0) {
        dx = upper ? P.nx : -1;
        sx = dx + (upper ? -1 : 1);
        dy = sy = k;
      } else {
        dy = upper ? P.ny : -1;
        sy = dy + (upper ? -1 : 1);
        dx = sx = k;
      }
      const long long d = tile_index(P, dx, dy), s = tile_index(P, sx, sy);
      const double h = F.H(st)[s], hu = F.HU(st)[s], hv = F.HV(st)[s];
      F.H(st)[d] = h;
      F.HU(st)[d] = dim == 0 ? -hu : hu;
      F.HV(st)[d] = dim == 1 ? -hv : hv;
    }
    __syncthreads();
  }
}

template <int SOLVER>
__global__ void __launch_bounds__(64)
k_amr_speed(const Task* __restrict__ tasks, const int* __restrict__ plist,
            const PatchDev* __restrict__ pd, Fields F, int cur, double* __restrict__ smax,
            const AmrClock* clock) {
  __shared__ MarchLds L;
  if (!amr_advance(*clock)) return;
  const Task t = tasks[blockIdx.x];
  EdgeBufs none{{nullptr, nullptr}, nullptr};
  MarchArgs A{plist, pd, F, nullptr, none, 1 - cur, 0.0, smax, nullptr, 1.0};
  // one slot per row segment (tasks may be row segments of their leaves)
  A.smax_stride = kMaxRowSeg;
  march_task<kSpeed, false, 0, SOLVER>(t, A, L);
}

// Global CFL minimum over the leaves (one workgroup), bounded by the step
// left to the end time (AMR driver: min(min over leaves, maxdt)).
__global__ void __launch_bounds__(256)
k_amr_dt(const int* __restrict__ leaves, int nleaves, const PatchDev* __restrict__ pd,
         const double* __restrict__ smax, const int* __restrict__ rseg, AmrClock* clock) {
  __shared__ double red[256];
  const bool adv = amr_advance(*clock);
  if (!adv) {
    if (threadIdx.x == 0) clock->advance = 0;
    return;
  }
  double m = 1.7976931348623157e308;
  for (int k = threadIdx.x; k < nleaves; k += blockDim.x) {
    const int s = leaves[k];
    const PatchDev P = pd[s];
    // the speed maximum of the leaf: over its row segments' slots
    const double* sm = smax + (long long)s * kMaxRowSeg;
    double v = sm[0];
    for (int j = 1; j < rseg[k]; ++j) v = dmax(v, sm[j]);
    m = dmin(m, max_timestep(v, P.dx, P.dy));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = dmin(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double cfl = red[0];
    const double md = clock->end_time - (clock->time + clock->dt);
    clock->bound = md < cfl ? md : cfl;
    clock->advance = 1;
  }
}

// TimeIntervals::isAllowedToAdvanceInTime (TimeIntervals.cpp:24-84) of each
// leaf: the step bound end - (t + dt) > 1e-8 and t + dt - constraint <= 1e-8,
// the constraint being the minimal end time over the neighbours listed in
// nptr/nbr (>= 0: a leaf slot; < 0: -1 - the slot of a virtual patch, whose
// end time is the minimum over the leaves below it: vleaf[vptr[v]..vptr[v+1])
// with v = vidx[slot]).  Counts the advancing leaves into *iter (zeroed by
// the last launch of the previous iteration) and stamps the virtual patches
// they read with the iteration number.
__global__ void __launch_bounds__(256)
k_lts_flags(const int* __restrict__ leaves, int nleaves, const int* __restrict__ nptr,
            const int* __restrict__ nbr, const int* __restrict__ vptr,
            const int* __restrict__ vleaf, const int* __restrict__ vidx,
            const int* __restrict__ rptr, const int* __restrict__ rsrc,
            const AmrClock* clock, AmrLts L) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nleaves) return;
  const int s = leaves[k];
  const double e = L.t[s] + L.dt[s];
  const double md = clock->end_time - e;
  double c = 1.7976931348623157e308;
  for (int j = nptr[k]; j < nptr[k + 1]; ++j) {
    const int q = nbr[j];
    if (q >= 0) {
      c = dmin(c, L.t[q] + L.dt[q]);
    } else {
      const int v = vidx[-1 - q];
      double m = 1.7976931348623157e308;
      for (int i = vptr[v]; i < vptr[v + 1]; ++i) {
        const int x = vleaf[i];
        m = dmin(m, L.t[x] + L.dt[x]);
      }
      c = dmin(c, m);
    }
  }
  const int adv = (md - 0.0 > kLaEps) && !(e - c > kLaEps);
  L.adv[s] = adv;
  if (adv) {
    atomicAdd(L.iter, 1);
    // the virtual patches its ghost faces read (rsrc[rptr[k] .. rptr[k+1]))
    const int stamp = *L.itn;
    for (int j = rptr[k]; j < rptr[k + 1]; ++j) L.need[rsrc[j]] = stamp;
  }
}

// The same decision in one workgroup of 1024 threads for plans of up to
// kLtsBlockLeaves leaves and kLtsBlockVirtuals virtual patches: first the end time of every
// virtual patch (one wavefront per virtual patch, its leaves over the lanes,
// an exact minimum), kept in LDS, then the leaves' flags reading it, so no
// thread walks the leaves below a virtual neighbour itself.
constexpr int kLtsBlockLeaves = 8192;
constexpr int kLtsBlockVirtuals = 4096;  // 32 KiB of LDS
__global__ void __launch_bounds__(1024)
k_lts_flags_block(const int* __restrict__ leaves, int nleaves, const int* __restrict__ nptr,
                  const int* __restrict__ nbr, const int* __restrict__ vptr, int nvirt,
                  const int* __restrict__ vleaf, const int* __restrict__ vidx,
                  const int* __restrict__ rptr, const int* __restrict__ rsrc,
                  const AmrClock* clock, AmrLts L) {
  extern __shared__ double vend[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  for (int v = wave; v < nvirt; v += nwaves) {
    double m = 1.7976931348623157e308;
    for (int i = vptr[v] + lane; i < vptr[v + 1]; i += 64) {
      const int x = vleaf[i];
      m = dmin(m, L.t[x] + L.dt[x]);
    }
    for (int o = 32; o >= 1; o >>= 1) m = dmin(m, __shfl_xor(m, o));
    if (lane == 0) vend[v] = m;
  }
  __syncthreads();
  const double end_time = clock->end_time;
  const int stamp = *L.itn;
  for (int k = threadIdx.x; k < nleaves; k += blockDim.x) {
    const int s = leaves[k];
    const double e = L.t[s] + L.dt[s];
    const double md = end_time - e;
    double c = 1.7976931348623157e308;
    for (int j = nptr[k]; j < nptr[k + 1]; ++j) {
      const int q = nbr[j];
      c = dmin(c, q >= 0 ? L.t[q] + L.dt[q] : vend[vidx[-1 - q]]);
    }
    const int adv = (md - 0.0 > kLaEps) && !(e - c > kLaEps);
    L.adv[s] = adv;
    if (adv) {
      atomicAdd(L.iter, 1);
      for (int j = rptr[k]; j < rptr[k + 1]; ++j) L.need[rsrc[j]] = stamp;
    }
  }
}

// uNew -> uOld and the walls (in call order) of one advancing leaf
// (DefaultTransfer::copyUNewToUOld; SWEKernel::fillBoundaryLayer,
// SWEKernel.cpp:176-237); widx[slot] = its entry in `walls`, -1 for none.
__global__ void __launch_bounds__(256)
k_lts_prep(const int* __restrict__ leaves, const int* __restrict__ widx,
           const AmrWalls* __restrict__ walls, const int* __restrict__ dims,
           const PatchDev* __restrict__ pd, Fields F, int cur, AmrLts L) {
  if (*L.iter == 0) return;
  const int s = leaves[blockIdx.x];
  if (!L.adv[s]) return;
  const PatchDev P = pd[s];
  for (int k = threadIdx.x; k < P.nx * P.ny; k += blockDim.x) {
    const long long i = tile_index(P, k / P.ny, k - (k / P.ny) * P.ny);
    F.H(1 - cur)[i] = F.H(cur)[i];
    F.HU(1 - cur)[i] = F.HU(cur)[i];
    F.HV(1 - cur)[i] = F.HV(cur)[i];
  }
  const int w = widx[s];
  if (w < 0) return;
  __syncthreads();
  const AmrWalls W = walls[w];
  const int st = 1 - cur;
  for (int q = 0; q < W.count; ++q) {
    const int dim = dims[2 * (W.first + q)], upper = dims[2 * (W.first + q) + 1];
    const int n_other = dim == 0 ? P.ny : P.nx;
    for (int k = -1 + (int)threadIdx.x; k < n_other + 1; k += blockDim.x) {
      int dx, dy, sx, sy;
      if (dim == 0) {
        dx = upper ? P.nx : -1;
        sx = dx + (upper ? -1 : 1);
        dy = sy = k;
      } else {
        dy = upper ? P.ny : -1;
        sy = dy + (upper ? -1 : 1);
        dx = sx = k;
      }
      const long long d = tile_index(P, dx, dy), sidx = tile_index(P, sx, sy);
      const double h = F.H(st)[sidx], hu = F.HU(st)[sidx], hv = F.HV(st)[sidx];
      F.H(st)[d] = h;
      F.HU(st)[d] = dim == 0 ? -hu : hu;
      F.HV(st)[d] = dim == 1 ? -hv : hv;
    }
    __syncthreads();
  }
}

// One task of leaves: CFL speeds from uOld with its ghost layer, each
// advancing leaf's step dt = fmin(CFL step, end - (t + dt))
// (SWEKernel.cpp:60-68), the update of the advancing leaves with their own
// steps, then their clocks (:83-87).  Zeroes *iter for the next iteration
// (no block of this launch reads it).
template <int SOLVER>
__global__ void __launch_bounds__(64)
k_lts_task(const Task* __restrict__ tasks, const int* __restrict__ plist,
           const PatchDev* __restrict__ pd, Fields F, int cur, double* __restrict__ smax,
           const AmrClock* clock, AmrLts T) {
  __shared__ MarchLds L;
  const Task t = tasks[blockIdx.x];
  const int lane = threadIdx.x & 63;
  const int W = pow2ceil(t.ny);
  const int seg = lane / W;
  const int col = lane - seg * W + 1;
  const bool segok = seg < t.nseg;
  const int p = plist[t.first + (segok ? seg : 0)];
  const bool lead = segok && col == 1 && T.adv[p] != 0;
  if (__any(lead ? 1 : 0)) {
    EdgeBufs none{{nullptr, nullptr}, nullptr};
    MarchArgs A{plist, pd, F, nullptr, none, 1 - cur, 0.0, smax, nullptr, 1.0};
    march_task<kSpeed, false, 0, SOLVER>(t, A, L);
    __syncthreads();
    if (lead) {
      const PatchDev P = pd[p];
      const double md = clock->end_time - (T.t[p] + T.dt[p]);
      T.stepdt[p] = fmin(max_timestep(smax[p], P.dx, P.dy), md);
    }
    __syncthreads();
    MarchArgs B{plist, pd, F, nullptr, none, cur, 0.0, nullptr, nullptr, 1.0};
    B.seg_dt = T.stepdt;
    B.seg_adv = T.adv;
    march_task<kUpdateSwap, false, 0, SOLVER, true>(t, B, L);
    if (lead) {
      const double dt = T.stepdt[p];
      if (dt - 0.0 > kLaEps) {
        T.t[p] = T.t[p] + T.dt[p];
        T.dt[p] = dt;
      }
      atomicAdd(T.steps, 1ull);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *T.iter = 0;
    *T.itn = *T.itn + 1;
  }
}

// The same iteration in row segments (the global step's segmented tasks):
// exact speeds of every task with an advancing leaf into one slot per row
// segment; the step of every advancing leaf, fmin(own CFL step, end - (t +
// dt)), from its segments' slots; the update of the advancing leaves (their
// uNew -> uOld ran in k_lts_prep, so it reads uOld and writes uNew) and
// their clocks.
template <int SOLVER>
__global__ void __launch_bounds__(64)
k_lts_speed(const Task* __restrict__ tasks, const int* __restrict__ plist,
            const PatchDev* __restrict__ pd, Fields F, int cur, double* __restrict__ smax,
            AmrLts T) {
  __shared__ MarchLds L;
  const Task t = tasks[blockIdx.x];
  const int lane = threadIdx.x & 63;
  const int W = pow2ceil(t.ny);
  const int seg = lane / W;
  const int col = lane - seg * W + 1;
  const bool segok = seg < t.nseg;
  const int p = plist[t.first + (segok ? seg : 0)];
  const bool lead = segok && col == 1 && T.adv[p] != 0;
  if (!__any(lead ? 1 : 0)) return;
  EdgeBufs none{{nullptr, nullptr}, nullptr};
  MarchArgs A{plist, pd, F, nullptr, none, 1 - cur, 0.0, smax, nullptr, 1.0};
  A.smax_stride = kMaxRowSeg;
  march_task<kSpeed, false, 0, SOLVER>(t, A, L);
}

__global__ void __launch_bounds__(256)
k_lts_dt(const int* __restrict__ leaves, int nleaves, const int* __restrict__ rseg,
         const PatchDev* __restrict__ pd, const double* __restrict__ smax,
         const AmrClock* clock, AmrLts T) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nleaves) return;
  const int p = leaves[k];
  if (!T.adv[p]) return;
  const double* sm = smax + (long long)p * kMaxRowSeg;
  double v = sm[0];
  for (int j = 1; j < rseg[k]; ++j) v = dmax(v, sm[j]);
  const PatchDev P = pd[p];
  const double md = clock->end_time - (T.t[p] + T.dt[p]);
  T.stepdt[p] = fmin(max_timestep(v, P.dx, P.dy), md);
}

template <int SOLVER>
__global__ void __launch_bounds__(64)
k_lts_update(const Task* __restrict__ tasks, const int* __restrict__ plist,
             const PatchDev* __restrict__ pd, Fields F, int cur, AmrLts T) {
  __shared__ MarchLds L;
  const Task t = tasks[blockIdx.x];
  const int lane = threadIdx.x & 63;
  const int W = pow2ceil(t.ny);
  const int seg = lane / W;
  const int col = lane - seg * W + 1;
  const bool segok = seg < t.nseg;
  const int p = plist[t.first + (segok ? seg : 0)];
  const bool lead = segok && col == 1 && T.adv[p] != 0;
  if (__any(lead ? 1 : 0)) {
    EdgeBufs none{{nullptr, nullptr}, nullptr};
    MarchArgs B{plist, pd, F, nullptr, none, cur, 0.0, nullptr, nullptr, 1.0};
    B.seg_dt = T.stepdt;
    B.seg_adv = T.adv;
    march_task<kUpdateFromOld, false, 0, SOLVER, true>(t, B, L);
    if (lead && t.sidx == 0) {
      // SWEKernel.cpp:83-87 for the leaf (once, by its first row segment)
      const double dt = T.stepdt[p];
      if (dt - 0.0 > kLaEps) {
        T.t[p] = T.t[p] + T.dt[p];
        T.dt[p] = dt;
      }
      atomicAdd(T.steps, 1ull);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *T.iter = 0;
    *T.itn = *T.itn + 1;
  }
}

template <int SOLVER>
__global__ void __launch_bounds__(64)
k_amr_solve(const Task* __restrict__ tasks, const int* __restrict__ plist,
            const PatchDev* __restrict__ pd, Fields F, int cur, AmrClock* clock) {
  __shared__ MarchLds L;
  // the dt phase decided; this launch's blocks read only advance and bound
  if (!clock->advance) return;
  const Task t = tasks[blockIdx.x];
  const double dt = clock->bound;
  EdgeBufs none{{nullptr, nullptr}, nullptr};
  // every leaf's step fmin(own CFL step, bound) equals the bound, which is
  // at most the minimum of the CFL steps.  uNew -> uOld ran before (k_amr_copy),
  // so the update reads uOld and writes uNew without the swap's uOld store
  // (the swap of SWEKernel.cpp:72-73 would store the values already there).
  MarchArgs A{plist, pd, F, nullptr, none, cur, dt, nullptr, nullptr, 1.0};
  march_task<kUpdateFromOld, false, 0, SOLVER>(t, A, L);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // SWEKernel.cpp:83-87 for every leaf
    if (dt - 0.0 > kLaEps) {
      clock->time = clock->time + clock->dt;
      clock->dt = dt;
    }
    clock->steps += 1;
  }
}

}  // namespace

struct pc_amr {
  pc_pool* pool = nullptr;
  std::vector<int> leaves, virtuals, plan_slots;
  std::vector<int> leaf_nx, leaf_ny;
  int ntasks = 0, nvblocks = 0, nghosts = 0, nwall_leaves = 0;
  Task* d_tasks = nullptr;
  // the global step's speed and update tasks: d_tasks with the rows of
  // every leaf split into row segments when the plan has few tasks; per
  // leaf (plan order) its number of row segments
  int nutasks = 0;
  Task* d_utasks = nullptr;
  int* d_rseg = nullptr;
  double* d_smax_seg = nullptr;
  int* d_plist = nullptr;
  int* d_leaves = nullptr;
  AmrVBlock* d_vblocks = nullptr;
  int* d_cptr = nullptr;
  int* d_cpairs = nullptr;
  AmrRestrict* d_pairs = nullptr;
  AmrGhost* d_ghosts = nullptr;
  AmrWalls* d_walls = nullptr;
  int* d_wall_dims = nullptr;
  double* d_smax = nullptr;
  AmrClock* d_clock = nullptr;
  char* d_tables = nullptr;  // one allocation holding every d_* table above
  size_t tables_cap = 0;
  double last_dt = 0.0;
  bool shared_clock = true;  // all leaves on one clock (global steps allowed)
  // local time stepping: per-slot clocks and flags, neighbour lists of the
  // leaves, leaves below each virtual patch
  std::vector<std::vector<int>> vleaves;  // per virtual patch, plan order
  bool lts_ready = false;
  size_t lts_ns = 0;  // slots covered by the per-slot arrays
  AmrLts lts{};
  int* d_vidx = nullptr;   // per slot: index of the virtual patch, -1 otherwise
  int* d_rptr = nullptr;   // per leaf (plan order): the virtual patches its faces read
  int* d_rsrc = nullptr;
  std::vector<AmrGhost> h_ghosts;
  int* d_widx = nullptr;   // per slot: entry of the leaf's walls, -1 for none
  std::vector<AmrWalls> h_walls;
  int* d_vptr = nullptr;
  int* d_vleaf = nullptr;
  int* d_nptr = nullptr;
  int* d_nbr = nullptr;
  hipGraph_t lgraph = nullptr;
  hipGraphExec_t lgexec = nullptr;
  const PatchDev* lg_pd = nullptr;
  Fields lg_F{};
  int lg_solver = -1, lg_cur = -1;
  // captured step and what it was captured with
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  const PatchDev* g_pd = nullptr;
  Fields g_F{};
  int g_solver = -1, g_cur = -1;
};

namespace {

template <class T>
int amr_upload(T** d, const std::vector<T>& h) {
  HIPCHK(hipMalloc(d, std::max<size_t>(h.size(), 1) * sizeof(T)));
  if (!h.empty()) HIPCHK(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return PC_OK;
}

void amr_free(pc_amr* a) {
  // the table buffer goes back to the pool for the next plan (the caller
  // has synchronised the stream)
  if (a->d_tables && !a->pool->amr_tables) {
    a->pool->amr_tables = a->d_tables;
    a->pool->amr_tables_cap = a->tables_cap;
  } else if (a->d_tables) {
    hipFree(a->d_tables);
  }
  a->d_tables = nullptr;
  if (a->gexec) hipGraphExecDestroy(a->gexec);
  if (a->graph) hipGraphDestroy(a->graph);
  void* lts[] = {a->lts.t,    a->lts.dt,   a->lts.adv, a->lts.stepdt, a->lts.iter, a->lts.steps,
                 a->lts.itn,  a->lts.need, a->d_vidx,  a->d_widx,     a->d_vptr,   a->d_vleaf,
                 a->d_nptr,   a->d_nbr,    a->d_rptr,  a->d_rsrc};
  for (void* q : lts) hipFree(q);
  if (a->lgexec) hipGraphExecDestroy(a->lgexec);
  if (a->lgraph) hipGraphDestroy(a->lgraph);
}

// The plan's slots must still be live with the same role and size.
int amr_check_slots(pc_amr* a) {
  pc_pool* p = a->pool;
  for (size_t i = 0; i < a->leaves.size(); ++i) {
    const int s = a->leaves[i];
    if (!valid_slot(p, s) || p->slots[s].desc.is_virtual ||
        p->slots[s].desc.subdivision_factor[0] != a->leaf_nx[i] ||
        p->slots[s].desc.subdivision_factor[1] != a->leaf_ny[i])
      return fail(PC_E_INVALID, "a leaf of the AMR plan changed or was released");
  }
  for (int s : a->virtuals)
    if (!valid_slot(p, s) || !p->slots[s].desc.is_virtual)
      return fail(PC_E_INVALID, "a virtual patch of the AMR plan changed or was released");
  return PC_OK;
}

int amr_cur(pc_amr* a, int* cur) {
  pc_pool* p = a->pool;
  const int c = cur_of(p, a->plan_slots[0]);
  for (int s : a->plan_slots)
    if (cur_of(p, s) != c)
      return fail(PC_E_INVALID, "AMR patches must share one state parity (one scenario)");
  *cur = c;
  return PC_OK;
}

}  // namespace

extern "C" {

int pc_amr_create(pc_pool* p, const int32_t* leaves, int nleaves, const int32_t* virtuals,
                  int nvirtuals, const int32_t* restrictions, int nrestrictions,
                  const pc_amr_ghost* ghosts, int nghosts, const pc_amr_wall* walls,
                  int nwalls, pc_amr** out) {
  if (!p || !out || nleaves < 1 || nvirtuals < 0 || nrestrictions < 0 || nghosts < 0 ||
      nwalls < 0 || !leaves || (nvirtuals && !virtuals) || (nrestrictions && !restrictions) ||
      (nghosts && !ghosts) || (nwalls && !walls))
    return fail(PC_E_INVALID, "invalid argument");
  hipSetDevice(p->device);
  CHK(prepare(p));
  std::vector<char> role(p->slots.size(), 0);  // 1 leaf, 2 virtual
  for (int i = 0; i < nleaves; ++i) {
    const int s = leaves[i];
    if (!valid_slot(p, s) || p->slots[s].desc.is_virtual || role[s])
      return fail(PC_E_INVALID, "leaves must be distinct live non-virtual patches");
    role[s] = 1;
  }
  for (int i = 0; i < nvirtuals; ++i) {
    const int s = virtuals[i];
    if (!valid_slot(p, s) || !p->slots[s].desc.is_virtual || role[s])
      return fail(PC_E_INVALID, "virtual patches must be distinct live virtual patches");
    role[s] = 2;
  }
  pc_amr* a = new pc_amr;
  a->pool = p;
  a->leaves.assign(leaves, leaves + nleaves);
  a->virtuals.assign(virtuals, virtuals + nvirtuals);
  a->plan_slots = a->leaves;
  a->plan_slots.insert(a->plan_slots.end(), a->virtuals.begin(), a->virtuals.end());
  // leaf clock: every leaf on one clock
  double t0, dt0;
  slot_clock(p, leaves[0], &t0, &dt0);
  for (int s : a->leaves) {
    double t, dt;
    slot_clock(p, s, &t, &dt);
    if (t != t0 || dt != dt0) a->shared_clock = false;
  }
  auto bail = [&](const char* msg) {
    amr_free(a);
    delete a;
    return fail(PC_E_INVALID, msg);
  };
  // leaf tasks: runs of equal-size leaves packed 64 / pow2(ny) per wavefront
  std::vector<Task> tasks;
  std::vector<int> plist;
  for (int s : a->leaves) {
    const auto& d = p->slots[s].desc;
    a->leaf_nx.push_back(d.subdivision_factor[0]);
    a->leaf_ny.push_back(d.subdivision_factor[1]);
    if (d.subdivision_factor[1] > 64) return bail("leaf patches must have ny <= 64");
  }
  for (size_t i = 0; i < a->leaves.size();) {
    const int nx = a->leaf_nx[i], ny = a->leaf_ny[i];
    int w = 1;
    while (w < ny) w <<= 1;
    const int per = 64 / w;
    Task t{(int)plist.size(), 0, nx, ny};
    while (i < a->leaves.size() && t.nseg < per && a->leaf_nx[i] == nx && a->leaf_ny[i] == ny) {
      plist.push_back(a->leaves[i]);
      ++t.nseg;
      ++i;
    }
    tasks.push_back(t);
  }
  a->ntasks = (int)tasks.size();
  // row segments of the global step's tasks: as in the pool's batched step,
  // a plan with fewer tasks than the GPU holds splits each leaf's rows into
  // 2 or 4 segments of at least 8 rows
  std::vector<Task> utasks;
  std::vector<int> rseg_slot(p->slots.size(), 1), rseg;
  {
    constexpr int kFill = 3 * 1024;
    int want = 1;
    while (want < kMaxRowSeg && (long long)a->ntasks * want * 2 <= kFill) want *= 2;
    for (const Task& t : tasks) {
      int r = want;
      while (r > 1 && t.nx / r < 8) r /= 2;
      for (int k = 0; k < r; ++k) {
        Task u = t;
        u.r0 = seg_first_row(k, t.nx, r);
        u.r1 = seg_first_row(k + 1, t.nx, r) - 1;
        u.sidx = k;
        utasks.push_back(u);
      }
      for (int k = 0; k < t.nseg; ++k) rseg_slot[plist[t.first + k]] = r;
    }
    for (int s : a->leaves) rseg.push_back(rseg_slot[s]);
  }
  a->nutasks = (int)utasks.size();
  // restriction pairs grouped by virtual patch, in the given order; per
  // virtual cell the list of pairs covering it
  std::vector<AmrVBlock> vblocks;
  std::vector<int> cptr{0}, cpairs;
  std::vector<AmrRestrict> pairs;
  // the restrictions of each destination slot, in the given order
  std::vector<int> rbeg(role.size() + 1, 0), rlist(nrestrictions);
  for (int r = 0; r < nrestrictions; ++r) {
    const int v = restrictions[2 * r + 1];
    if (v >= 0 && v < (int)role.size()) ++rbeg[v + 1];
  }
  for (size_t k = 1; k < rbeg.size(); ++k) rbeg[k] += rbeg[k - 1];
  {
    std::vector<int> fill(rbeg.begin(), rbeg.end() - 1);
    for (int r = 0; r < nrestrictions; ++r) {
      const int v = restrictions[2 * r + 1];
      if (v >= 0 && v < (int)role.size()) rlist[fill[v]++] = r;
    }
  }
  std::vector<int> ccount;
  for (int vi = 0; vi < nvirtuals; ++vi) {
    const int v = virtuals[vi];
    const int pbeg = (int)pairs.size();
    const auto& dd = p->slots[v].desc;
    const PatchDev D = make_patchdev(p->slots[v]);
    for (int ri = rbeg[v]; ri < rbeg[v + 1]; ++ri) {
      const int r = rlist[ri];
      const int s = restrictions[2 * r];
      if (s < 0 || s >= (int)role.size() || role[s] != 1)
        return bail("a restriction source is not a leaf of the plan");
      const auto& sd = p->slots[s].desc;
      const PatchDev S = make_patchdev(p->slots[s]);
      AmrRestrict R{};
      R.src = s;
      const int z[2] = {0, 0};
      const int sn[2] = {sd.subdivision_factor[0], sd.subdivision_factor[1]};
      const int dn[2] = {dd.subdivision_factor[0], dd.subdivision_factor[1]};
      map_to_patch(z, sn, S, sd, D, dd, 1e-12, R.dOff, R.dSize);
      for (int d = 0; d < 2; ++d) {
        if (R.dOff[d] < 0 || R.dOff[d] + R.dSize[d] > dn[d])
          return bail("restriction source is not inside the destination");
        R.rSize[d] = sn[d];
        R.srcSub[d] = sd.size[d] * (1.0 / (double)sn[d]);
        R.dstSub[d] = dd.size[d] * (1.0 / (double)dn[d]);
        R.spos[d] = sd.position[d];
        R.dpos[d] = dd.position[d];
      }
      R.dstRegion = R.dstSub[0] * R.dstSub[1];
      pairs.push_back(R);
    }
    a->vleaves.emplace_back();
    for (int pi = pbeg; pi < (int)pairs.size(); ++pi) a->vleaves.back().push_back(pairs[pi].src);
    const int nx = dd.subdivision_factor[0], ny = dd.subdivision_factor[1];
    const int base = (int)cptr.size() - 1;
    for (int c0 = 0; c0 < nx * ny; c0 += kRestrictCells)
      vblocks.push_back(AmrVBlock{v, c0, base, 0});
    // per cell the pairs whose region covers it, in pair order: count, then
    // fill pair by pair (each pair marks its region's cells)
    ccount.assign((size_t)nx * ny, 0);
    for (int pi = pbeg; pi < (int)pairs.size(); ++pi) {
      const AmrRestrict& R = pairs[pi];
      for (int x = R.dOff[0]; x < R.dOff[0] + R.dSize[0]; ++x)
        for (int y = R.dOff[1]; y < R.dOff[1] + R.dSize[1]; ++y) ++ccount[(size_t)x * ny + y];
    }
    const size_t cbeg = cpairs.size();
    for (int c = 0; c < nx * ny; ++c) cptr.push_back(cptr.back() + ccount[c]);
    cpairs.resize(cptr.back());
    for (int c = 0; c < nx * ny; ++c) ccount[c] = cptr[base + c] - (int)cbeg;  // cursors
    for (int pi = pbeg; pi < (int)pairs.size(); ++pi) {
      const AmrRestrict& R = pairs[pi];
      for (int x = R.dOff[0]; x < R.dOff[0] + R.dSize[0]; ++x)
        for (int y = R.dOff[1]; y < R.dOff[1] + R.dSize[1]; ++y)
          cpairs[cbeg + ccount[(size_t)x * ny + y]++] = pi;
    }
  }
  for (int r = 0; r < nrestrictions; ++r) {
    const int v = restrictions[2 * r + 1];
    if (v < 0 || v >= (int)role.size() || role[v] != 2)
      return bail("a restriction destination is not a virtual patch of the plan");
  }
  a->nvblocks = (int)vblocks.size();
  // ghost faces
  std::vector<AmrGhost> gh;
  for (int i = 0; i < nghosts; ++i) {
    const pc_amr_ghost& G = ghosts[i];
    const int s = G.src, d = G.dst;
    if (s < 0 || d < 0 || s >= (int)role.size() || d >= (int)role.size() || role[d] != 1 ||
        !role[s])
      return bail("a ghost entry needs a leaf destination and a plan source");
    const auto& sd = p->slots[s].desc;
    const auto& dd = p->slots[d].desc;
    AmrGhost g{};
    g.src = s;
    g.dst = d;
    g.sx = G.size[0];
    g.sy = G.size[1];
    g.dox = G.dst_offset[0];
    g.doy = G.dst_offset[1];
    for (int k = 0; k < 2; ++k)
      if (G.size[k] < 0 || G.dst_offset[k] < -1 ||
          G.dst_offset[k] + G.size[k] > dd.subdivision_factor[k] + 1)
        return bail("ghost destination region out of bounds");
    // destinations are strips of the ghost ring (entries run concurrently
    // and read interior cells only)
    bool ring = false;
    for (int k = 0; k < 2; ++k)
      ring = ring || (G.size[k] == 1 && (G.dst_offset[k] == -1 ||
                                         G.dst_offset[k] == dd.subdivision_factor[k]));
    if (!ring) return bail("ghost destination must be a strip of the ghost layer");
    if (G.kind == 0) {
      g.kind = sd.is_virtual ? kAmrTransferVirtual : kAmrTransferLeaf;
      g.sox = G.src_offset[0];
      g.soy = G.src_offset[1];
      for (int k = 0; k < 2; ++k)
        if (G.src_offset[k] < 0 || G.src_offset[k] + G.size[k] > sd.subdivision_factor[k])
          return bail("transfer source region out of bounds");
    } else if (G.kind == 1) {
      if (role[s] != 1) return bail("interpolation sources must be leaves");
      g.kind = kAmrInterpolate;
      const PatchDev S = make_patchdev(p->slots[s]), D = make_patchdev(p->slots[d]);
      // the cells read (the grown source region, SWEKernel.cpp:252-263)
      // must lie in the source
      int so[2], ss[2];
      map_to_patch(G.dst_offset, G.size, D, dd, S, sd, 1e-12, so, ss);
      for (int k = 0; k < 2; ++k) {
        const int n = sd.subdivision_factor[k];
        if (so[k] > 0) {
          so[k] = so[k] - 1;
          ss[k] = std::min(n, ss[k] + 2);
        } else {
          ss[k] = std::min(n, ss[k] + 1);
        }
        if (so[k] < 0 || so[k] + ss[k] > n + 1 || n < 2)
          return bail("interpolation region is not covered by the source patch");
        g.sn[k] = n;
      }
      g.spos[0] = sd.position[0];
      g.spos[1] = sd.position[1];
      g.ssub[0] = S.dx;
      g.ssub[1] = S.dy;
      g.sinv[0] = 1.0 / S.dx;
      g.sinv[1] = 1.0 / S.dy;
      g.dpos[0] = dd.position[0];
      g.dpos[1] = dd.position[1];
      g.dsub[0] = D.dx;
      g.dsub[1] = D.dy;
    } else {
      return bail("ghost kind must be 0 (transfer) or 1 (interpolation)");
    }
    gh.push_back(g);
  }
  a->nghosts = (int)gh.size();
  a->h_ghosts = gh;
  // walls grouped per leaf in call order
  std::vector<AmrWalls> wl;
  std::vector<int> wdims;
  for (int i = 0; i < nwalls;) {
    const int s = walls[i].slot;
    if (s < 0 || s >= (int)role.size() || role[s] != 1) return bail("walls belong to leaves");
    AmrWalls W{s, (int)(wdims.size() / 2), 0, 0};
    while (i < nwalls && walls[i].slot == s) {
      if ((walls[i].dimension != 0 && walls[i].dimension != 1) ||
          (walls[i].upper != 0 && walls[i].upper != 1))
        return bail("wall dimension and side must be 0 or 1");
      wdims.push_back(walls[i].dimension);
      wdims.push_back(walls[i].upper);
      ++W.count;
      ++i;
    }
    wl.push_back(W);
  }
  a->nwall_leaves = (int)wl.size();
  a->h_walls = wl;
  AmrClock c{};
  c.time = t0;
  c.dt = dt0;
  // every table in one device allocation, filled by one copy (a regrid
  // makes a new plan: one allocation, usually the previous plan's, and one
  // DMA instead of fifteen of each)
  PinnedTables st;
  const size_t o_tasks = st.add(tasks), o_utasks = st.add(utasks), o_rseg = st.add(rseg);
  const size_t o_smax_seg = st.add(nullptr, p->slots.size() * kMaxRowSeg * sizeof(double));
  const size_t o_plist = st.add(plist), o_leaves = st.add(a->leaves);
  const size_t o_vblocks = st.add(vblocks), o_cptr = st.add(cptr), o_cpairs = st.add(cpairs);
  const size_t o_pairs = st.add(pairs), o_ghosts = st.add(gh), o_walls = st.add(wl);
  const size_t o_wall_dims = st.add(wdims);
  const size_t o_smax = st.add(nullptr, p->slots.size() * sizeof(double));
  const std::vector<AmrClock> clk{c};
  const size_t o_clock = st.add(clk);
  hipError_t e = hipSuccess;
  if (p->amr_tables && p->amr_tables_cap >= st.total) {
    a->d_tables = p->amr_tables;  // idle: its plan was destroyed after a sync
    a->tables_cap = p->amr_tables_cap;
    p->amr_tables = nullptr;
    p->amr_tables_cap = 0;
  } else {
    a->tables_cap = 2 * st.total;  // room for a grown grid
    e = hipMalloc(&a->d_tables, a->tables_cap);
  }
  if (e != hipSuccess) {
    amr_free(a);
    delete a;
    return fail(PC_E_HIP, std::string("plan tables: ") + hipGetErrorString(e));
  }
  const int urc = st.upload(p, a->d_tables);
  if (urc != PC_OK) {
    amr_free(a);
    delete a;
    return urc;
  }
  char* const base = a->d_tables;
  a->d_tasks = reinterpret_cast<Task*>(base + o_tasks);
  a->d_utasks = reinterpret_cast<Task*>(base + o_utasks);
  a->d_rseg = reinterpret_cast<int*>(base + o_rseg);
  a->d_smax_seg = reinterpret_cast<double*>(base + o_smax_seg);
  a->d_plist = reinterpret_cast<int*>(base + o_plist);
  a->d_leaves = reinterpret_cast<int*>(base + o_leaves);
  a->d_vblocks = reinterpret_cast<AmrVBlock*>(base + o_vblocks);
  a->d_cptr = reinterpret_cast<int*>(base + o_cptr);
  a->d_cpairs = reinterpret_cast<int*>(base + o_cpairs);
  a->d_pairs = reinterpret_cast<AmrRestrict*>(base + o_pairs);
  a->d_ghosts = reinterpret_cast<AmrGhost*>(base + o_ghosts);
  a->d_walls = reinterpret_cast<AmrWalls*>(base + o_walls);
  a->d_wall_dims = reinterpret_cast<int*>(base + o_wall_dims);
  a->d_smax = reinterpret_cast<double*>(base + o_smax);
  a->d_clock = reinterpret_cast<AmrClock*>(base + o_clock);
  *out = a;
  return PC_OK;
}

}  // extern "C"

namespace {

void amr_launch_step(pc_amr* a, Fields F, int cur) {
  pc_pool* p = a->pool;
  const int nl = (int)a->leaves.size();
  const AmrLts none{};  // unused by the global phases
  if (a->nvblocks > 0)
    hipLaunchKernelGGL(k_amr_restrict<false>, dim3(a->nvblocks), dim3(64), 0, p->stream,
                       a->d_vblocks, a->d_cptr, a->d_cpairs, a->d_pairs, p->d_pd, F, cur,
                       a->d_clock, none);
  hipLaunchKernelGGL(k_amr_copy, dim3(nl), dim3(256), 0, p->stream, a->d_leaves, p->d_pd, F,
                     cur, a->d_clock);
  if (a->nghosts > 0)
    hipLaunchKernelGGL(k_amr_ghosts<false>, dim3((a->nghosts + 3) / 4), dim3(256), 0, p->stream,
                       a->d_ghosts, a->nghosts, p->d_pd, F, cur, a->d_clock, none);
  if (a->nwall_leaves > 0)
    hipLaunchKernelGGL(k_amr_walls, dim3(a->nwall_leaves), dim3(64), 0, p->stream, a->d_walls,
                       a->d_wall_dims, p->d_pd, F, cur, a->d_clock);
  if (p->solver)
    hipLaunchKernelGGL(k_amr_speed<1>, dim3(a->nutasks), dim3(64), 0, p->stream, a->d_utasks,
                       a->d_plist, p->d_pd, F, cur, a->d_smax_seg, a->d_clock);
  else
    hipLaunchKernelGGL(k_amr_speed<0>, dim3(a->nutasks), dim3(64), 0, p->stream, a->d_utasks,
                       a->d_plist, p->d_pd, F, cur, a->d_smax_seg, a->d_clock);
  hipLaunchKernelGGL(k_amr_dt, dim3(1), dim3(256), 0, p->stream, a->d_leaves, nl, p->d_pd,
                     a->d_smax_seg, a->d_rseg, a->d_clock);
  if (p->solver)
    hipLaunchKernelGGL(k_amr_solve<1>, dim3(a->nutasks), dim3(64), 0, p->stream, a->d_utasks,
                       a->d_plist, p->d_pd, F, cur, a->d_clock);
  else
    hipLaunchKernelGGL(k_amr_solve<0>, dim3(a->nutasks), dim3(64), 0, p->stream, a->d_utasks,
                       a->d_plist, p->d_pd, F, cur, a->d_clock);
}

// One grid iteration with local time stepping: advance flags from the
// clocks at its start; restriction of every leaf on its own interval; the
// ghost faces of the advancing leaves (every neighbour read before any leaf
// changes); uNew -> uOld and walls; then CFL speeds, steps, update and
// clocks, in three launches over row segments or one over whole leaves.
void amr_launch_lts(pc_amr* a, Fields F, int cur) {
  pc_pool* p = a->pool;
  const int nl = (int)a->leaves.size();
  const AmrLts& L = a->lts;
  const int nv = (int)a->vleaves.size();
  if (nl <= kLtsBlockLeaves && nv <= kLtsBlockVirtuals)
    hipLaunchKernelGGL(k_lts_flags_block, dim3(1), dim3(1024), nv * sizeof(double), p->stream,
                       a->d_leaves, nl, a->d_nptr, a->d_nbr, a->d_vptr, nv, a->d_vleaf,
                       a->d_vidx, a->d_rptr, a->d_rsrc, a->d_clock, L);
  else
    hipLaunchKernelGGL(k_lts_flags, dim3((nl + 255) / 256), dim3(256), 0, p->stream,
                       a->d_leaves, nl, a->d_nptr, a->d_nbr, a->d_vptr, a->d_vleaf, a->d_vidx,
                       a->d_rptr, a->d_rsrc, a->d_clock, L);
  if (a->nvblocks > 0)
    hipLaunchKernelGGL(k_amr_restrict<true>, dim3(a->nvblocks), dim3(64), 0, p->stream,
                       a->d_vblocks, a->d_cptr, a->d_cpairs, a->d_pairs, p->d_pd, F, cur,
                       a->d_clock, L);
  if (a->nghosts > 0)
    hipLaunchKernelGGL(k_amr_ghosts<true>, dim3((a->nghosts + 3) / 4), dim3(256), 0, p->stream,
                       a->d_ghosts, a->nghosts, p->d_pd, F, cur, a->d_clock, L);
  hipLaunchKernelGGL(k_lts_prep, dim3(nl), dim3(256), 0, p->stream, a->d_leaves, a->d_widx,
                     a->d_walls, a->d_wall_dims, p->d_pd, F, cur, L);
  if (a->nutasks == a->ntasks) {
    // whole leaves: speeds, steps, update and clocks in one launch
    if (p->solver)
      hipLaunchKernelGGL(k_lts_task<1>, dim3(a->ntasks), dim3(64), 0, p->stream, a->d_tasks,
                         a->d_plist, p->d_pd, F, cur, a->d_smax, a->d_clock, L);
    else
      hipLaunchKernelGGL(k_lts_task<0>, dim3(a->ntasks), dim3(64), 0, p->stream, a->d_tasks,
                         a->d_plist, p->d_pd, F, cur, a->d_smax, a->d_clock, L);
    return;
  }
  if (p->solver)
    hipLaunchKernelGGL(k_lts_speed<1>, dim3(a->nutasks), dim3(64), 0, p->stream, a->d_utasks,
                       a->d_plist, p->d_pd, F, cur, a->d_smax_seg, L);
  else
    hipLaunchKernelGGL(k_lts_speed<0>, dim3(a->nutasks), dim3(64), 0, p->stream, a->d_utasks,
                       a->d_plist, p->d_pd, F, cur, a->d_smax_seg, L);
  hipLaunchKernelGGL(k_lts_dt, dim3((nl + 255) / 256), dim3(256), 0, p->stream, a->d_leaves, nl,
                     a->d_rseg, p->d_pd, a->d_smax_seg, a->d_clock, L);
  if (p->solver)
    hipLaunchKernelGGL(k_lts_update<1>, dim3(a->nutasks), dim3(64), 0, p->stream, a->d_utasks,
                       a->d_plist, p->d_pd, F, cur, L);
  else
    hipLaunchKernelGGL(k_lts_update<0>, dim3(a->nutasks), dim3(64), 0, p->stream, a->d_utasks,
                       a->d_plist, p->d_pd, F, cur, L);
}

bool same_fields(const Fields& x, const Fields& y) {
  return x.h[0] == y.h[0] && x.h[1] == y.h[1] && x.hu[0] == y.hu[0] && x.hu[1] == y.hu[1] &&
         x.hv[0] == y.hv[0] && x.hv[1] == y.hv[1] && x.b == y.b;
}

// (Re)capture the step graph when the pool's arrays, the solver or the
// state parity differ from those it was captured with.
int amr_graph(pc_amr* a, Fields F, int cur) {
  pc_pool* p = a->pool;
  if (a->gexec && a->g_pd == p->d_pd && same_fields(a->g_F, F) && a->g_solver == p->solver &&
      a->g_cur == cur)
    return PC_OK;
  if (a->gexec) HIPCHK(hipGraphExecDestroy(a->gexec));
  if (a->graph) HIPCHK(hipGraphDestroy(a->graph));
  a->gexec = nullptr;
  a->graph = nullptr;
  HIPCHK(hipStreamSynchronize(p->stream));
  HIPCHK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
  amr_launch_step(a, F, cur);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(p->stream, &g);
  if (e != hipSuccess) return fail(PC_E_HIP, std::string("capture: ") + hipGetErrorString(e));
  a->graph = g;
  HIPCHK(hipGraphInstantiate(&a->gexec, g, nullptr, nullptr, 0));
  a->g_pd = p->d_pd;
  a->g_F = F;
  a->g_solver = p->solver;
  a->g_cur = cur;
  return PC_OK;
}

int amr_lts_graph(pc_amr* a, Fields F, int cur) {
  pc_pool* p = a->pool;
  if (a->lgexec && a->lg_pd == p->d_pd && same_fields(a->lg_F, F) &&
      a->lg_solver == p->solver && a->lg_cur == cur)
    return PC_OK;
  if (a->lgexec) HIPCHK(hipGraphExecDestroy(a->lgexec));
  if (a->lgraph) HIPCHK(hipGraphDestroy(a->lgraph));
  a->lgexec = nullptr;
  a->lgraph = nullptr;
  HIPCHK(hipStreamSynchronize(p->stream));
  HIPCHK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
  amr_launch_lts(a, F, cur);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(p->stream, &g);
  if (e != hipSuccess) return fail(PC_E_HIP, std::string("capture: ") + hipGetErrorString(e));
  a->lgraph = g;
  HIPCHK(hipGraphInstantiate(&a->lgexec, g, nullptr, nullptr, 0));
  a->lg_pd = p->d_pd;
  a->lg_F = F;
  a->lg_solver = p->solver;
  a->lg_cur = cur;
  return PC_OK;
}

}  // namespace

extern "C" {

int pc_amr_lts_init(pc_amr* a, const int32_t* nptr, const int32_t* nbr) {
  if (!a || !nptr || (nptr[a->leaves.size()] > 0 && !nbr))
    return fail(PC_E_INVALID, "invalid argument");
  pc_pool* p = a->pool;
  hipSetDevice(p->device);
  CHK(prepare(p));
  CHK(amr_check_slots(a));
  const int nl = (int)a->leaves.size();
  std::vector<char> role(p->slots.size(), 0);
  for (int s : a->leaves) role[s] = 1;
  for (int s : a->virtuals) role[s] = 2;
  std::vector<int> np(nptr, nptr + nl + 1), nb;
  if (np[0] != 0) return fail(PC_E_INVALID, "neighbour lists must start at 0");
  for (int k = 0; k < nl; ++k) {
    if (np[k + 1] < np[k]) return fail(PC_E_INVALID, "neighbour list offsets must not decrease");
    for (int j = np[k]; j < np[k + 1]; ++j) {
      const int q = nbr[j];
      if (q < 0 || q >= (int)role.size() || !role[q])
        return fail(PC_E_INVALID, "a neighbour is not a patch of the plan");
      nb.push_back(role[q] == 1 ? q : -1 - q);
    }
  }
  std::vector<int> vptr{0}, vleaf;
  for (const auto& v : a->vleaves) {
    vleaf.insert(vleaf.end(), v.begin(), v.end());
    vptr.push_back((int)vleaf.size());
  }
  const size_t ns = p->slots.size();
  std::vector<double> t(ns, 0.0), dt(ns, 0.0);
  for (int s : a->leaves) slot_clock(p, s, &t[s], &dt[s]);
  void* old[] = {a->lts.t,    a->lts.dt,   a->lts.adv, a->lts.stepdt, a->lts.iter, a->lts.steps,
                 a->lts.itn,  a->lts.need, a->d_vidx,  a->d_widx,     a->d_vptr,   a->d_vleaf,
                 a->d_nptr,   a->d_nbr,    a->d_rptr,  a->d_rsrc};
  for (void* q : old) hipFree(q);
  a->lts = AmrLts{};
  a->d_vidx = a->d_widx = a->d_vptr = a->d_vleaf = a->d_nptr = a->d_nbr = nullptr;
  a->d_rptr = a->d_rsrc = nullptr;
  if (a->lgexec) hipGraphExecDestroy(a->lgexec);
  if (a->lgraph) hipGraphDestroy(a->lgraph);
  a->lgexec = nullptr;
  a->lgraph = nullptr;
  CHK(amr_upload(&a->lts.t, t));
  CHK(amr_upload(&a->lts.dt, dt));
  CHK(amr_upload(&a->lts.adv, std::vector<int>(ns, 0)));
  CHK(amr_upload(&a->lts.stepdt, std::vector<double>(ns, 0.0)));
  CHK(amr_upload(&a->lts.iter, std::vector<int>(1, 0)));
  CHK(amr_upload(&a->lts.steps, std::vector<unsigned long long>(1, 0)));
  std::vector<int> vidx(ns, -1), widx(ns, -1);
  for (size_t v = 0; v < a->virtuals.size(); ++v) vidx[a->virtuals[v]] = (int)v;
  for (size_t w = 0; w < a->h_walls.size(); ++w) widx[a->h_walls[w].slot] = (int)w;
  // per leaf the virtual patches its ghost faces read
  std::vector<int> lidx(ns, -1), rptr{0}, rsrc;
  for (int k = 0; k < nl; ++k) lidx[a->leaves[k]] = k;
  std::vector<std::vector<int>> reads(nl);
  for (const AmrGhost& g : a->h_ghosts)
    if (g.kind == kAmrTransferVirtual) reads[lidx[g.dst]].push_back(g.src);
  for (int k = 0; k < nl; ++k) {
    rsrc.insert(rsrc.end(), reads[k].begin(), reads[k].end());
    rptr.push_back((int)rsrc.size());
  }
  CHK(amr_upload(&a->d_rptr, rptr));
  CHK(amr_upload(&a->d_rsrc, rsrc));
  CHK(amr_upload(&a->lts.itn, std::vector<int>(1, 1)));
  CHK(amr_upload(&a->lts.need, std::vector<int>(ns, 0)));
  CHK(amr_upload(&a->d_vidx, vidx));
  CHK(amr_upload(&a->d_widx, widx));
  CHK(amr_upload(&a->d_vptr, vptr));
  CHK(amr_upload(&a->d_vleaf, vleaf));
  CHK(amr_upload(&a->d_nptr, np));
  CHK(amr_upload(&a->d_nbr, nb));
  a->lts_ns = ns;
  a->lts_ready = true;
  return PC_OK;
}

int pc_amr_lts_step(pc_amr* a, int k, double end_time, long long* leaf_steps) {
  if (!a || k < 0) return fail(PC_E_INVALID, "invalid argument");
  if (!a->lts_ready) return fail(PC_E_INVALID, "pc_amr_lts_init has not been called");
  pc_pool* p = a->pool;
  hipSetDevice(p->device);
  CHK(prepare(p));
  CHK(amr_check_slots(a));
  int cur;
  CHK(amr_cur(a, &cur));
  mark_state_changed(p);
  a->shared_clock = false;
  Fields F = fields_of(p);
  CHK(amr_lts_graph(a, F, cur));
  HIPCHK(hipMemcpyAsync(&a->d_clock->end_time, &end_time, sizeof(double),
                        hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipMemsetAsync(a->lts.steps, 0, sizeof(unsigned long long), p->stream));
  for (int it = 0; it < k; ++it) HIPCHK(hipGraphLaunch(a->lgexec, p->stream));
  unsigned long long n = 0;
  const size_t ns = a->lts_ns;
  std::vector<double> t(ns), dt(ns);
  HIPCHK(hipMemcpyAsync(&n, a->lts.steps, sizeof(n), hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipMemcpyAsync(t.data(), a->lts.t, ns * sizeof(double), hipMemcpyDeviceToHost,
                        p->stream));
  HIPCHK(hipMemcpyAsync(dt.data(), a->lts.dt, ns * sizeof(double), hipMemcpyDeviceToHost,
                        p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  // the leaves' descriptors follow their device clocks
  for (int s : a->leaves) {
    p->slots[s].desc.time = t[s];
    p->slots[s].desc.timestep_size = dt[s];
    p->slots[s].clock_from_scenario = false;
  }
  if (leaf_steps) *leaf_steps = (long long)n;
  return PC_OK;
}

int pc_amr_step(pc_amr* a, int k, double end_time, double* last_dt) {
  if (!a || k < 0) return fail(PC_E_INVALID, "invalid argument");
  if (!a->shared_clock)
    return fail(PC_E_INVALID, "global steps need every leaf of the plan on one clock");
  pc_pool* p = a->pool;
  hipSetDevice(p->device);
  CHK(prepare(p));
  CHK(amr_check_slots(a));
  int cur;
  CHK(amr_cur(a, &cur));
  mark_state_changed(p);
  Fields F = fields_of(p);
  CHK(amr_graph(a, F, cur));
  HIPCHK(hipMemcpyAsync(&a->d_clock->end_time, &end_time, sizeof(double),
                        hipMemcpyHostToDevice, p->stream));
  for (int it = 0; it < k; ++it) HIPCHK(hipGraphLaunch(a->gexec, p->stream));
  AmrClock c;
  HIPCHK(hipMemcpyAsync(&c, a->d_clock, sizeof(c), hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  // the leaves' descriptors follow the device clock (as pc_solve_timestep
  // would have left them)
  for (int s : a->leaves) {
    p->slots[s].desc.time = c.time;
    p->slots[s].desc.timestep_size = c.dt;
    p->slots[s].clock_from_scenario = false;
  }
  a->last_dt = (k > 0 && c.advance) ? c.bound : 0.0;
  if (last_dt) *last_dt = a->last_dt;
  return PC_OK;
}

int pc_amr_clock(pc_amr* a, double* time, double* timestep_size, long long* steps) {
  if (!a) return fail(PC_E_INVALID, "invalid argument");
  hipSetDevice(a->pool->device);
  AmrClock c;
  HIPCHK(hipMemcpyAsync(&c, a->d_clock, sizeof(c), hipMemcpyDeviceToHost, a->pool->stream));
  HIPCHK(hipStreamSynchronize(a->pool->stream));
  if (time) *time = c.time;
  if (timestep_size) *timestep_size = c.dt;
  if (steps) *steps = c.steps;
  return PC_OK;
}

// Face, corner and wall entries of every leaf of a tree of patches with
// refinement factor 3 (host only, no device work): the rows pc_amr_create
// takes.  Patch k has level[k] >= 1 and index (ij[2k], ij[2k+1]) on its
// level; level 1 is a px x py grid.  For each leaf in `leaves` order
// (indices into the patch arrays): its four faces x-, x+, y-, y+ (a face
// outside the domain is a wall), then its four corners (-,-), (-,+), (+,-),
// (+,+) (skipped outside the domain); a face or corner is a transfer from
// the patch of the leaf's level there, else an interpolation from the
// nearest coarser patch covering it (the compositor's choice,
// GhostLayerCompositor.cpp:280-320, GhostlayerCompositorFunctors.cpp:87-141);
// then its walls in the order left, top, right, bottom
// (SolveTimestep.cpp:94-131).  Entries name patches by handle[k].
int pc_amr_face_plan(int px, int py, int n, int npatch, const int32_t* level,
                     const int32_t* ij, const int32_t* handle, const int32_t* leaves,
                     int nleaves, pc_amr_ghost* ghosts, int cap_ghosts, int* nghosts,
                     pc_amr_wall* walls, int cap_walls, int* nwalls) {
  if (px < 1 || py < 1 || n < 1 || npatch < 0 || nleaves < 0 || !nghosts || !nwalls ||
      (npatch && (!level || !ij || !handle)) || (nleaves && !leaves))
    return fail(PC_E_INVALID, "invalid argument");
  // index grid of every level: px 3^(l-1) x py 3^(l-1) patches
  constexpr int kMaxLevel = 12;
  long long gw[kMaxLevel + 1][2];
  for (int l = 1; l <= kMaxLevel; ++l) {
    gw[l][0] = l == 1 ? px : 3 * gw[l - 1][0];
    gw[l][1] = l == 1 ? py : 3 * gw[l - 1][1];
  }
  // patch lookup per level: a dense table where the level's grid is small
  // against the patch count, else a hash map (25 bits per index in its key)
  auto key = [](long long l, long long i, long long j) { return (l << 50) | (i << 25) | j; };
  std::unordered_map<long long, int> at;
  std::vector<int> dense[kMaxLevel + 1];
  int nat = 0;
  for (int k = 0; k < npatch; ++k) {
    const int l = level[k];
    if (l < 1 || l > kMaxLevel || ij[2 * k] < 0 || ij[2 * k + 1] < 0 ||
        ij[2 * k] >= gw[l][0] || ij[2 * k + 1] >= gw[l][1] || ij[2 * k] >= (1 << 25) ||
        ij[2 * k + 1] >= (1 << 25))
      return fail(PC_E_INVALID, "invalid patch level or index");
  }
  const long long dense_cap = 16LL * npatch + 65536;
  for (int k = 0; k < npatch; ++k) {
    const int l = level[k];
    if (dense[l].empty() && gw[l][0] <= dense_cap / gw[l][1])
      dense[l].assign(gw[l][0] * gw[l][1], -1);
    if (!dense[l].empty())
      dense[l][ij[2 * k] * gw[l][1] + ij[2 * k + 1]] = k;
    else
      ++nat;
  }
  if (nat) {
    at.reserve((size_t)nat * 2);
    for (int k = 0; k < npatch; ++k)
      if (dense[level[k]].empty()) at[key(level[k], ij[2 * k], ij[2 * k + 1])] = k;
  }
  auto find = [&](int lev, long long i, long long j) {
    if (!dense[lev].empty()) return dense[lev][i * gw[lev][1] + j];
    if (!nat) return -1;
    auto it = at.find(key(lev, i, j));
    return it != at.end() ? it->second : -1;
  };
  // the patch at (lev, i, j), else the nearest coarser one covering it;
  // -1 outside the domain or where nothing covers it
  auto covering = [&](int lev, long long i, long long j) {
    if (i < 0 || j < 0 || i >= gw[lev][0] || j >= gw[lev][1]) return -1;
    for (; lev >= 1; --lev, i /= 3, j /= 3) {
      const int k = find(lev, i, j);
      if (k >= 0) return k;
    }
    return -1;
  };
  const int steps[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};  // x-, x+, y-, y+
  const int geom[4][6] = {  // size, source offset, destination offset per face
      {1, n, n - 1, 0, -1, 0}, {1, n, 0, 0, n, 0}, {n, 1, 0, n - 1, 0, -1}, {n, 1, 0, 0, 0, n}};
  const int corners[4][2] = {{-1, -1}, {-1, 1}, {1, -1}, {1, 1}};
  const int wall_face[4] = {0, 3, 1, 2};  // left, top, right, bottom as face indices
  int ng = 0, nw = 0;
  for (int li = 0; li < nleaves; ++li) {
    const int p = leaves[li];
    if (p < 0 || p >= npatch) return fail(PC_E_INVALID, "invalid leaf index");
    const int lev = level[p];
    const long long i = ij[2 * p], j = ij[2 * p + 1];
    const long long wx = gw[lev][0], wy = gw[lev][1];
    bool wall[4];
    auto emit = [&](int src, int sx, int sy, int sox, int soy, int dox, int doy) {
      if (ghosts && ng < cap_ghosts) {
        pc_amr_ghost& g = ghosts[ng];
        g.kind = level[src] == lev ? 0 : 1;
        g.src = handle[src];
        g.dst = handle[p];
        g.size[0] = sx;
        g.size[1] = sy;
        g.src_offset[0] = sox;
        g.src_offset[1] = soy;
        g.dst_offset[0] = dox;
        g.dst_offset[1] = doy;
      }
      ++ng;
    };
    for (int f = 0; f < 4; ++f) {
      const long long qi = i + steps[f][0], qj = j + steps[f][1];
      wall[f] = qi < 0 || qj < 0 || qi >= wx || qj >= wy;
      if (wall[f]) continue;
      const int src = covering(lev, qi, qj);
      if (src < 0) return fail(PC_E_INVALID, "face region not covered");
      const int* g = geom[f];
      emit(src, g[0], g[1], g[2], g[3], g[4], g[5]);
    }
    for (const auto& c : corners) {
      const int src = covering(lev, i + c[0], j + c[1]);
      if (src < 0) continue;  // outside the domain: the walls fill it
      emit(src, 1, 1, c[0] == -1 ? n - 1 : 0, c[1] == -1 ? n - 1 : 0, c[0] == -1 ? -1 : n,
           c[1] == -1 ? -1 : n);
    }
    for (int w = 0; w < 4; ++w) {
      const int f = wall_face[w];
      if (!wall[f]) continue;
      if (walls && nw < cap_walls) {
        walls[nw].slot = handle[p];
        walls[nw].dimension = f / 2;
        walls[nw].upper = f % 2;
      }
      ++nw;
    }
  }
  *nghosts = ng;
  *nwalls = nw;
  if ((ghosts && ng > cap_ghosts) || (walls && nw > cap_walls))
    return fail(PC_E_INVALID, "output capacity too small");
  return PC_OK;
}

int pc_amr_destroy(pc_amr* a) {
  if (!a) return fail(PC_E_INVALID, "invalid argument");
  hipSetDevice(a->pool->device);
  hipStreamSynchronize(a->pool->stream);
  amr_free(a);
  delete a;
  return PC_OK;
}

}  // extern "C"
