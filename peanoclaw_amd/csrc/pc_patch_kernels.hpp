// Per-patch kernels behind the Numerics-mirroring entry points and the
// grid-level transfers (reflect, same-level transfer with time factor,
// d-linear interpolation, area-weighted restriction, h <-> h+b transform,
// refinement criterion).  All FP64, same operation order as the formulas
// they restate (file:line given per kernel).
#pragma once

#include "pc_kernels.hpp"

namespace pcamd {

// a[i] = i: the one-patch lists of the per-patch kernels (plist + slot)
__global__ void k_iota(int* __restrict__ a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}

__device__ __forceinline__ long long tile_index(const PatchDev& P, int x, int y) {
  // cell (x, y) with x, y in [-1, n]
  return P.off + (long long)(x + 1) * P.stride + (y + 1);
}

// Row segments of the one-patch kernels (one wavefront each): at least 2
// rows each, at most kPatchSeg.  A single call is latency-bound, so the
// rows are spread over as many wavefronts as the patch allows.
constexpr int kPatchSeg = 8;
__host__ __device__ __forceinline__ int patch_row_segments(int nx) {
  const int s = nx / 2;
  return s < 1 ? 1 : (s > kPatchSeg ? kPatchSeg : s);
}

// Completion word of a synchronous per-call kernel: `seq` stored to
// out[kDoneWord] (fine-grained host memory) after the results, with release
// ordering at system scope, so the host sees the results once it sees seq.
constexpr int kDoneWord = 4;
__device__ __forceinline__ void signal_done(double* out, unsigned long long seq) {
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(out + kDoneWord), seq,
                     __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// SWEKernel::solveTimestep on one patch (SWEKernel.cpp:50-88): speed of
// every edge of uOld, dt = fmin(CFL, maxDt), update, swap.  Both the speeds
// and the update run in blockDim/64 row segments, one wavefront each: the
// speeds' maxima meet in `segmax` (kPatchSeg doubles), every wavefront then
// takes dt from them.  The update reads uOld and writes uNew only, so the
// segments run side by side; the old uNew, which becomes uOld, waits in
// `keep` (3 n^2 doubles) until every segment has read uOld.  `plist` holds
// plist[slot] == slot in global memory (the march reads its patch list with
// global loads).
template <int SOLVER>
__global__ void __launch_bounds__(64 * kPatchSeg)
k_patch_solve(int slot, const int* __restrict__ plist, const PatchDev* __restrict__ pd,
              Fields F, int cur, double maxDt, double* __restrict__ segmax,
              double* __restrict__ out, double* __restrict__ keep, unsigned long long seq) {
  __shared__ MarchLds L[kPatchSeg];
  const int w = (int)threadIdx.x >> 6, nw = (int)blockDim.x >> 6;
  const PatchDev P = pd[slot];
  Task t;
  t.first = slot; t.nseg = 1; t.nx = P.nx; t.ny = P.ny;
  if (nw > 1) {
    t.r0 = seg_first_row(w, P.nx, nw);
    t.r1 = seg_first_row(w + 1, P.nx, nw) - 1;
    t.sidx = w;
  }
  EdgeBufs none{{nullptr, nullptr}, nullptr};
  MarchArgs A{plist, pd, F, nullptr, none, 1 - cur, 0.0, segmax, nullptr, 1.0};
  A.smax_stride = 0;  // segment w's maximum to segmax[w]
  march_task<kSpeed, false, 0, SOLVER>(t, A, L[w]);
  const int n = P.nx * P.ny;
  for (int k = (int)threadIdx.x; k < n; k += (int)blockDim.x) {
    const long long i = tile_index(P, k / P.ny, k - (k / P.ny) * P.ny);
    keep[k] = F.H(cur)[i];
    keep[n + k] = F.HU(cur)[i];
    keep[2 * n + k] = F.HV(cur)[i];
  }
  __syncthreads();
  double s = segmax[0];
  for (int k = 1; k < nw; ++k) s = dmax(s, segmax[k]);
  const double cfl = max_timestep(s, P.dx, P.dy);
  const double dt = fmin(cfl, maxDt);
  // the host needs only dt and the estimate: it is released here, and the
  // update below finishes on the stream before any later work on the pool
  if (threadIdx.x == 0) {
    out[0] = dt;
    out[1] = cfl;
    signal_done(out, seq);
  }
  A.par = cur;
  A.dt = dt;
  A.smax = nullptr;
  march_task<kUpdateFromOld, false, 0, SOLVER>(t, A, L[w]);
  __syncthreads();
  for (int k = (int)threadIdx.x; k < n; k += (int)blockDim.x) {
    const long long i = tile_index(P, k / P.ny, k - (k / P.ny) * P.ny);
    F.H(1 - cur)[i] = keep[k];
    F.HU(1 - cur)[i] = keep[n + k];
    F.HV(1 - cur)[i] = keep[2 * n + k];
  }
}

// The CFL step of one patch from uOld (state 1 - cur) with its ghost layer,
// without the update: out[0] = max_timestep(max edge speed); row segments
// as in k_patch_solve.
template <int SOLVER>
__global__ void __launch_bounds__(64 * kPatchSeg)
k_patch_cfl(int slot, const int* __restrict__ plist, const PatchDev* __restrict__ pd,
            Fields F, int cur, double* __restrict__ segmax, double* __restrict__ out,
            unsigned long long seq) {
  __shared__ MarchLds L[kPatchSeg];
  const int w = (int)threadIdx.x >> 6, nw = (int)blockDim.x >> 6;
  const PatchDev P = pd[slot];
  Task t;
  t.first = slot; t.nseg = 1; t.nx = P.nx; t.ny = P.ny;
  if (nw > 1) {
    t.r0 = seg_first_row(w, P.nx, nw);
    t.r1 = seg_first_row(w + 1, P.nx, nw) - 1;
    t.sidx = w;
  }
  EdgeBufs none{{nullptr, nullptr}, nullptr};
  MarchArgs A{plist, pd, F, nullptr, none, 1 - cur, 0.0, segmax, nullptr, 1.0};
  A.smax_stride = 0;
  march_task<kSpeed, false, 0, SOLVER>(t, A, L[w]);
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = segmax[0];
    for (int k = 1; k < nw; ++k) s = dmax(s, segmax[k]);
    out[0] = max_timestep(s, P.dx, P.dy);
    signal_done(out, seq);
  }
}

// SWEKernel::fillBoundaryLayer (SWEKernel.cpp:176-237) on uOld (state st).
__global__ void k_fill_boundary(const PatchDev* __restrict__ pd, int slot, Fields F,
                                int st, int dim, int upper) {
  const PatchDev P = pd[slot];
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
}

// DefaultTransfer::transferGhostlayer (DefaultTransfer.cpph:15-101).
__global__ void k_transfer(const PatchDev* __restrict__ pd, int srcSlot, int dstSlot,
                           Fields F, int srcCur, int dstCur, int sx, int sy,
                           int sox, int soy, int dox, int doy, double tf,
                           double inv) {
  const PatchDev S = pd[srcSlot], D = pd[dstSlot];
  for (int k = threadIdx.x + blockIdx.x * blockDim.x; k < sx * sy;
       k += blockDim.x * gridDim.x) {
    const int x = k / sy, y = k - (k / sy) * sy;
    const long long si = tile_index(S, sox + x, soy + y);
    const long long di = tile_index(D, dox + x, doy + y);
    const int so = 1 - srcCur, dOld = 1 - dstCur;
    F.H(dOld)[di] = F.H(srcCur)[si] * tf + F.H(so)[si] * inv;
    F.HU(dOld)[di] = F.HU(srcCur)[si] * tf + F.HU(so)[si] * inv;
    F.HV(dOld)[di] = F.HV(srcCur)[si] * tf + F.HV(so)[si] * inv;
  }
}

// copyUNewToUOld (mode 0) / swapUNewAndUOld (mode 1), DefaultTransfer.cpph:103-127
__global__ void k_copy_swap(const PatchDev* __restrict__ pd, int slot, Fields F,
                            int cur, int mode) {
  const PatchDev P = pd[slot];
  for (int k = threadIdx.x + blockIdx.x * blockDim.x; k < P.nx * P.ny;
       k += blockDim.x * gridDim.x) {
    const int x = k / P.ny, y = k - (k / P.ny) * P.ny;
    const long long i = tile_index(P, x, y);
    double* a[3] = {F.H(cur), F.HU(cur), F.HV(cur)};
    double* o[3] = {F.H(1 - cur), F.HU(1 - cur), F.HV(1 - cur)};
    for (int q = 0; q < 3; ++q) {
      const double vn = a[q][i];
      if (mode == 1) a[q][i] = o[q][i];
      o[q][i] = vn;
    }
  }
}

// SWEKernel::transformWaterHeight (SWEKernel.cpp:21-48) over a region.
// uNew has no ghost layer in the reference's layout (q, x, y fastest, n^2
// per unknown), so a region cell (x, y) with x = nx or y = ny - reached by
// the grown source region of interpolateSolution (SWEKernel.cpp:252-262) -
// addresses element x*ny + y of that array: unknown q = idx / (nx ny) at
// (x', y') = ((idx mod nx ny) / ny, idx mod ny).  With unew_layout set the
// kernel writes that element, as the reference does; b is read at (x, y).
__device__ __forceinline__ void transform_cell(const PatchDev& P, const Fields& F, int st,
                                               int ox, int oy, int sy, double sign,
                                               int unew_layout, int k) {
  const int x = ox + k / sy, y = oy + (k - (k / sy) * sy);
  const long long ib = tile_index(P, x, y);
  double* f = F.H(st);
  long long i = ib;
  if (unew_layout && (x >= P.nx || y >= P.ny)) {
    const int lin = x * P.ny + y;
    const int area = P.nx * P.ny;
    const int q = lin / area, r = lin - q * area;
    f = q == 0 ? F.H(st) : (q == 1 ? F.HU(st) : F.HV(st));
    i = tile_index(P, r / P.ny, r - (r / P.ny) * P.ny);
  }
  f[i] = f[i] + sign * F.b[ib];
}

__global__ void k_transform(const PatchDev* __restrict__ pd, int slot, Fields F,
                            int st, int ox, int oy, int sx, int sy, double sign,
                            int unew_layout) {
  const PatchDev P = pd[slot];
  for (int k = threadIdx.x + blockIdx.x * blockDim.x; k < sx * sy;
       k += blockDim.x * gridDim.x)
    transform_cell(P, F, st, ox, oy, sy, sign, unew_layout, k);
}

// One region transform of a batch (the arguments of one k_transform launch).
struct TransformItem {
  int slot, st, ox, oy, sx, sy, unew_layout, pad;
  double sign;
};

// Batched k_transform: item blockIdx.y, its region cells over blockIdx.x.
// The items of one launch touch disjoint memory.
__global__ void k_transform_batch(const PatchDev* __restrict__ pd, Fields F,
                                  const TransformItem* __restrict__ items) {
  const TransformItem it = items[blockIdx.y];
  const PatchDev P = pd[it.slot];
  for (int k = threadIdx.x + blockIdx.x * blockDim.x; k < it.sx * it.sy;
       k += blockDim.x * gridDim.x)
    transform_cell(P, F, it.st, it.ox, it.oy, it.sy, it.sign, it.unew_layout, k);
}

struct InterpArgs {
  int srcSlot, dstSlot, srcCur, dstCur;
  int sx, sy, ox, oy;
  int toUOld;
  double tf;
  double spos[2], ssub[2], sinv[2], dpos[2], dsub[2];
  int sn[2];
};

// DefaultInterpolationTemplate::interpolateSolution
// (DefaultInterpolation.cpph:29-270): destination cell k of the region.
__device__ __forceinline__ void interp_cell(const PatchDev& S, const PatchDev& D,
                                            const Fields& F, const InterpArgs& a, int k) {
  const int x = a.ox + k / a.sy, y = a.oy + (k - (k / a.sy) * a.sy);
  const long long di = tile_index(D, x, y);
  const int dOld = 1 - a.dstCur;
  double* dq[3];
  if (a.toUOld) {
    dq[0] = F.H(dOld); dq[1] = F.HU(dOld); dq[2] = F.HV(dOld);
  } else {
    dq[0] = F.H(a.dstCur); dq[1] = F.HU(a.dstCur); dq[2] = F.HV(a.dstCur);
  }
  F.H(dOld)[di] = 0.0;
  F.HU(dOld)[di] = 0.0;
  F.HV(dOld)[di] = 0.0;
  double c[2];
  c[0] = a.dpos[0] + a.dsub[0] * (double)x + a.dsub[0] * 0.5;
  c[1] = a.dpos[1] + a.dsub[1] * (double)y + a.dsub[1] * 0.5;
  int si[2];
  for (int d = 0; d < 2; ++d) {
    const double base = a.spos[d] + a.ssub[d] * 0.5;
    const double loc = c[d] - base;
    int v = (int)floor(loc * a.sinv[d]);
    v = v < a.sn[d] - 2 ? v : a.sn[d] - 2;
    v = v > 0 ? v : 0;
    si[d] = v;
  }
  const double tf = a.tf;
  for (int o = 0; o < 4; ++o) {
    const int off[2] = {o & 1, (o >> 1) & 1};
    const int nb[2] = {si[0] + off[0], si[1] + off[1]};
    double np[2];
    np[0] = a.spos[0] + a.ssub[0] * (double)nb[0] + a.ssub[0] * 0.5;
    np[1] = a.spos[1] + a.ssub[1] * (double)nb[1] + a.ssub[1] * 0.5;
    double sf = 1.0;
    for (int d = 0; d < 2; ++d) {
      const double sign = (double)(1 - 2 * off[d]);
      sf *= 1.0 - (sign * (c[d] - np[d])) * a.sinv[d];
    }
    const long long si_ = tile_index(S, nb[0], nb[1]);
    const double* so[3] = {F.H(1 - a.srcCur), F.HU(1 - a.srcCur), F.HV(1 - a.srcCur)};
    const double* sn[3] = {F.H(a.srcCur), F.HU(a.srcCur), F.HV(a.srcCur)};
    for (int q = 0; q < 3; ++q) {
      const double vo = so[q][si_], vn = sn[q][si_];
      if (a.toUOld)
        dq[q][di] = dq[q][di] + sf * (vo * (1.0 - tf) + vn * tf);
      else
        dq[q][di] = dq[q][di] + sf * vo * (1.0 - tf) + sf * vn * tf;
    }
  }
}

// One thread per destination cell.
__global__ void k_interpolate(const PatchDev* __restrict__ pd, Fields F, InterpArgs a) {
  const PatchDev S = pd[a.srcSlot], D = pd[a.dstSlot];
  for (int k = threadIdx.x + blockIdx.x * blockDim.x; k < a.sx * a.sy;
       k += blockDim.x * gridDim.x)
    interp_cell(S, D, F, a, k);
}

// Batched k_interpolate: call blockIdx.y, its destination cells over
// blockIdx.x; the calls of one launch touch disjoint patches.
__global__ void k_interpolate_batch(const PatchDev* __restrict__ pd, Fields F,
                                    const InterpArgs* __restrict__ items) {
  const InterpArgs a = items[blockIdx.y];
  const PatchDev S = pd[a.srcSlot], D = pd[a.dstSlot];
  for (int k = threadIdx.x + blockIdx.x * blockDim.x; k < a.sx * a.sy;
       k += blockDim.x * gridDim.x)
    interp_cell(S, D, F, a, k);
}

struct RestrictArgs {
  int srcSlot, dstSlot, srcCur, dstCur;
  int rOff[2], rSize[2];  // region in source cells
  int dOff[2], dSize[2];  // mapped destination region
  double tfOld, tfNew;
  double srcSub[2], dstSub[2], dstRegion, eps;
  double spos[2], dpos[2];
};

// DefaultRestrictionTemplate::restrictRegion (DefaultRestriction.cpph:16-106):
// destination cell k of the mapped region, source cells in the reference's
// order.
__device__ __forceinline__ void restrict_cell(const PatchDev& S, const PatchDev& D,
                                              const Fields& F, const RestrictArgs& a, int k) {
  const int di0 = a.dOff[0] + k / a.dSize[1];
  const int di1 = a.dOff[1] + (k - (k / a.dSize[1]) * a.dSize[1]);
  double dp[2] = {a.dpos[0] + (double)di0 * a.dstSub[0],
                  a.dpos[1] + (double)di1 * a.dstSub[1]};
  int cOff[2], cSize[2];
  for (int d = 0; d < 2; ++d) {
    double ot = (dp[d] - a.spos[d] + a.eps);
    ot /= a.srcSub[d];
    const int o = (int)ot;
    double st = dp[d] + a.dstSub[d] - a.spos[d] - a.eps;
    st /= a.srcSub[d];
    st -= ot;
    st += 1.0;
    const int s = (int)st;
    const int upper = o + s, rUpper = a.rOff[d] + a.rSize[d];
    cOff[d] = o > a.rOff[d] ? o : a.rOff[d];
    cSize[d] = (upper < rUpper ? upper : rUpper) - cOff[d];
  }
  const long long di = tile_index(D, di0, di1);
  const int dOldSt = 1 - a.dstCur, sOldSt = 1 - a.srcCur;
  double accO[3] = {F.H(dOldSt)[di], F.HU(dOldSt)[di], F.HV(dOldSt)[di]};
  double accN[3] = {F.H(a.dstCur)[di], F.HU(a.dstCur)[di], F.HV(a.dstCur)[di]};
  for (int sx = 0; sx < cSize[0]; ++sx)
    for (int sy = 0; sy < cSize[1]; ++sy) {
      const int s0 = cOff[0] + sx, s1 = cOff[1] + sy;
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
      const double vo[3] = {F.H(sOldSt)[si], F.HU(sOldSt)[si], F.HV(sOldSt)[si]};
      const double vn[3] = {F.H(a.srcCur)[si], F.HU(a.srcCur)[si], F.HV(a.srcCur)[si]};
      for (int q = 0; q < 3; ++q) {
        const double dO = (vo[q] * (1.0 - a.tfOld) + vn[q] * a.tfOld) * region / a.dstRegion;
        const double dN = (vo[q] * (1.0 - a.tfNew) + vn[q] * a.tfNew) * region / a.dstRegion;
        accO[q] = accO[q] + dO;
        accN[q] = accN[q] + dN;
      }
    }
  F.H(dOldSt)[di] = accO[0]; F.HU(dOldSt)[di] = accO[1]; F.HV(dOldSt)[di] = accO[2];
  F.H(a.dstCur)[di] = accN[0]; F.HU(a.dstCur)[di] = accN[1]; F.HV(a.dstCur)[di] = accN[2];
}

// One thread per destination cell.
__global__ void k_restrict(const PatchDev* __restrict__ pd, Fields F, RestrictArgs a) {
  const PatchDev S = pd[a.srcSlot], D = pd[a.dstSlot];
  const int n = a.dSize[0] * a.dSize[1];
  for (int k = threadIdx.x + blockIdx.x * blockDim.x; k < n; k += blockDim.x * gridDim.x)
    restrict_cell(S, D, F, a, k);
}

// Batched k_restrict: region blockIdx.y, its destination cells over
// blockIdx.x; the regions of one launch write disjoint destination patches.
__global__ void k_restrict_batch(const PatchDev* __restrict__ pd, Fields F,
                                 const RestrictArgs* __restrict__ items) {
  const RestrictArgs a = items[blockIdx.y];
  const PatchDev S = pd[a.srcSlot], D = pd[a.dstSlot];
  const int n = a.dSize[0] * a.dSize[1];
  for (int k = threadIdx.x + blockIdx.x * blockDim.x; k < n; k += blockDim.x * gridDim.x)
    restrict_cell(S, D, F, a, k);
}

// Patch::switchToLeaf (Patch.cpp:793-830) of a batch of patches (patch
// blockIdx.y): the [0, 1]-basis values of the restriction evaluated at the
// leaf's interval, uOld = vo (1 - fOld) + vn fOld and uNew = vo (1 - fNew)
// + vn fNew over the interior cells.
struct LeafSwitch {
  int slot, cur;
  double fOld, fNew;
};
__global__ void k_switch_to_leaf(const PatchDev* __restrict__ pd, Fields F,
                                 const LeafSwitch* __restrict__ items) {
  const LeafSwitch it = items[blockIdx.y];
  const PatchDev P = pd[it.slot];
  for (int k = threadIdx.x + blockIdx.x * blockDim.x; k < P.nx * P.ny;
       k += blockDim.x * gridDim.x) {
    const long long i = tile_index(P, k / P.ny, k - (k / P.ny) * P.ny);
    double* o[3] = {F.H(1 - it.cur), F.HU(1 - it.cur), F.HV(1 - it.cur)};
    double* n[3] = {F.H(it.cur), F.HU(it.cur), F.HV(it.cur)};
    for (int q = 0; q < 3; ++q) {
      const double vo = o[q][i], vn = n[q][i];
      o[q][i] = vo * (1.0 - it.fOld) + vn * it.fOld;
      n[q][i] = vo * (1.0 - it.fNew) + vn * it.fNew;
    }
  }
}

// Zero uNew's interior and all of uOld (ghost ring included) of a batch of
// patches (patch blockIdx.y, state `cur[blockIdx.y]` current); b is kept.
__global__ void k_clear_patches(const PatchDev* __restrict__ pd, Fields F,
                                const int* __restrict__ slots, const int* __restrict__ curs) {
  const int slot = slots[blockIdx.y], cur = curs[blockIdx.y];
  const PatchDev P = pd[slot];
  const int w = P.stride, area = (P.nx + 2) * w;  // row padding included
  for (int k = threadIdx.x + blockIdx.x * blockDim.x; k < area; k += blockDim.x * gridDim.x) {
    const long long i = P.off + k;
    F.H(1 - cur)[i] = 0.0;
    F.HU(1 - cur)[i] = 0.0;
    F.HV(1 - cur)[i] = 0.0;
    const int x = k / w, y = k - x * w;
    if (x >= 1 && x <= P.nx && y >= 1 && y <= P.ny) {
      F.H(cur)[i] = 0.0;
      F.HU(cur)[i] = 0.0;
      F.HV(cur)[i] = 0.0;
    }
  }
}

// b of a batch of tiles from a staging array: tile blockIdx.y takes
// (nx+2)(ny+2) values (rows of ny+2) from src + soff[blockIdx.y].
__global__ void k_scatter_b(const PatchDev* __restrict__ pd, Fields F,
                            const int* __restrict__ slots, const long long* __restrict__ soff,
                            const double* __restrict__ src) {
  const PatchDev P = pd[slots[blockIdx.y]];
  const double* s = src + soff[blockIdx.y];
  const int w = P.ny + 2, area = (P.nx + 2) * w;
  for (int k = threadIdx.x + blockIdx.x * blockDim.x; k < area; k += blockDim.x * gridDim.x) {
    const int x = k / w, y = k - x * w;
    F.b[P.off + (long long)x * P.stride + y] = s[k];
  }
}

// BreakingDamSWEScenario::computeDemandedMeshWidth (BreakingDam.cpp:85-139)
// and SolveTimestep::leaveCell's decision (SolveTimestep.cpp:765-786):
// one workgroup of kRefineThreads per patch over uNew (a maximum, so the
// order of the reduction does not change the result).
constexpr int kRefineThreads = 256;
__global__ void __launch_bounds__(kRefineThreads)
k_refine_flags(const int* __restrict__ slots, int nslots, const PatchDev* __restrict__ pd,
               Fields F, const ScnDev* __restrict__ scn, double threshold,
               double minW, double maxW, int* __restrict__ flags,
               double* __restrict__ demanded, double* __restrict__ done = nullptr,
               unsigned long long seq = 0) {
  const int k = blockIdx.x;
  if (k >= nslots) return;
  const int slot = slots[k];
  const PatchDev P = pd[slot];
  const int cur = scn[P.scn].stepc & 1;
  const double* H = F.H(cur);
  __shared__ double wave_max[kRefineThreads / 64];
  double g = 0.0;
  const int nxy = (P.nx - 1) * (P.ny - 1);
  for (int e = threadIdx.x; e < nxy; e += kRefineThreads) {
    const int xi = e / (P.ny - 1), yi = e - (e / (P.ny - 1)) * (P.ny - 1);
    const double q0 = H[tile_index(P, xi, yi)];
    const double q0x = (H[tile_index(P, xi + 1, yi)] - q0) / P.dx;
    const double q0y = (H[tile_index(P, xi, yi + 1)] - q0) / P.dy;
    g = fmax(g, sqrt(fabs(q0x) * fabs(q0x) + fabs(q0y) * fabs(q0y)));
  }
  for (int m = 32; m >= 1; m >>= 1) g = fmax(g, __shfl_xor(g, m));
  if ((threadIdx.x & 63) == 0) wave_max[threadIdx.x >> 6] = g;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kRefineThreads / 64; ++w) g = fmax(g, wave_max[w]);
    double dw0, dw1;
    if (fabs(minW - maxW) <= kLaEps) {
      dw0 = dw1 = minW;
    } else if (g > threshold) {
      dw0 = dw1 = minW;
    } else if (g < threshold) {
      dw0 = dw1 = maxW;
    } else {
      dw0 = P.dx; dw1 = P.dy;
    }
    int f;
    if (P.dx - dw0 > kLaEps || P.dy - dw1 > kLaEps) {
      f = 1;
    } else if (!((P.dx * 3.0) - dw0 > kLaEps || (P.dy * 3.0) - dw1 > kLaEps)) {
      f = 0;
    } else {
      f = 2;
    }
    flags[k] = f;
    demanded[2 * k] = dw0;
    demanded[2 * k + 1] = dw1;
    if (done) signal_done(done, seq);  // a single-patch call (one workgroup)
  }
}

// Per-patch sum of h over the interior of the current state.
__global__ void __launch_bounds__(64)
k_patch_sums(const int* __restrict__ slots, int nslots, const PatchDev* __restrict__ pd,
             Fields F, const ScnDev* __restrict__ scn, double* __restrict__ out) {
  const int k = blockIdx.x;
  if (k >= nslots) return;
  const PatchDev P = pd[slots[k]];
  const int cur = scn[P.scn].stepc & 1;
  double s = 0.0, comp = 0.0;
  for (int e = threadIdx.x; e < P.nx * P.ny; e += 64) {
    const double v = F.H(cur)[tile_index(P, e / P.ny, e % P.ny)];
    const double t = s + v;
    comp += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
    s = t;
  }
  s += comp;
  for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
  if (threadIdx.x == 0) out[k] = s;
}

__global__ void k_reset_counter(int* c) { *c = 0; }

// The edge solver on a batch of independent edges (diagnostics/tests).
template <int SOLVER>
__global__ void k_edge_batch(long long n, const double* __restrict__ in,
                             double* __restrict__ out) {
  for (long long e = threadIdx.x + (long long)blockIdx.x * blockDim.x; e < n;
       e += (long long)blockDim.x * gridDim.x) {
    const double* v = in + 6 * e;
    double dhL, dhR, dqL, dqR, sp;
    if (SOLVER == 0) {
      const CellQ L = cell_dir(cell_base(v[0], v[4]), v[2]);
      const CellQ R = cell_dir(cell_base(v[1], v[5]), v[3]);
      edge_solve(L, R, dhL, dhR, dqL, dqR);
      sp = edge_speed(L, R);
    } else {
      aug_solve(v[0], v[1], v[2], v[3], v[4], v[5], dhL, dhR, dqL, dqR);
      sp = aug_speed(v[0], v[1], v[2], v[3], v[4], v[5]);
    }
    double* o = out + 5 * e;
    o[0] = dhL; o[1] = dhR; o[2] = dqL; o[3] = dqR;
    o[4] = sp;
  }
}

// ---- flux correction (DefaultFluxCorrection; off by default in the reference) ----
//
// Per slot, the reference's flux array (LinearizationZYXQ.h, flux part):
// faces (x lower, x upper, y lower, y upper), each 3 unknowns x the cells
// along the face, unknown slowest; kFluxPerSlot doubles reserved per slot.
constexpr int kFluxPerSlot = 2 * 3 * (64 + 64);

__device__ __forceinline__ int flux_face_offset(int nx, int ny, int dim, int direction) {
  const int o1 = ny * 3, o2 = o1 + ny * 3, o3 = o2 + nx * 3;
  const int f = 2 * dim + (1 + direction) / 2;
  return f == 0 ? 0 : f == 1 ? o1 : f == 2 ? o2 : o3;
}

// Numerics::computeFlux (Numerics.h:185-203) as the SWE build compiles it:
// the unknowns times the normal momentum, the momenta times h once more,
// and 0.5 * g * h * h with g = 1.0 added to the normal momentum.
__device__ __forceinline__ void ref_face_flux(int dim, const double u[3], double f[3]) {
  const double un = dim == 0 ? u[1] : u[2];
  f[0] = u[0] * un;
  f[1] = u[1] * un;
  f[2] = u[2] * un;
  f[1] *= u[0];
  f[2] *= u[0];
  const double g = 1.0;
  const double pr = 0.5 * g * u[0] * u[0];
  if (dim == 0)
    f[1] += pr;
  else
    f[2] += pr;
}

// DefaultFluxCorrectionTemplate::computeFluxes (DefaultFluxCorrection.cpph:
// 44-86): one thread per face cell, 0.5 * (F(ghost uOld) + F(interior uOld));
// the upwind term of :71-72 is a separate statement the reference discards.
// The interior / ghost cells of an upper face sit at ny - 1 / ny, as the
// reference's face iterator reads getSubdivisionFactor(+1) (square patches).
__global__ void __launch_bounds__(64)
k_compute_fluxes(PatchDev P, Fields F, int old_state, double* __restrict__ flux) {
  const int nx = P.nx, ny = P.ny;
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * (nx + ny)) return;
  int dim, dir;
  if (e < ny) {
    dim = 0; dir = -1;
  } else if (e < 2 * ny) {
    dim = 0; dir = 1; e -= ny;
  } else if (e < 2 * ny + nx) {
    dim = 1; dir = -1; e -= 2 * ny;
  } else {
    dim = 1; dir = 1; e -= 2 * ny + nx;
  }
  const int nf = dim == 0 ? ny : nx;
  const int ni = dir == -1 ? 0 : ny - 1, ng = dir == -1 ? -1 : ny;
  const long long ii = dim == 0 ? tile_index(P, ni, e) : tile_index(P, e, ni);
  const long long ig = dim == 0 ? tile_index(P, ng, e) : tile_index(P, e, ng);
  const double inner[3] = {F.H(old_state)[ii], F.HU(old_state)[ii], F.HV(old_state)[ii]};
  const double outer[3] = {F.H(old_state)[ig], F.HU(old_state)[ig], F.HV(old_state)[ig]};
  double fi[3], fg[3];
  ref_face_flux(dim, outer, fg);
  ref_face_flux(dim, inner, fi);
  const int fo = flux_face_offset(nx, ny, dim, dir);
  for (int q = 0; q < 3; ++q) flux[fo + q * nf + e] = 0.5 * (fg[q] + fi[q]);
}

// What applyCorrection reads of a patch besides its tile.
struct FluxGeom {
  double pos[2], sub[2];
  int nx, ny;
};

__device__ __forceinline__ double fmax_std(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double fmin_std(double a, double b) { return (b < a) ? b : a; }

// DefaultFluxCorrectionTemplate::applyCorrection (DefaultFluxCorrection.cpph:
// 88-164) with HyperplaneRegion::getInterfaceRegionForSubcell
// (HyperplaneRegion.cpp:67-103) and calculateOverlappingRegion (:13-41), as
// the reference computes them: the destination region of a source cell is
// taken against the destination's first tangential subcell (the subcell
// size is used as the patch size), and the destination flux is read from
// the start of its face whatever the region's offset.  Thread q updates
// unknown q in the reference's order (source cells, then region cells).
__global__ void __launch_bounds__(64)
k_apply_flux_correction(FluxGeom S, PatchDev SP, const double* __restrict__ sflux, FluxGeom D,
                        PatchDev DP, const double* __restrict__ dflux, Fields F, int dst_cur,
                        int dimension, int direction, int rdim, double time_overlap) {
  const int q = threadIdx.x;
  if (q >= 3) return;
  const int sd = rdim == 0 ? 1 : 0;
  const double eps = 1e-10;
  const double dvol = D.sub[0] * D.sub[1];
  const int nfs = dimension == 0 ? S.ny : S.nx, nfd = dimension == 0 ? D.ny : D.nx;
  const int fos = flux_face_offset(S.nx, S.ny, dimension, direction);
  const int fod = flux_face_offset(D.nx, D.ny, dimension, -direction);
  const int ni = -direction == -1 ? 0 : D.ny - 1;
  double* const U = q == 0 ? F.H(dst_cur) : q == 1 ? F.HU(dst_cur) : F.HV(dst_cur);
  for (int ks = 0; ks < nfs; ++ks) {
    const double sF = sflux[fos + q * nfs + ks];
    const double pos = D.pos[sd], size = D.sub[sd], sub = D.sub[sd];
    double npos = S.pos[sd];
    npos += S.sub[sd] * (double)ks;
    const double lo = fmax_std(pos, npos);
    const double hi = fmin_std(pos + size, npos + S.sub[sd]);
    const int loC = (int)((lo - pos + eps) / sub);
    const int hiC = (int)ceil((hi - pos - eps) / sub);
    for (int j = 0; j < hiC - loC; ++j) {
      const int t = loC + j;
      if (t < 0 || t >= nfd) continue;
      const double dF = dflux[fod + q * nfd + j];
      // calculateOverlappingRegion over the axis other than `dimension`
      const int od = dimension == 0 ? 1 : 0;
      double region = 1.0;
      const double iv = fmin_std(S.pos[od] + ((double)ks + 1.0) * S.sub[od],
                                 D.pos[od] + ((double)t + 1.0) * D.sub[od]) -
                        fmax_std(S.pos[od] + (double)ks * S.sub[od], D.pos[od] + (double)t * D.sub[od]);
      region *= iv;
      if (region < 0.0) region = 0.0;
      const long long i = dimension == 0 ? tile_index(DP, ni, t) : tile_index(DP, t, ni);
      U[i] = U[i] + (sF - dF) * time_overlap * region / dvol;
    }
  }
  (void)SP;
}

// ---- served requests ----

constexpr int kMaxHills = 8;

// Initial state of one dam-break request on [0,1]^2, as
// BreakingDamSWEScenario::initializePatch / getWaterHeight
// (BreakingDam.cpp:53-83, 165-176) with the request's centre, radius and
// heights: at the cell centre X = position + (i + 1/2) * subcell size,
// r = sqrt((X - cx)^2 + (Y - cy)^2), h = column height if r <= radius else
// `depth`; hu = hv = 0; b = sum of Gaussian hills
// a_k exp(-|x - c_k|^2 / s2_k) (0 without hills), evaluated at every cell
// centre of the tile including the ghost ring.
struct RequestInit {
  double dx;  // request grid width 1 / cells_per_side (patch positions)
  double cx, cy, radius, col_h, depth;
  int nh, P, n, cur;
  double hx[kMaxHills], hy[kMaxHills], hs2[kMaxHills], ha[kMaxHills];
};

// One workgroup per patch; slots[k] holds grid patch (k / P, k % P).
__global__ void __launch_bounds__(256)
k_init_request(const int* __restrict__ slots, int nslots, const PatchDev* __restrict__ pd,
               Fields F, RequestInit a) {
  const int k = blockIdx.x;
  if (k >= nslots) return;
  const PatchDev P = pd[slots[k]];
  const int pi = k / a.P, pj = k - (k / a.P) * a.P;
  // the patch's position as its description holds it (i * n * dx)
  const double px = (double)(pi * a.n) * a.dx, py = (double)(pj * a.n) * a.dx;
  const int s = a.n + 2;
  for (int c = threadIdx.x; c < s * s; c += blockDim.x) {
    const int x = c / s, y = c - (c / s) * s;
    const double X = px + ((double)(x - 1) + 0.5) * P.dx;
    const double Y = py + ((double)(y - 1) + 0.5) * P.dy;
    double b = 0.0;
    for (int m = 0; m < a.nh; ++m) {
      const double ex = X - a.hx[m], ey = Y - a.hy[m];
      b += a.ha[m] * exp(-(ex * ex + ey * ey) / a.hs2[m]);
    }
    const double ex = X - a.cx, ey = Y - a.cy;
    const double r = sqrt(ex * ex + ey * ey);
    const long long i = P.off + (long long)x * P.stride + y;
    F.b[i] = b;
    F.H(a.cur)[i] = a.col_h * (double)(r <= a.radius) + a.depth * (double)(r > a.radius);
    F.HU(a.cur)[i] = 0.0;
    F.HV(a.cur)[i] = 0.0;
  }
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Order-independent 64-bit digest of the interior bits of h, hu, hv of a
// grid of patches (slots in grid order): sum over cells of a mixed hash of
// (cell index, bits of h, hu, hv).  Equal states give equal digests; any
// changed bit changes the digest except with probability ~2^-64.
__global__ void __launch_bounds__(256)
k_hash_request(const int* __restrict__ slots, int nslots, const PatchDev* __restrict__ pd,
               Fields F, int cur, unsigned long long* __restrict__ out) {
  __shared__ unsigned long long red[256];
  const int k = blockIdx.x;
  if (k >= nslots) return;
  const PatchDev P = pd[slots[k]];
  const int nx = P.nx, ny = P.ny;
  unsigned long long acc = 0;
  for (int c = threadIdx.x; c < nx * ny; c += blockDim.x) {
    const int x = c / ny, y = c - (c / ny) * ny;
    const long long i = tile_index(P, x, y);
    unsigned long long v = mix64((unsigned long long)k * (unsigned long long)(nx * ny) + c);
    v = mix64(v ^ (unsigned long long)__double_as_longlong(F.H(cur)[i]));
    v = mix64(v ^ (unsigned long long)__double_as_longlong(F.HU(cur)[i]));
    v = mix64(v ^ (unsigned long long)__double_as_longlong(F.HV(cur)[i]));
    acc += v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int m = 128; m > 0; m >>= 1) {
    if ((int)threadIdx.x < m) red[threadIdx.x] += red[threadIdx.x + m];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

// Plain device copy, one 16-byte element per thread (no grid-stride loop):
// the bandwidth probe whose byte count is known exactly, n2 = number of
// double2 elements.
__global__ void __launch_bounds__(256) k_copy_probe(long long n2,
                                                    const double2* __restrict__ src,
                                                    double2* __restrict__ dst) {
  const long long i = threadIdx.x + (long long)blockIdx.x * blockDim.x;
  if (i < n2) dst[i] = src[i];
}

// Flat stream with the update's traffic mix per element: 4 reads (h, hu, hv,
// b), 3 writes (h, hu, hv); one pair of elements per thread in 16-byte lanes
// (the last thread takes an odd element alone).
__global__ void __launch_bounds__(256) k_stream_probe(long long n, const double* __restrict__ h,
                                                      const double* __restrict__ hu,
                                                      const double* __restrict__ hv,
                                                      const double* __restrict__ b,
                                                      double* __restrict__ oh,
                                                      double* __restrict__ ohu,
                                                      double* __restrict__ ohv) {
  const long long i = threadIdx.x + (long long)blockIdx.x * blockDim.x;
  const long long n2 = n / 2;
  if (i < n2) {
    const double2 x = reinterpret_cast<const double2*>(h)[i];
    const double2 y = reinterpret_cast<const double2*>(hu)[i];
    const double2 z = reinterpret_cast<const double2*>(hv)[i];
    const double2 w = reinterpret_cast<const double2*>(b)[i];
    double2 a, c, d;
    a.x = x.x + 0.5 * w.x;
    a.y = x.y + 0.5 * w.y;
    c.x = y.x * a.x;
    c.y = y.y * a.y;
    d.x = z.x * a.x;
    d.y = z.y * a.y;
    reinterpret_cast<double2*>(oh)[i] = a;
    reinterpret_cast<double2*>(ohu)[i] = c;
    reinterpret_cast<double2*>(ohv)[i] = d;
  } else if (i == n2 && (n & 1)) {
    const long long k = n - 1;
    const double x = h[k] + 0.5 * b[k];
    oh[k] = x;
    ohu[k] = hu[k] * x;
    ohv[k] = hv[k] * x;
  }
}

}  // namespace pcamd
