// Device-side arithmetic of the SWE patch update for gfx950 (FP64).
//
// The f-wave edge solver and the block accumulation follow the reference's
// call sites (tools/SWE_WaveAccumulationBlock.cpp:142-214) and its update
// (:247-296).  All arithmetic is FP64, compiled with -ffp-contract=off, so
// every +,-,*,/ and sqrt is one correctly rounded operation in the order
// written here; the CPU restatement performs the same operations in the
// same order, which makes results bitwise reproducible between the two.
//
// Work per edge is reduced by computing the per-cell terms once per cell
// (u = q/h, sqrt(h), sqrt(g h), q u + (g/2 h) h) instead of once per edge
// side; these are pure functions of the cell's state, so reusing them does
// not change any result bit.
#pragma once

#include <hip/hip_runtime.h>

namespace pcamd {

constexpr double kGravity = 9.81;
constexpr double kHalfG = 0.5 * 9.81;
constexpr double kDryTol = 0.01;   // f-wave dry tolerance
constexpr double kZeroTol = 1e-7;  // zero wave-speed tolerance
constexpr double kCfl = 0.4;       // SWE_WaveAccumulationBlock.cpp:236
constexpr double kMinSpeed = 0.00001;  // :225
constexpr double kClampH = 0.1;    // :270 velocity clamp depth
constexpr double kLaEps = 1e-8;    // tarch::la comparison tolerance

// One cell as seen by one sweep direction (q = normal momentum).
struct CellQ {
  double h, q, b;
  double u, sq, c, f;
};

// FP64 square root and quotient without the exponent-scaling steps of the
// general IEEE sequences.  sqrt_nr: hardware 1/sqrt estimate refined by one
// Goldschmidt step and two Newton corrections on the remainder x - g*g;
// div_nr: hardware reciprocal refined by two Newton steps, quotient
// corrected once with the remainder a - b*q (Markstein's correction),
// special operands (zeros, infinities, NaN) resolved by v_div_fixup.  Both
// return the correctly rounded result (bit-identical to sqrt() and a / b)
// whenever the scaling would be the identity: x in [2^-767, DBL_MAX] for
// sqrt_nr; for div_nr b and a / b normal and well inside the exponent range,
// a zero or normal.  Every use below stays in that range for a wet cell
// (h >= 0.01; momenta, speeds and Roe averages of moderate size); the terms
// of dry cells fall outside it (h = 0) and are replaced by the dry
// substitution before anything reads them.
__device__ __forceinline__ double sqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double hh = y * 0.5;
  const double r = fma(-hh, g, 0.5);
  g = fma(g, r, g);
  hh = fma(hh, r, hh);
  double d = fma(-g, g, x);
  g = fma(d, hh, g);
  d = fma(-g, g, x);
  g = fma(d, hh, g);
  return g;
}

__device__ __forceinline__ double div_nr(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  const double q = a * r;
  const double rem = fma(-b, q, a);
  return __builtin_amdgcn_div_fixup(fma(rem, r, q), b, a);
}

__device__ __forceinline__ double dmin(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return a < b ? b : a; }

// Per-cell terms shared by both sweeps.
struct CellBase {
  double h, b, sq, c, p;  // p = (g/2 h) h
};

__device__ __forceinline__ CellBase cell_base(double h, double b) {
  CellBase r;
  r.h = h;
  r.b = b;
  r.sq = sqrt_nr(h);
  r.c = sqrt_nr(kGravity * h);
  r.p = (kHalfG * h) * h;
  return r;
}

__device__ __forceinline__ CellQ cell_dir(const CellBase& cb, double q) {
  CellQ r;
  r.h = cb.h;
  r.q = q;
  r.b = cb.b;
  r.u = div_nr(q, cb.h);
  r.sq = cb.sq;
  r.c = cb.c;
  r.f = q * r.u + cb.p;
  return r;
}

// Replace a dry side by the mirror image of the wet side (wall).
__device__ __forceinline__ void apply_dry(CellQ& L, CellQ& R, bool& leftDry,
                                          bool& rightDry, bool& bothDry) {
  const bool lw = L.h >= kDryTol;
  const bool rw = R.h >= kDryTol;
  rightDry = lw && !rw;
  leftDry = !lw && rw;
  bothDry = !lw && !rw;
  if (rightDry) {
    R.h = L.h; R.q = -L.q; R.b = L.b; R.u = -L.u; R.sq = L.sq; R.c = L.c; R.f = L.f;
  }
  if (leftDry) {
    L.h = R.h; L.q = -R.q; L.b = R.b; L.u = -R.u; L.sq = R.sq; L.c = R.c; L.f = R.f;
  }
}

__device__ __forceinline__ void wave_speeds(const CellQ& L, const CellQ& R,
                                            double& s1, double& s2) {
  const double hRoe = 0.5 * (L.h + R.h);
  const double uRoe = div_nr(L.u * L.sq + R.u * R.sq, L.sq + R.sq);
  const double cRoe = sqrt_nr(kGravity * hRoe);
  s1 = dmin(L.u - L.c, uRoe - cRoe);
  s2 = dmax(R.u + R.c, uRoe + cRoe);
}

// Edge speed only: max(|s1|, |s2|), 0 when both sides are dry.
__device__ __forceinline__ double edge_speed(CellQ L, CellQ R) {
  bool ld, rd, bd;
  apply_dry(L, R, ld, rd, bd);
  double s1, s2;
  wave_speeds(L, R, s1, s2);
  const double s = dmax(fabs(s1), fabs(s2));
  return bd ? 0.0 : s;
}

// Full f-wave edge solve: net updates of the left and right cell.
__device__ __forceinline__ void edge_solve(CellQ L, CellQ R, double& dhL,
                                           double& dhR, double& dqL,
                                           double& dqR) {
  bool leftDry, rightDry, bothDry;
  apply_dry(L, R, leftDry, rightDry, bothDry);
  double s1, s2;
  wave_speeds(L, R, s1, s2);
  const double d0 = R.q - L.q;
  const double d1 = (R.f - L.f) + (kHalfG * (L.h + R.h)) * (R.b - L.b);
  const double inv = div_nr(1.0, s2 - s1);
  const double beta1 = (s2 * d0 - d1) * inv;
  const double beta2 = (d1 - s1 * d0) * inv;
  const double z1 = beta1 * s1, z2 = beta2 * s2;
  double a = 0.0, b = 0.0, c = 0.0, d = 0.0;  // dhL, dhR, dqL, dqR
  if (s1 < -kZeroTol) {
    a += beta1; c += z1;
  } else if (s1 > kZeroTol) {
    b += beta1; d += z1;
  } else {
    a += 0.5 * beta1; c += 0.5 * z1;
    b += 0.5 * beta1; d += 0.5 * z1;
  }
  if (s2 < -kZeroTol) {
    a += beta2; c += z2;
  } else if (s2 > kZeroTol) {
    b += beta2; d += z2;
  } else {
    a += 0.5 * beta2; c += 0.5 * z2;
    b += 0.5 * beta2; d += 0.5 * z2;
  }
  if (leftDry || bothDry) { a = 0.0; c = 0.0; }
  if (rightDry || bothDry) { b = 0.0; d = 0.0; }
  dhL = a; dhR = b; dqL = c; dqR = d;
}

// Wave-uniform fast paths.  When every lane of the wavefront has two wet
// cells (h >= dryTol) the dry substitution is the identity, and when every
// lane's waves are subsonic (s1 < -zeroTol, s2 > zeroTol) the distribution
// is "wave 1 left, wave 2 right".  The operations that remain are exactly
// those the general path performs in that case, so results are unchanged
// bit for bit; only the selects of the unused branches are skipped.
__device__ __forceinline__ bool wave_all(bool c) { return __all(c ? 1 : 0) != 0; }

__device__ __forceinline__ double edge_speed_wet(const CellQ& L, const CellQ& R) {
  double s1, s2;
  wave_speeds(L, R, s1, s2);
  return dmax(fabs(s1), fabs(s2));
}

__device__ __forceinline__ void edge_solve_wet(const CellQ& L, const CellQ& R, double& dhL,
                                               double& dhR, double& dqL, double& dqR) {
  double s1, s2;
  wave_speeds(L, R, s1, s2);
  const double d0 = R.q - L.q;
  const double d1 = (R.f - L.f) + (kHalfG * (L.h + R.h)) * (R.b - L.b);
  const double inv = div_nr(1.0, s2 - s1);
  const double beta1 = (s2 * d0 - d1) * inv;
  const double beta2 = (d1 - s1 * d0) * inv;
  const double z1 = beta1 * s1, z2 = beta2 * s2;
  if (wave_all(s1 < -kZeroTol && s2 > kZeroTol)) {
    dhL = 0.0 + beta1; dqL = 0.0 + z1;
    dhR = 0.0 + beta2; dqR = 0.0 + z2;
    return;
  }
  double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
  if (s1 < -kZeroTol) {
    a += beta1; c += z1;
  } else if (s1 > kZeroTol) {
    b += beta1; d += z1;
  } else {
    a += 0.5 * beta1; c += 0.5 * z1;
    b += 0.5 * beta1; d += 0.5 * z1;
  }
  if (s2 < -kZeroTol) {
    a += beta2; c += z2;
  } else if (s2 > kZeroTol) {
    b += beta2; d += z2;
  } else {
    a += 0.5 * beta2; c += 0.5 * z2;
    b += 0.5 * beta2; d += 0.5 * z2;
  }
  dhL = a; dhR = b; dqL = c; dqR = d;
}

// Dispatch on wave-wide wetness (all lanes must call it together).
__device__ __forceinline__ void edge_solve_any(const CellQ& L, const CellQ& R, double& dhL,
                                               double& dhR, double& dqL, double& dqR) {
  if (wave_all(L.h >= kDryTol && R.h >= kDryTol))
    edge_solve_wet(L, R, dhL, dhR, dqL, dqR);
  else
    edge_solve(L, R, dhL, dhR, dqL, dqR);
}

__device__ __forceinline__ double edge_speed_any(const CellQ& L, const CellQ& R) {
  if (wave_all(L.h >= kDryTol && R.h >= kDryTol)) return edge_speed_wet(L, R);
  return edge_speed(L, R);
}

// ---- cheap upper bounds of edge speeds (pruned CFL pass) ----
//
// For the Einfeldt speeds above, |s1|, |s2| <= max(|uL|,|uR|) + max(cL,cR):
// the Roe velocity is a convex combination of uL and uR, and
// sqrt(g (hL+hR)/2) <= max(cL, cR).  A dry side is replaced by the mirror
// of the wet side (same |u| and c), and two dry sides give speed 0, so a dry
// cell may contribute (0, 0).  The terms are evaluated in FP32 with the
// hardware reciprocal and square root (relative error below 1e-6); callers
// compare with a margin of 2^-10, far above that error.
struct BoundTerms {
  float ax, ay, c;  // |hu/h|, |hv/h|, sqrt(g h); all 0 for a dry cell
};

__device__ __forceinline__ BoundTerms bound_terms(double h, double hu, double hv) {
  BoundTerms t;
  const float hf = (float)h;
  const float inv = __builtin_amdgcn_rcpf(hf);
  const bool wet = h >= kDryTol;
  t.ax = wet ? fabsf((float)hu) * inv : 0.0f;
  t.ay = wet ? fabsf((float)hv) * inv : 0.0f;
  t.c = wet ? __builtin_amdgcn_sqrtf(9.81f * hf) : 0.0f;
  return t;
}

constexpr float kBoundMargin = 1.0f + 1.0f / 1024.0f;

// True unless the bound proves the edge slower than `thr` (NaN -> true).
__device__ __forceinline__ bool bound_candidate(float aL, float cL, float aR, float cR,
                                                float thr) {
  const float b = (fmaxf(aL, aR) + fmaxf(cL, cR)) * kBoundMargin;
  return !(b < thr);
}

// SWE_WaveAccumulationBlock::computeNumericalFluxes (:225-239): CFL step of
// one patch from its maximum edge speed (DBL_MAX where the float reference
// uses FLT_MAX).
__device__ __host__ __forceinline__ double max_timestep(double s, double dx,
                                                        double dy) {
  if (s > kMinSpeed) {
    double a = dx / s, b = dy / s;
    double t = b < a ? b : a;
    t *= kCfl;
    return t;
  }
  return 1.7976931348623157e308;
}

}  // namespace pcamd
