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
// is "wave 1 left, wave 2 right".  The operations that remain are those the
// general path performs in that case (with the exceptions noted where they
// occur, which cannot change a cell's result), so results are unchanged bit
// for bit; only the selects of the unused branches are skipped.
// Wave votes over the active lanes, read straight from the compare mask
// (no per-lane 0/1 value and second compare, as __all / __any make).
__device__ __forceinline__ bool wave_all(bool c) {
  return __builtin_amdgcn_ballot_w64(c) == __builtin_amdgcn_read_exec();
}
__device__ __forceinline__ bool wave_any(bool c) { return __builtin_amdgcn_ballot_w64(c) != 0; }
// Both conditions on every active lane outside the mask `ign` (one mask per
// compare, combined as masks; `ign` holds active lanes only).
__device__ __forceinline__ bool wave_all2(bool c0, bool c1, unsigned long long ign = 0) {
  return ((__builtin_amdgcn_ballot_w64(c0) & __builtin_amdgcn_ballot_w64(c1)) | ign) ==
         __builtin_amdgcn_read_exec();
}

// wave_speeds for two wet cells with the hardware min / max: they differ
// from the selects of wave_speeds only for a NaN or for +0 against -0, and
// neither occurs here (c >= sqrt(g dryTol) > 0, so u - c and uRoe - cRoe
// are zero only as x - x = +0, and u + c likewise).
__device__ __forceinline__ void wave_speeds_wet(const CellQ& L, const CellQ& R,
                                                double& s1, double& s2) {
  const double hRoe = 0.5 * (L.h + R.h);
  const double uRoe = div_nr(L.u * L.sq + R.u * R.sq, L.sq + R.sq);
  const double cRoe = sqrt_nr(kGravity * hRoe);
  s1 = __builtin_fmin(L.u - L.c, uRoe - cRoe);
  s2 = __builtin_fmax(R.u + R.c, uRoe + cRoe);
}

__device__ __forceinline__ double edge_speed_wet(const CellQ& L, const CellQ& R) {
  double s1, s2;
  wave_speeds_wet(L, R, s1, s2);
  return __builtin_fmax(fabs(s1), fabs(s2));
}

// `ign`: lanes holding dummy edges whose results are discarded; they take
// no part in the wave's choice of path (their values may be anything).
__device__ __forceinline__ void edge_solve_wet(const CellQ& L, const CellQ& R, double& dhL,
                                               double& dhR, double& dqL, double& dqR,
                                               unsigned long long ign = 0) {
  double s1, s2;
  wave_speeds_wet(L, R, s1, s2);
  const double d0 = R.q - L.q;
  const double d1 = (R.f - L.f) + (kHalfG * (L.h + R.h)) * (R.b - L.b);
  const double inv = div_nr(1.0, s2 - s1);
  const double beta1 = (s2 * d0 - d1) * inv;
  const double beta2 = (d1 - s1 * d0) * inv;
  const double z1 = beta1 * s1, z2 = beta2 * s2;
  if (wave_all2(s1 < -kZeroTol, s2 > kZeroTol, ign)) {
    // The general path's sums start from +0, so they return +0 where these
    // are -0.  The march (the only caller) uses the four values only in
    // sums that start from +0 or whose other term is never -0 (the cell's
    // accumulators begin as 0.0 + x), so the cell results are the same bits
    // either way, and the additions are left out.
    dhL = beta1; dqL = z1;
    dhR = beta2; dqR = z2;
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

// Dispatch on wave-wide wetness (all lanes must call it together); the
// lanes of `ign` hold dummy edges, as in edge_solve_wet.
__device__ __forceinline__ void edge_solve_any(const CellQ& L, const CellQ& R, double& dhL,
                                               double& dhR, double& dqL, double& dqR,
                                               unsigned long long ign = 0) {
  if (wave_all2(L.h >= kDryTol, R.h >= kDryTol, ign))
    edge_solve_wet(L, R, dhL, dhR, dqL, dqR, ign);
  else
    edge_solve(L, R, dhL, dhR, dqL, dqR);
}

__device__ __forceinline__ double edge_speed_any(const CellQ& L, const CellQ& R) {
  if (wave_all2(L.h >= kDryTol, R.h >= kDryTol)) return edge_speed_wet(L, R);
  return edge_speed(L, R);
}

// ---- augmented Riemann solver ----
//
// SWE's solverType 2 (SConstruct:295-296): the augmented Riemann solver of
// D. L. George, J. Comput. Phys. 227 (2008) 3089-3113, in the form of its
// published implementation (GeoClaw riemann_aug_JCP / riemanntype and the
// rpn2 wet/dry handling), normal direction only, one pass of the
// steady-state wave and of the middle-state estimate, no rarefaction
// entropy corrector.  Three waves with speeds l1 = min(Einfeldt s1, middle
// 2-speed), l2 = (l1 + l3) / 2, l3 = max(Einfeldt s2, middle 1-speed) and
// eigenvectors (1, l1, l1^2), (0, 0, 1), (1, l3, l3^2); the jump
// (dh - ddh, dhu, dphi - ddphi) minus the steady-state wave is solved by
// Cramer's rule.  A side is wet when h > dryTol; a dry side becomes a wall
// (mirrored state, the waves into it removed) when the wet side cannot
// reach its bathymetry, otherwise it is inundated.  Plain sqrt() and '/'
// (correctly rounded) throughout: dry sides produce zeros and the middle
// state may be 0, outside the range of sqrt_nr / div_nr.
constexpr double kAugCritTol = 1e-6;

// Middle-state speeds of the Riemann problem without bathymetry: two
// rarefactions in closed form; two shocks one Newton step on sqrt(h) from
// max(hL, hR); shock + rarefaction one chord step on [min, max].
__device__ __forceinline__ void aug_middle(double hL, double hR, double uL, double uR,
                                           double& hm, double& s1m, double& s2m) {
  const double g = kGravity;
  const double hmin = dmin(hR, hL), hmax = dmax(hR, hL);
  const double du = uR - uL;
  if (hmin <= kDryTol) {
    hm = 0.0;
    s1m = ((uR + uL) - 2.0 * sqrt(g * hR)) + 2.0 * sqrt(g * hL);
    s2m = s1m;
    return;
  }
  const double cmax = sqrt(g * hmax);
  const double fmin = du + 2.0 * (sqrt(g * hmin) - cmax);
  const double fmax = du + (hmax - hmin) * sqrt(((0.5 * g) * (hmax + hmin)) / (hmax * hmin));
  if (fmin > 0.0) {
    const double cL = sqrt(g * hL), cR = sqrt(g * hR);
    const double w = dmax(0.0, 2.0 * (cL + cR) - du);
    hm = (1.0 / (16.0 * g)) * (w * w);
    const double cm = sqrt(g * hm);
    s1m = (uL + 2.0 * cL) - 3.0 * cm;
    s2m = (uR - 2.0 * cR) + 3.0 * cm;
  } else if (fmax <= 0.0) {
    const double h0 = hmax;
    const double gL = sqrt((0.5 * g) * (1.0 / h0 + 1.0 / hL));
    const double gR = sqrt((0.5 * g) * (1.0 / h0 + 1.0 / hR));
    const double f0 = (du + (h0 - hL) * gL) + (h0 - hR) * gR;
    const double h02 = 4.0 * (h0 * h0);
    const double dfdh = (gL - (g * (h0 - hL)) / (h02 * gL)) + (gR - (g * (h0 - hR)) / (h02 * gR));
    const double slope = (2.0 * sqrt(h0)) * dfdh;
    const double r = sqrt(h0) - f0 / slope;
    hm = r * r;
    const double u1m = uL - (hm - hL) * sqrt((0.5 * g) * (1.0 / hm + 1.0 / hL));
    const double u2m = uR + (hm - hR) * sqrt((0.5 * g) * (1.0 / hm + 1.0 / hR));
    const double cm = sqrt(g * hm);
    s1m = u1m - cm;
    s2m = u2m + cm;
  } else {
    const double lo = hmin;
    const double flo = (du + 2.0 * (sqrt(g * lo) - cmax)) +
                       (lo - hmin) * sqrt((0.5 * g) * (1.0 / lo + 1.0 / hmin));
    const double slope = (fmax - flo) / (hmax - lo);
    hm = lo - flo / slope;
    const double cm = sqrt(g * hm);
    if (hL > hR) {
      const double w = uL + 2.0 * sqrt(g * hL);
      s1m = w - 3.0 * cm;
      s2m = w - cm;
    } else {
      const double w = uR - 2.0 * sqrt(g * hR);
      s2m = w + 3.0 * cm;
      s1m = w + cm;
    }
  }
}

__device__ __forceinline__ double det3(double a00, double a01, double a02, double a10,
                                       double a11, double a12, double a20, double a21,
                                       double a22) {
  return (a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20)) +
         a02 * (a10 * a21 - a11 * a20);
}

// Wet/dry handling and wave speeds; returns false for two dry sides.  On
// return the states are the modified ones (zeroed dry side, mirrored wall
// state, lowered bathymetry of an inundated dry side).
struct AugEdge {
  double hL, hR, huL, huR, bL, bR, uL, uR, phiL, phiR, cL, cR;
  double l1, l2, l3, s1m, s2m;
  double w0, w1, w2;
};

__device__ __forceinline__ bool aug_setup(double hL, double hR, double huL, double huR,
                                          double bL, double bR, AugEdge& e) {
  const double g = kGravity;
  if (!(hL > kDryTol) && !(hR > kDryTol)) return false;
  double uL, uR, phiL, phiR;
  if (hR > kDryTol) {
    uR = huR / hR;
    phiR = (0.5 * g) * (hR * hR) + (huR * huR) / hR;
  } else {
    hR = 0.0; huR = 0.0; uR = 0.0; phiR = 0.0;
  }
  if (hL > kDryTol) {
    uL = huL / hL;
    phiL = (0.5 * g) * (hL * hL) + (huL * huL) / hL;
  } else {
    hL = 0.0; huL = 0.0; uL = 0.0; phiL = 0.0;
  }
  double w0 = 1.0, w1 = 1.0, w2 = 1.0;
  double hs, t1, t2;
  if (hR <= kDryTol) {
    aug_middle(hL, hL, uL, -uL, hs, t1, t2);
    if (dmax(hL, hs) + bL < bR) {
      w1 = 0.0; w2 = 0.0;
      hR = hL; huR = -huL; bR = bL; phiR = phiL; uR = -uL;
    } else if (hL + bL < bR) {
      bR = hL + bL;
    }
  } else if (hL <= kDryTol) {
    aug_middle(hR, hR, -uR, uR, hs, t1, t2);
    if (dmax(hR, hs) + bR < bL) {
      w0 = 0.0; w1 = 0.0;
      hL = hR; huL = -huR; bL = bR; phiL = phiR; uL = -uR;
    } else if (hR + bR < bL) {
      bL = hR + bR;
    }
  }
  const double cL = sqrt(g * hL), cR = sqrt(g * hR);
  const double uhat = (cL * uL + cR * uR) / (cR + cL);
  const double chat = sqrt(g * (0.5 * (hR + hL)));
  const double sE1 = dmin(uL - cL, uhat - chat);
  const double sE2 = dmax(uR + cR, uhat + chat);
  double hm, s1m, s2m;
  aug_middle(hL, hR, uL, uR, hm, s1m, s2m);
  e.l1 = dmin(sE1, s2m);
  e.l3 = dmax(sE2, s1m);
  e.l2 = 0.5 * (e.l1 + e.l3);
  e.hL = hL; e.hR = hR; e.huL = huL; e.huR = huR; e.bL = bL; e.bR = bR;
  e.uL = uL; e.uR = uR; e.phiL = phiL; e.phiR = phiR; e.cL = cL; e.cR = cR;
  e.s1m = s1m; e.s2m = s2m;
  e.w0 = w0; e.w1 = w1; e.w2 = w2;
  return true;
}

__device__ __forceinline__ double aug_speed(double hL, double hR, double huL, double huR,
                                            double bL, double bR) {
  AugEdge e;
  if (!aug_setup(hL, hR, huL, huR, bL, bR, e)) return 0.0;
  double s = 0.0;
  s = dmax(s, fabs(e.l1 * e.w0));
  s = dmax(s, fabs(e.l2 * e.w1));
  s = dmax(s, fabs(e.l3 * e.w2));
  return s;
}

__device__ __forceinline__ void aug_split(double sw, double fh, double fq, double& a, double& b,
                                          double& c, double& d) {
  if (sw < -kZeroTol) {
    a += fh; c += fq;
  } else if (sw > kZeroTol) {
    b += fh; d += fq;
  } else {
    a += 0.5 * fh; c += 0.5 * fq;
    b += 0.5 * fh; d += 0.5 * fq;
  }
}

__device__ __forceinline__ void aug_solve(double hL0, double hR0, double huL0, double huR0,
                                          double bL0, double bR0, double& dhL, double& dhR,
                                          double& dqL, double& dqR) {
  const double g = kGravity;
  const double crit = kAugCritTol;
  AugEdge e;
  double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
  if (aug_setup(hL0, hR0, huL0, huR0, bL0, bR0, e)) {
    const double hL = e.hL, hR = e.hR, uL = e.uL, uR = e.uR, cL = e.cL, cR = e.cR;
    const double sE1 = e.l1, sE2 = e.l3, l1 = e.l1, l3 = e.l3;
    const double delh = hR - hL, delhu = e.huR - e.huL, delphi = e.phiR - e.phiL;
    const double delb = e.bR - e.bL;
    const double hstarHLL =
        dmax((((e.huL - e.huR) + sE2 * hR) - sE1 * hL) / (sE2 - sE1), 0.0);
    const double hbar = dmax(0.5 * (hL + hR), 0.0);
    const double s1s2bar = 0.25 * ((uL + uR) * (uL + uR)) - g * hbar;
    const double s1s2tilde = dmax(0.0, uL * uR) - g * hbar;
    const bool sonic = fabs(s1s2bar) <= crit || s1s2bar * s1s2tilde <= crit ||
                       (s1s2bar * sE1) * sE2 <= crit || dmin(fabs(sE1), fabs(sE2)) < crit ||
                       (sE1 < 0.0 && e.s1m > 0.0) || (sE2 > 0.0 && e.s2m < 0.0) ||
                       (uL + cL) * (uR + cR) < 0.0 || (uL - cL) * (uR - cR) < 0.0;
    double ddh = sonic ? -delb : ((delb * g) * hbar) / s1s2bar;
    if (sE1 < -crit && sE2 > crit) {
      ddh = dmin(ddh, (hstarHLL * (sE2 - sE1)) / sE2);
      ddh = dmax(ddh, (hstarHLL * (sE2 - sE1)) / sE1);
    } else if (sE1 >= crit) {
      ddh = dmin(ddh, (hstarHLL * (sE2 - sE1)) / sE1);
      ddh = dmax(ddh, -hL);
    } else if (sE2 <= -crit) {
      ddh = dmin(ddh, hR);
      ddh = dmax(ddh, (hstarHLL * (sE2 - sE1)) / sE2);
    }
    double ddphi =
        sonic ? ((-g) * hbar) * delb : ((((-delb) * g) * hbar) * s1s2tilde) / s1s2bar;
    ddphi = dmin(ddphi, g * dmax((-hL) * delb, (-hR) * delb));
    ddphi = dmax(ddphi, g * dmin((-hL) * delb, (-hR) * delb));
    const double d0 = delh - ddh, d1 = delhu, d2 = delphi - ddphi;
    const double r20 = l1 * l1, r22 = l3 * l3;
    const double det = det3(1.0, 0.0, 1.0, l1, 0.0, l3, r20, 1.0, r22);
    const double beta0 = det3(d0, 0.0, 1.0, d1, 0.0, l3, d2, 1.0, r22) / det;
    const double beta1 = det3(1.0, d0, 1.0, l1, d1, l3, r20, d2, r22) / det;
    const double beta2 = det3(1.0, 0.0, d0, l1, 0.0, d1, r20, 1.0, d2) / det;
    aug_split(l1 * e.w0, (beta0 * l1) * e.w0, (beta0 * r20) * e.w0, a, b, c, d);
    aug_split(e.l2 * e.w1, (beta1 * 0.0) * e.w1, (beta1 * 1.0) * e.w1, a, b, c, d);
    aug_split(l3 * e.w2, (beta2 * l3) * e.w2, (beta2 * r22) * e.w2, a, b, c, d);
  }
  dhL = a; dhR = b; dqL = c; dqR = d;
}

// Edge solver selection for the patch kernels: 0 = f-wave (with its
// wave-uniform fast paths), 1 = augmented Riemann.
template <int SOLVER>
struct EdgeSolver {
  __device__ static __forceinline__ double speed(const CellQ& L, const CellQ& R) {
    return edge_speed_any(L, R);
  }
  __device__ static __forceinline__ void solve(const CellQ& L, const CellQ& R, double& dhL,
                                               double& dhR, double& dqL, double& dqR) {
    edge_solve_any(L, R, dhL, dhR, dqL, dqR);
  }
  // Lanes of `ign` hold dummy edges (any operands, results discarded), and
  // the caller has taken the wetness vote: `all_wet` says that every lane
  // outside `ign` has two wet cells.
  static constexpr bool kDummyLanesFree = true;
  __device__ static __forceinline__ void solve_voted(bool all_wet, const CellQ& L,
                                                     const CellQ& R, unsigned long long ign,
                                                     double& dhL, double& dhR, double& dqL,
                                                     double& dqR) {
    if (all_wet)
      edge_solve_wet(L, R, dhL, dhR, dqL, dqR, ign);
    else
      edge_solve(L, R, dhL, dhR, dqL, dqR);
  }
};

template <>
struct EdgeSolver<1> {
  __device__ static __forceinline__ double speed(const CellQ& L, const CellQ& R) {
    return aug_speed(L.h, R.h, L.q, R.q, L.b, R.b);
  }
  __device__ static __forceinline__ void solve(const CellQ& L, const CellQ& R, double& dhL,
                                               double& dhR, double& dqL, double& dqR) {
    aug_solve(L.h, R.h, L.q, R.q, L.b, R.b, dhL, dhR, dqL, dqR);
  }
  // Dummy edges get real (if meaningless) operands: the solver is not
  // exercised on arbitrary states.  It takes no wetness vote.
  static constexpr bool kDummyLanesFree = false;
  __device__ static __forceinline__ void solve_voted(bool, const CellQ& L, const CellQ& R,
                                                     unsigned long long, double& dhL,
                                                     double& dhR, double& dqL, double& dqR) {
    aug_solve(L.h, R.h, L.q, R.q, L.b, R.b, dhL, dhR, dqL, dqR);
  }
};

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

// max of two bound terms.  The terms are +0, positive, +inf or a NaN with a
// clear sign bit (|hu| and the hardware square root of a positive value, or
// 0 for a dry cell); for those the IEEE order is the order of the bit
// patterns as signed integers, and a NaN wins, so a NaN term still makes the
// edge a candidate.  One integer max instead of fmaxf's two operand
// canonicalisations and the max.
__device__ __forceinline__ float fmax_nonneg(float a, float b) {
  return __int_as_float(max(__float_as_int(a), __float_as_int(b)));
}

// True unless the bound proves the edge slower than `thr` (NaN -> true).
__device__ __forceinline__ bool bound_candidate(float aL, float cL, float aR, float cR,
                                                float thr) {
  const float b = (fmax_nonneg(aL, aR) + fmax_nonneg(cL, cR)) * kBoundMargin;
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
