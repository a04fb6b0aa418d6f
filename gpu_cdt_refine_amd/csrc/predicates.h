// Robust geometric predicates for the refinement path, shared by the HIP
// kernels (device) and the host-side CDT builder.
//
// Method: Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast
// Robust Geometric Predicates" (CMU-CS-96-140).  Each predicate first
// evaluates the determinant in plain FP64 and accepts its sign when it exceeds
// the published forward error bound (stage A).  Otherwise the determinant is
// evaluated exactly with floating-point expansions.  Products use FMA for the
// exact error term (two_product = a*b + fma(a,b,-a*b)), which is exact on both
// the host and gfx950.
//
// Between the FP64 filter and the exact stage sits a double-double stage
// evaluated entirely in registers (see incircle_dd_sign).
//
// Optional first stage (predicate mode 1): a semi-static FP32
// filter.  Coordinates differences are rounded to FP32 and the determinant is
// evaluated in FP32; its sign is accepted only when it exceeds an error bound
// that includes the FP64->FP32 rounding of every difference (relative 2^-24
// each), so the accepted sign equals the exact sign.
//
// Reference counterparts (all replaced, none translated):
//   cuda_fast / cuda_ccw / orient2dexact   cudaCCW.h:445-563
//   cuda_inCircle / cuda_inCircle_exact    cudaCCW.h:570-757
// The reference never uploads its error-bound constants (cudaCCW.h:41 is never
// written), so its "exact" predicates degrade to plain FP64; here the bounds are
// compile-time constants and the exact stage is always reachable.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GCDT_HD __host__ __device__ __forceinline__
#define GCDT_HD_NOINLINE static __host__ __device__ __noinline__
#else
#define GCDT_HD static inline
#define GCDT_HD_NOINLINE static
#endif

#include <math.h>
#include <stdint.h>

namespace gcdt {

// Round-to-nearest FP64 unit roundoff 2^-53.
#define GCDT_EPS (1.1102230246251565e-16)
#define GCDT_CCW_BOUND_A ((3.0 + 16.0 * GCDT_EPS) * GCDT_EPS)
#define GCDT_ICC_BOUND_A ((10.0 + 96.0 * GCDT_EPS) * GCDT_EPS)

// Workspace (in doubles) that the exact incircle needs.  Callers on the
// device pass a per-thread slice of a global-memory arena so that the exact
// stage does not inflate every kernel's private (scratch) segment.
#define GCDT_INCIRCLE_WS 1024

// Per-thread counters for the predicate stages (filter hit rates).
struct PredStats {
  unsigned long long fp32_hits, fp64_hits, exact_calls, dd_hits;
};

GCDT_HD void two_sum(double a, double b, double &x, double &y) {
  x = a + b;
  double bv = x - a;
  double av = x - bv;
  y = (a - av) + (b - bv);
}

GCDT_HD void two_diff(double a, double b, double &x, double &y) {
  x = a - b;
  double bv = a - x;
  double av = x + bv;
  y = (a - av) + (bv - b);
}

GCDT_HD void two_product(double a, double b, double &x, double &y) {
  x = a * b;
  y = fma(a, b, -x);
}

// e += b (single component); e has elen nonoverlapping components in
// increasing magnitude order.  Zero components are eliminated.  Returns the
// new length.  In-place growth: h may alias e (processed left to right).
GCDT_HD int grow_expansion_zeroelim(int elen, double *e, double b) {
  double q = b, qnew, hh;
  int hindex = 0;
  for (int i = 0; i < elen; i++) {
    two_sum(q, e[i], qnew, hh);
    q = qnew;
    if (hh != 0.0) e[hindex++] = hh;
  }
  if (q != 0.0 || hindex == 0) e[hindex++] = q;
  return hindex;
}

// h = e + f (Shewchuk's Fast-Expansion-Sum with zero elimination).  h must
// not alias e or f; capacity elen + flen.
GCDT_HD int fast_expansion_sum_zeroelim(int elen, const double *e, int flen,
                                        const double *f, double *h) {
  // Merge e and f by increasing magnitude, carrying a running sum Q; every
  // roundoff term that is nonzero is emitted.  Bounds-checked reads only.
  double Q, Qnew, hh, enow, fnow;
  int eindex = 0, findex = 0, hindex = 0;
  enow = e[0];
  fnow = f[0];
  if ((fnow > enow) == (fnow > -enow)) {
    Q = enow;
    eindex++;
    enow = (eindex < elen) ? e[eindex] : 0.0;
  } else {
    Q = fnow;
    findex++;
    fnow = (findex < flen) ? f[findex] : 0.0;
  }
  if ((eindex < elen) && (findex < flen)) {
    if ((fnow > enow) == (fnow > -enow)) {
      Qnew = enow + Q;
      hh = Q - (Qnew - enow);
      eindex++;
      enow = (eindex < elen) ? e[eindex] : 0.0;
    } else {
      Qnew = fnow + Q;
      hh = Q - (Qnew - fnow);
      findex++;
      fnow = (findex < flen) ? f[findex] : 0.0;
    }
    Q = Qnew;
    if (hh != 0.0) h[hindex++] = hh;
    while ((eindex < elen) && (findex < flen)) {
      if ((fnow > enow) == (fnow > -enow)) {
        two_sum(Q, enow, Qnew, hh);
        eindex++;
        enow = (eindex < elen) ? e[eindex] : 0.0;
      } else {
        two_sum(Q, fnow, Qnew, hh);
        findex++;
        fnow = (findex < flen) ? f[findex] : 0.0;
      }
      Q = Qnew;
      if (hh != 0.0) h[hindex++] = hh;
    }
  }
  while (eindex < elen) {
    two_sum(Q, enow, Qnew, hh);
    eindex++;
    enow = (eindex < elen) ? e[eindex] : 0.0;
    Q = Qnew;
    if (hh != 0.0) h[hindex++] = hh;
  }
  while (findex < flen) {
    two_sum(Q, fnow, Qnew, hh);
    findex++;
    fnow = (findex < flen) ? f[findex] : 0.0;
    Q = Qnew;
    if (hh != 0.0) h[hindex++] = hh;
  }
  if ((Q != 0.0) || (hindex == 0)) h[hindex++] = Q;
  return hindex;
}

// h = b * e (Shewchuk's Scale-Expansion with zero elimination); capacity 2*elen.
GCDT_HD int scale_expansion_zeroelim(int elen, const double *e, double b,
                                     double *h) {
  double Q, sum, hh, product1, product0;
  int hindex = 0;
  two_product(e[0], b, Q, hh);
  if (hh != 0.0) h[hindex++] = hh;
  for (int eindex = 1; eindex < elen; eindex++) {
    two_product(e[eindex], b, product1, product0);
    two_sum(Q, product0, sum, hh);
    if (hh != 0.0) h[hindex++] = hh;
    double Qn = product1 + sum;  // fast_two_sum(product1, sum)
    hh = sum - (Qn - product1);
    Q = Qn;
    if (hh != 0.0) h[hindex++] = hh;
  }
  if ((Q != 0.0) || (hindex == 0)) h[hindex++] = Q;
  return hindex;
}

GCDT_HD double expansion_sign(int elen, const double *e) {
  // Components are in increasing magnitude; the last nonzero one decides.
  for (int i = elen - 1; i >= 0; i--)
    if (e[i] != 0.0) return e[i] > 0.0 ? 1.0 : -1.0;
  return 0.0;
}

// Exact orient2d from raw coordinates: 6 exact products summed into a
// 12-component expansion held in registers.
GCDT_HD double orient2d_exact(double ax, double ay, double bx, double by,
                              double cx, double cy) {
  double e[12];
  int n = 0;
  double p1, p0;
  two_product(ax, by, p1, p0);
  n = grow_expansion_zeroelim(n, e, p0);
  n = grow_expansion_zeroelim(n, e, p1);
  two_product(-ay, bx, p1, p0);
  n = grow_expansion_zeroelim(n, e, p0);
  n = grow_expansion_zeroelim(n, e, p1);
  two_product(bx, cy, p1, p0);
  n = grow_expansion_zeroelim(n, e, p0);
  n = grow_expansion_zeroelim(n, e, p1);
  two_product(-by, cx, p1, p0);
  n = grow_expansion_zeroelim(n, e, p0);
  n = grow_expansion_zeroelim(n, e, p1);
  two_product(cx, ay, p1, p0);
  n = grow_expansion_zeroelim(n, e, p0);
  n = grow_expansion_zeroelim(n, e, p1);
  two_product(-cy, ax, p1, p0);
  n = grow_expansion_zeroelim(n, e, p0);
  n = grow_expansion_zeroelim(n, e, p1);
  return expansion_sign(n, e);
}

// Sign-exact orient2d: >0 if c lies to the left of a->b (a,b,c CCW).
// Returns the FP64 determinant when stage A certifies its sign, else +-1/0.
GCDT_HD double orient2d(double ax, double ay, double bx, double by, double cx,
                        double cy) {
  double detleft = (ax - cx) * (by - cy);
  double detright = (ay - cy) * (bx - cx);
  double det = detleft - detright;
  double detsum;
  if (detleft > 0.0) {
    if (detright <= 0.0) return det;
    detsum = detleft + detright;
  } else if (detleft < 0.0) {
    if (detright >= 0.0) return det;
    detsum = -detleft - detright;
  } else {
    return orient2d_exact(ax, ay, bx, by, cx, cy);
  }
  double errbound = GCDT_CCW_BOUND_A * detsum;
  if ((det >= errbound) || (-det >= errbound)) return det;
  return orient2d_exact(ax, ay, bx, by, cx, cy);
}

// The value (not only the sign) of an adaptive orient2d, stage by stage as
// in Shewchuk's orient2d / orient2dadapt (CMU-CS-96-140, section 4), which is
// what Triangle's counterclockwise() returns (triangle.cpp:5155-5291) and
// what its findcircumcenter divides by (triangle.cpp:6556-6563).  Stage A:
// the FP64 determinant when its filter certifies it; B: the estimate of the
// exact expansion of the rounded differences' products; C: that estimate
// corrected by the differences' tails; D: the largest component of the exact
// expansion.  The filter part is inline; the rest is out of line (its 40
// doubles of expansions stay off the hot kernels' register budget).
#define GCDT_CCW_BOUND_B ((2.0 + 12.0 * GCDT_EPS) * GCDT_EPS)
#define GCDT_CCW_BOUND_C ((9.0 + 64.0 * GCDT_EPS) * GCDT_EPS * GCDT_EPS)
#define GCDT_RESULT_ERR ((3.0 + 8.0 * GCDT_EPS) * GCDT_EPS)

// (a1*b1 - a2*b2) as a 4-component expansion u[0..3] in increasing magnitude:
// the two exact products (x1 + x0) - (y1 + y0), subtracting y0 then y1, each
// by a difference and a two-sum (Shewchuk's Two-Two-Diff).
GCDT_HD void prod_diff_expansion(double a1, double b1, double a2, double b2, double *u) {
  double x1, x0, y1, y0, j, z, d, d2, top;
  two_product(a1, b1, x1, x0);
  two_product(a2, b2, y1, y0);
  two_diff(x0, y0, d, u[0]);
  two_sum(x1, d, j, z);
  two_diff(z, y1, d2, u[1]);
  two_sum(j, d2, top, u[2]);
  u[3] = top;
}

GCDT_HD_NOINLINE double orient2d_adapt_value(double ax, double ay, double bx, double by,
                                             double cx, double cy, double detsum) {
  const double acx = ax - cx, bcx = bx - cx, acy = ay - cy, bcy = by - cy;
  double B[4], C1[8], C2[12], D[16], u[4];
  prod_diff_expansion(acx, bcy, acy, bcx, B);
  double det = ((B[0] + B[1]) + B[2]) + B[3];
  double err = GCDT_CCW_BOUND_B * detsum;
  if (det >= err || -det >= err) return det;
  double t, acxt, bcxt, acyt, bcyt;
  two_diff(ax, cx, t, acxt);
  two_diff(bx, cx, t, bcxt);
  two_diff(ay, cy, t, acyt);
  two_diff(by, cy, t, bcyt);
  if (acxt == 0.0 && acyt == 0.0 && bcxt == 0.0 && bcyt == 0.0) return det;
  err = GCDT_CCW_BOUND_C * detsum + GCDT_RESULT_ERR * fabs(det);
  det += (acx * bcyt + bcy * acxt) - (acy * bcxt + bcx * acyt);
  if (det >= err || -det >= err) return det;
  prod_diff_expansion(acxt, bcy, acyt, bcx, u);
  int n1 = fast_expansion_sum_zeroelim(4, B, 4, u, C1);
  prod_diff_expansion(acx, bcyt, acy, bcxt, u);
  int n2 = fast_expansion_sum_zeroelim(n1, C1, 4, u, C2);
  prod_diff_expansion(acxt, bcyt, acyt, bcxt, u);
  int n3 = fast_expansion_sum_zeroelim(n2, C2, 4, u, D);
  return D[n3 - 1];
}

// Stage A of the value: true when the FP64 determinant is returned as is
// (*det), false when the adaptive stages must run (*detsum is their input).
GCDT_HD bool orient2d_value_filter(double ax, double ay, double bx, double by, double cx,
                                   double cy, double *det, double *detsum) {
  const double l = (ax - cx) * (by - cy), r = (ay - cy) * (bx - cx);
  *det = l - r;
  if (l > 0.0) {
    if (r <= 0.0) return true;
    *detsum = l + r;
  } else if (l < 0.0) {
    if (r >= 0.0) return true;
    *detsum = -l - r;
  } else {
    return true;
  }
  const double err = GCDT_CCW_BOUND_A * *detsum;
  return *det >= err || -*det >= err;
}

GCDT_HD double orient2d_value(double ax, double ay, double bx, double by, double cx,
                              double cy) {
  double det, detsum = 0.0;
  if (orient2d_value_filter(ax, ay, bx, by, cx, cy, &det, &detsum)) return det;
  return orient2d_adapt_value(ax, ay, bx, by, cx, cy, detsum);
}

// Range of the FP32 stages: every coordinate difference must be 0 or have a
// magnitude in [2^-lo, 2^hi], so that no FP32 difference, product or degree-4
// term is subnormal or infinite and the relative error bounds below hold.
// (An FP32 sum that cancels into the subnormal range is exact.)
GCDT_HD bool fp32_in_range(float x, float lo, float hi) {
  const float a = fabsf(x);
  return a == 0.0f || (a >= lo && a <= hi);
}

// FP32 semi-static stage for orient2d.  Returns +1/-1 when certified, 0 when
// undecided.  The differences are formed in FP64 (relative error <= u64) and
// rounded to FP32 (relative error <= u32 = 2^-24 each); products and the final
// difference add three FP32 roundings.  The bound below covers all of them
// with margin: |det - det_f32| <= (4*u32 + 8*u64 + O(u32^2)) * (|l| + |r|).
// Differences outside [2^-60, 2^60] defer (products stay normal).
GCDT_HD int orient2d_fp32_sign(double ax, double ay, double bx, double by,
                               double cx, double cy) {
  float acx = (float)(ax - cx), bcy = (float)(by - cy);
  float acy = (float)(ay - cy), bcx = (float)(bx - cx);
  const float lo = 0x1p-60f, hi = 0x1p60f;
  if (!(fp32_in_range(acx, lo, hi) && fp32_in_range(bcy, lo, hi) &&
        fp32_in_range(acy, lo, hi) && fp32_in_range(bcx, lo, hi)))
    return 0;
  float l = acx * bcy, r = acy * bcx;
  float det = l - r;
  float bound = 2.4e-7f * (fabsf(l) + fabsf(r));  // 4.03 * 2^-24 rounded up
  // Guard against FP32 underflow / denormals: require a normal-range bound.
  if (bound < 1e-30f) return 0;
  if (det > bound) return 1;
  if (-det > bound) return -1;
  return 0;
}

// --- exact incircle -------------------------------------------------------
// d inside circle(a,b,c) (a,b,c CCW) => positive.
//
// Stage B (common): when the six coordinate differences relative to d are
// exact (no FP64 rounding), the determinant is a polynomial in those six
// doubles; it is accumulated exactly from 3 x (4 x 8)-component products.
// Stage C (general): the 4x4 lifted determinant in raw coordinates, expanded
// along the lift column: sum_i +-lift_i * orient(minor_i), each orient a <=12
// component expansion and each lift <=4 components.
// ws: workspace of GCDT_INCIRCLE_WS doubles.

GCDT_HD_NOINLINE double incircle_exact(double ax, double ay, double bx,
                                       double by, double cx, double cy,
                                       double dx, double dy, double *ws) {
  double *acc = ws;          // [0, 400)
  double *acc2 = ws + 400;   // [400, 800)
  double *tmp = ws + 800;    // [800, 1000)
  double adx, bdx, cdx, ady, bdy, cdy, t0, t1, t2, t3, t4, t5;
  two_diff(ax, dx, adx, t0);
  two_diff(bx, dx, bdx, t1);
  two_diff(cx, dx, cdx, t2);
  two_diff(ay, dy, ady, t3);
  two_diff(by, dy, bdy, t4);
  two_diff(cy, dy, cdy, t5);
  int n = 0;
  if (t0 == 0.0 && t1 == 0.0 && t2 == 0.0 && t3 == 0.0 && t4 == 0.0 &&
      t5 == 0.0) {
    // Stage B: exact differences.
    double u[3][2] = {{bdx, cdy}, {cdx, ady}, {adx, bdy}};
    double v[3][2] = {{cdx, bdy}, {adx, cdy}, {bdx, ady}};
    double lx[3] = {adx, bdx, cdx}, ly[3] = {ady, bdy, cdy};
    for (int k = 0; k < 3; k++) {
      double m[4], lift[4], p1, p0, q1, q0;
      // minor = u0*u1 - v0*v1 (4 components)
      two_product(u[k][0], u[k][1], p1, p0);
      two_product(-v[k][0], v[k][1], q1, q0);
      int mn = 0;
      mn = grow_expansion_zeroelim(mn, m, p0);
      mn = grow_expansion_zeroelim(mn, m, p1);
      mn = grow_expansion_zeroelim(mn, m, q0);
      mn = grow_expansion_zeroelim(mn, m, q1);
      // lift = lx^2 + ly^2 (4 components)
      two_product(lx[k], lx[k], p1, p0);
      two_product(ly[k], ly[k], q1, q0);
      int ln = 0;
      ln = grow_expansion_zeroelim(ln, lift, p0);
      ln = grow_expansion_zeroelim(ln, lift, p1);
      ln = grow_expansion_zeroelim(ln, lift, q0);
      ln = grow_expansion_zeroelim(ln, lift, q1);
      for (int i = 0; i < ln; i++) {
        int tn = scale_expansion_zeroelim(mn, m, lift[i], tmp);
        if (n == 0) {
          for (int j = 0; j < tn; j++) acc[j] = tmp[j];
          n = tn;
        } else {
          int nn = fast_expansion_sum_zeroelim(n, acc, tn, tmp, acc2);
          double *sw = acc; acc = acc2; acc2 = sw;
          n = nn;
        }
      }
    }
    return expansion_sign(n, acc);
  }
  // Stage C: raw-coordinate expansion along the lift column.
  //   det = +la*O(b,c,d) - lb*O(a,c,d) + lc*O(a,b,d) - ld*O(a,b,c)
  // (cofactor expansion of |x y x^2+y^2 1| along the lift column), with
  // O(p,q,r) the orient2d determinant.  For a,b,c CCW on the unit circle and
  // d at its center this is 2*area(a,b,c) > 0: "inside => positive".
  const double px[4] = {ax, bx, cx, dx}, py[4] = {ay, by, cy, dy};
  for (int k = 0; k < 4; k++) {
    int i0 = (k == 0) ? 1 : 0;
    int i1 = (k <= 1) ? 2 : 1;
    int i2 = (k <= 2) ? 3 : 2;
    double o[12], p1, p0;
    int on = 0;
    const double qx[3] = {px[i0], px[i1], px[i2]};
    const double qy[3] = {py[i0], py[i1], py[i2]};
    for (int j = 0; j < 3; j++) {
      int j1 = (j + 1) % 3;
      two_product(qx[j], qy[j1], p1, p0);
      on = grow_expansion_zeroelim(on, o, p0);
      on = grow_expansion_zeroelim(on, o, p1);
      two_product(-qy[j], qx[j1], p1, p0);
      on = grow_expansion_zeroelim(on, o, p0);
      on = grow_expansion_zeroelim(on, o, p1);
    }
    double sgn = (k % 2 == 0) ? 1.0 : -1.0;
    double lift[4];
    int ln = 0;
    two_product(px[k], px[k], p1, p0);
    ln = grow_expansion_zeroelim(ln, lift, sgn * p0);
    ln = grow_expansion_zeroelim(ln, lift, sgn * p1);
    two_product(py[k], py[k], p1, p0);
    ln = grow_expansion_zeroelim(ln, lift, sgn * p0);
    ln = grow_expansion_zeroelim(ln, lift, sgn * p1);
    for (int i = 0; i < ln; i++) {
      int tn = scale_expansion_zeroelim(on, o, lift[i], tmp);
      if (n == 0) {
        for (int j = 0; j < tn; j++) acc[j] = tmp[j];
        n = tn;
      } else {
        int nn = fast_expansion_sum_zeroelim(n, acc, tn, tmp, acc2);
        double *sw = acc; acc = acc2; acc2 = sw;
        n = nn;
      }
    }
  }
  return expansion_sign(n, acc);
}

// FP64 stage A of incircle.  Returns the determinant when certified; sets
// *ok = false otherwise.
GCDT_HD double incircle_fp64(double ax, double ay, double bx, double by,
                             double cx, double cy, double dx, double dy,
                             bool *ok) {
  double adx = ax - dx, bdx = bx - dx, cdx = cx - dx;
  double ady = ay - dy, bdy = by - dy, cdy = cy - dy;
  double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
  double alift = adx * adx + ady * ady;
  double cdxady = cdx * ady, adxcdy = adx * cdy;
  double blift = bdx * bdx + bdy * bdy;
  double adxbdy = adx * bdy, bdxady = bdx * ady;
  double clift = cdx * cdx + cdy * cdy;
  double det = alift * (bdxcdy - cdxbdy) + blift * (cdxady - adxcdy) +
               clift * (adxbdy - bdxady);
  double permanent = (fabs(bdxcdy) + fabs(cdxbdy)) * alift +
                     (fabs(cdxady) + fabs(adxcdy)) * blift +
                     (fabs(adxbdy) + fabs(bdxady)) * clift;
  double errbound = GCDT_ICC_BOUND_A * permanent;
  *ok = (det > errbound) || (-det > errbound);
  return det;
}

// FP32 semi-static stage for incircle: +1/-1 certified, 0 undecided.  Each of
// the six differences carries relative error <= u32 + u64 after rounding to
// FP32; the degree-4 evaluation adds at most 7 FP32 roundings per term.  The
// bound 1.2e-6 * permanent (about 20 * 2^-24) covers both with margin.
// Differences outside [2^-30, 2^30] defer: degree-4 terms then stay within
// [2^-120, 2^120], inside FP32's normal range.
GCDT_HD int incircle_fp32_sign(double ax, double ay, double bx, double by,
                               double cx, double cy, double dx, double dy) {
  float adx = (float)(ax - dx), bdx = (float)(bx - dx), cdx = (float)(cx - dx);
  float ady = (float)(ay - dy), bdy = (float)(by - dy), cdy = (float)(cy - dy);
  const float lo = 0x1p-30f, hi = 0x1p30f;
  if (!(fp32_in_range(adx, lo, hi) && fp32_in_range(bdx, lo, hi) &&
        fp32_in_range(cdx, lo, hi) && fp32_in_range(ady, lo, hi) &&
        fp32_in_range(bdy, lo, hi) && fp32_in_range(cdy, lo, hi)))
    return 0;
  float bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
  float alift = adx * adx + ady * ady;
  float cdxady = cdx * ady, adxcdy = adx * cdy;
  float blift = bdx * bdx + bdy * bdy;
  float adxbdy = adx * bdy, bdxady = bdx * ady;
  float clift = cdx * cdx + cdy * cdy;
  float det = alift * (bdxcdy - cdxbdy) + blift * (cdxady - adxcdy) +
              clift * (adxbdy - bdxady);
  float permanent = (fabsf(bdxcdy) + fabsf(cdxbdy)) * alift +
                    (fabsf(cdxady) + fabsf(adxcdy)) * blift +
                    (fabsf(adxbdy) + fabsf(bdxady)) * clift;
  // Finite, normal-range magnitudes only (FP32 overflow/underflow => defer).
  if (!(permanent < 1e30f) || permanent < 1e-25f) return 0;
  float bound = 1.2e-6f * permanent;
  if (det > bound) return 1;
  if (-det > bound) return -1;
  return 0;
}

// --- double-double stage (registers only) ---------------------------------
// Between the FP64 filter and the exact expansions: the coordinate
// differences are formed exactly as (hi, lo) pairs and the determinant is
// evaluated in double-double arithmetic (about 106-bit significands).  With
// u = 2^-53, |lo| <= u|hi| and FMA-exact products, a double-double product has
// absolute error <= 8u^2 |a||b| and a sum <= 2u^2 (|a| + |b|); carried through
// lift_k * minor_k and the three-term sum this bounds the error by
// 48 u^2 * permanent < 2^-100 * permanent.  The sign is accepted when |det|
// exceeds 2^-94 * permanent (a 2^6 margin), with the permanent held in the
// normal FP64 range so that nothing overflows and underflow stays far below
// the bound.  Only (near-)exact degeneracies reach the expansion stage.
struct DD {
  double hi, lo;
};

GCDT_HD DD dd_from_diff(double a, double b) {
  DD r;
  two_diff(a, b, r.hi, r.lo);
  return r;
}

GCDT_HD DD dd_fast(double a, double b) {  // exact renormalisation (two_sum)
  DD r;
  two_sum(a, b, r.hi, r.lo);
  return r;
}

GCDT_HD DD dd_add(DD a, DD b) {
  double s, e, t, f;
  two_sum(a.hi, b.hi, s, e);
  two_sum(a.lo, b.lo, t, f);
  e += t;
  DD r = dd_fast(s, e);
  r.lo += f;
  return dd_fast(r.hi, r.lo);
}

GCDT_HD DD dd_neg(DD a) {
  a.hi = -a.hi;
  a.lo = -a.lo;
  return a;
}

GCDT_HD DD dd_mul(DD a, DD b) {
  double p, e;
  two_product(a.hi, b.hi, p, e);
  e += a.hi * b.lo + a.lo * b.hi;
  return dd_fast(p, e);
}

// +1 / -1 certified, 0 undecided.
GCDT_HD int incircle_dd_sign(double ax, double ay, double bx, double by,
                             double cx, double cy, double dx, double dy) {
  DD adx = dd_from_diff(ax, dx), ady = dd_from_diff(ay, dy);
  DD bdx = dd_from_diff(bx, dx), bdy = dd_from_diff(by, dy);
  DD cdx = dd_from_diff(cx, dx), cdy = dd_from_diff(cy, dy);
  DD alift = dd_add(dd_mul(adx, adx), dd_mul(ady, ady));
  DD blift = dd_add(dd_mul(bdx, bdx), dd_mul(bdy, bdy));
  DD clift = dd_add(dd_mul(cdx, cdx), dd_mul(cdy, cdy));
  DD ma = dd_add(dd_mul(bdx, cdy), dd_neg(dd_mul(cdx, bdy)));
  DD mb = dd_add(dd_mul(cdx, ady), dd_neg(dd_mul(adx, cdy)));
  DD mc = dd_add(dd_mul(adx, bdy), dd_neg(dd_mul(bdx, ady)));
  DD det = dd_add(dd_add(dd_mul(alift, ma), dd_mul(blift, mb)), dd_mul(clift, mc));
  double permanent = (fabs(bdx.hi * cdy.hi) + fabs(cdx.hi * bdy.hi)) * alift.hi +
                     (fabs(cdx.hi * ady.hi) + fabs(adx.hi * cdy.hi)) * blift.hi +
                     (fabs(adx.hi * bdy.hi) + fabs(bdx.hi * ady.hi)) * clift.hi;
  // keep every intermediate far from underflow/overflow, else defer
  if (!(permanent > 1e-250 && permanent < 1e250)) return 0;
  double bound = permanent * 5.048709793414476e-29;  // 2^-94
  if (det.hi > bound) return 1;
  if (det.hi < -bound) return -1;
  return 0;
}

GCDT_HD int orient2d_dd_sign(double ax, double ay, double bx, double by,
                             double cx, double cy) {
  DD acx = dd_from_diff(ax, cx), bcy = dd_from_diff(by, cy);
  DD acy = dd_from_diff(ay, cy), bcx = dd_from_diff(bx, cx);
  DD det = dd_add(dd_mul(acx, bcy), dd_neg(dd_mul(acy, bcx)));
  double permanent = fabs(acx.hi * bcy.hi) + fabs(acy.hi * bcx.hi);
  if (!(permanent > 1e-280 && permanent < 1e280)) return 0;
  double bound = permanent * 5.048709793414476e-29;  // 2^-94
  if (det.hi > bound) return 1;
  if (det.hi < -bound) return -1;
  return 0;
}

// Sign-exact incircle: >0 if d strictly inside circle(a,b,c), a,b,c CCW.
// use_fp32 enables the FP32 semi-static stage.  stats may be null.
// Rare stages, out of line so the hot filters keep a small register
// footprint: 0 = exact zero; stats->dd_hits / exact_calls counted here.
GCDT_HD_NOINLINE int incircle_slow(double ax, double ay, double bx, double by,
                                   double cx, double cy, double dx, double dy,
                                   double *ws, PredStats *stats) {
  int sd = incircle_dd_sign(ax, ay, bx, by, cx, cy, dx, dy);
  if (sd != 0) {
    if (stats) stats->dd_hits++;
    return sd;
  }
  if (stats) stats->exact_calls++;
  double s = incircle_exact(ax, ay, bx, by, cx, cy, dx, dy, ws);
  return s > 0.0 ? 1 : (s < 0.0 ? -1 : 0);
}

// Filter stages only: the sign, or PRED_UNSURE when neither filter decides.
constexpr int PRED_UNSURE = 2;

GCDT_HD int incircle_filtered(double ax, double ay, double bx, double by,
                              double cx, double cy, double dx, double dy,
                              bool use_fp32, PredStats *stats) {
  if (use_fp32) {
    int s = incircle_fp32_sign(ax, ay, bx, by, cx, cy, dx, dy);
    if (s != 0) {
      if (stats) stats->fp32_hits++;
      return s;
    }
  }
  bool ok;
  double det = incircle_fp64(ax, ay, bx, by, cx, cy, dx, dy, &ok);
  if (ok) {
    if (stats) stats->fp64_hits++;
    return det > 0.0 ? 1 : -1;
  }
  return PRED_UNSURE;
}

GCDT_HD int incircle_sign(double ax, double ay, double bx, double by,
                          double cx, double cy, double dx, double dy,
                          double *ws, bool use_fp32, PredStats *stats) {
  int s = incircle_filtered(ax, ay, bx, by, cx, cy, dx, dy, use_fp32, stats);
  if (s != PRED_UNSURE) return s;
  return incircle_slow(ax, ay, bx, by, cx, cy, dx, dy, ws, stats);
}

GCDT_HD_NOINLINE int orient2d_slow(double ax, double ay, double bx, double by,
                                   double cx, double cy, PredStats *stats) {
  int sd = orient2d_dd_sign(ax, ay, bx, by, cx, cy);
  if (sd != 0) {
    if (stats) stats->dd_hits++;
    return sd;
  }
  if (stats) stats->exact_calls++;
  double s = orient2d_exact(ax, ay, bx, by, cx, cy);
  return s > 0.0 ? 1 : (s < 0.0 ? -1 : 0);
}

GCDT_HD int orient2d_filtered(double ax, double ay, double bx, double by,
                              double cx, double cy, bool use_fp32,
                              PredStats *stats) {
  if (use_fp32) {
    int s = orient2d_fp32_sign(ax, ay, bx, by, cx, cy);
    if (s != 0) {
      if (stats) stats->fp32_hits++;
      return s;
    }
  }
  double detleft = (ax - cx) * (by - cy);
  double detright = (ay - cy) * (bx - cx);
  double det = detleft - detright;
  double detsum = fabs(detleft) + fabs(detright);
  if ((det > GCDT_CCW_BOUND_A * detsum) || (-det > GCDT_CCW_BOUND_A * detsum)) {
    if (stats) stats->fp64_hits++;
    return det > 0.0 ? 1 : -1;
  }
  return PRED_UNSURE;
}

GCDT_HD int orient2d_sign(double ax, double ay, double bx, double by,
                          double cx, double cy, bool use_fp32,
                          PredStats *stats) {
  int s = orient2d_filtered(ax, ay, bx, by, cx, cy, use_fp32, stats);
  if (s != PRED_UNSURE) return s;
  return orient2d_slow(ax, ay, bx, by, cx, cy, stats);
}

}  // namespace gcdt
