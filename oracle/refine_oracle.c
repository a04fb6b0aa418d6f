/*
 * Sequential restatement of the bulk-synchronous CDT quality refinement
 * (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * What it restates, and from where:
 *   - round structure: encroached subsegments are split first, then bad
 *     triangles are marked and split, repeat until none remain
 *     (reference refine.cu:6831-7024 GPU_Refine_Quality; Triangle
 *     enforcequality triangle.cpp:13646-13738)
 *   - bad-triangle test incl. the Miller-Pav-Walkington exemption
 *     (refine.cu:376-497 kernelMarkBadTriangles; triangle.cpp:7220-7384)
 *   - circumcenter / off-center (refine.cu:3780-3868; triangle.cpp:6523-6618)
 *     with the off-center constant kept in double (the reference truncates it
 *     to int through splitTriangles' parameter, refine.cu:5132)
 *   - encroachment test, diametral circle (Ruppert) or lens (Chew)
 *     (refine.cu:115-139; triangle.cpp:7106-7203)
 *   - segment split point with concentric power-of-two shells and one
 *     collinearity correction step (refine.cu:2686-2725; triangle.cpp:13368-13415)
 *   - point location by orientation walk that stops at subsegments, and
 *     rejection of Steiner points that land on / behind a subsegment or
 *     encroach one (refine.cu:3908-4184, 5061-5089; triangle.cpp:8290-8325)
 *   - conflict resolution: a candidate claims its Bowyer-Watson cavity and the
 *     triangles across the cavity boundary; the lowest candidate index wins
 *     every contested triangle (the reference's atomicMin rule,
 *     refine.cu:4159-4172, 4296)
 *
 * Insertion itself is cavity retriangulation (a star from the new vertex over
 * the cavity boundary), which leaves the mesh constrained-Delaunay in one step
 * (the reference reaches the same mesh by 1->3 / 2->4 splits followed by
 * rounds of edge flips, refine.cu:2328-2490).  The exact tie rules below
 * (candidate order, cavity order, star ordering, slot assignment) are the
 * contract the GPU path reproduces.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#include "oracle.h"

#define TF_ALIVE 1
#define TF_CHECK 2
#define TF_BAD 4
#define TF_GIVEUP 8
#define PF_INPUT 1
#define PF_SEGV 2
#define PF_FREE 4
#define PF_DEAD 8
#define PF_NEW 16 /* inserted in the run's last round (set at export) */

#define ST_VALID 0
#define ST_REJ_ENC 1
#define ST_REJ_BLOCKED 2
#define ST_OVERFLOW 3
#define ST_GIVEUP 4
#define ST_DEFER 6 /* not evaluated this pass (see the vertex deletion below) */

#define WALK_LIMIT (1 << 22)

typedef struct {
  int np, capp;
  double *P;
  uint8_t *pflag;
  int *ppar;
  int nt, capt;
  int *TV, *TN, *TS;
  uint8_t *tflag;
  int ns, caps;
  int *SV, *SP, *ST;
  uint8_t *senc;
  int nseg_in;
  int *SEG;
  double goodangle, lens, offconst;
  int offcenter, ruppert;
  /* batched refinement: instance of every triangle and per-instance
   * parameters (goodangle, lens, offconst, offcenter, ruppert) */
  int *tinst;
  const double *ipar;
} Mesh;

struct orc_result {
  int np, nt, ns;
  double *P;
  int *T, *N, *S;
  uint8_t *pflag;
  int *I; /* instance of every output triangle (batched refinement) */
};

static void *xrealloc(void *p, size_t n) {
  void *q = realloc(p, n);
  if (!q && n) abort();
  return q;
}

static void grow_points(Mesh *m, int need) {
  if (need <= m->capp) return;
  int c = m->capp ? m->capp : 16;
  while (c < need) c *= 2;
  m->P = xrealloc(m->P, sizeof(double) * 2 * (size_t)c);
  m->pflag = xrealloc(m->pflag, (size_t)c);
  m->ppar = xrealloc(m->ppar, sizeof(int) * (size_t)c);
  m->capp = c;
}

static void grow_tris(Mesh *m, int need) {
  if (need <= m->capt) return;
  int c = m->capt ? m->capt : 16;
  while (c < need) c *= 2;
  m->TV = xrealloc(m->TV, sizeof(int) * 3 * (size_t)c);
  m->TN = xrealloc(m->TN, sizeof(int) * 3 * (size_t)c);
  m->TS = xrealloc(m->TS, sizeof(int) * 3 * (size_t)c);
  m->tflag = xrealloc(m->tflag, (size_t)c);
  if (m->ipar) m->tinst = xrealloc(m->tinst, sizeof(int) * (size_t)c);
  m->capt = c;
}

static void grow_segs(Mesh *m, int need) {
  if (need <= m->caps) return;
  int c = m->caps ? m->caps : 16;
  while (c < need) c *= 2;
  m->SV = xrealloc(m->SV, sizeof(int) * 2 * (size_t)c);
  m->SP = xrealloc(m->SP, sizeof(int) * (size_t)c);
  m->ST = xrealloc(m->ST, sizeof(int) * (size_t)c);
  m->senc = xrealloc(m->senc, (size_t)c);
  m->caps = c;
}

static const double *PT(const Mesh *m, int v) { return m->P + 2 * (size_t)v; }

/* Make triangle t's instance parameters current (batched refinement). */
static void use_inst(Mesh *m, int t) {
  if (!m->ipar) return;
  const double *q = m->ipar + 5 * (size_t)m->tinst[t];
  m->goodangle = q[0];
  m->lens = q[1];
  m->offconst = q[2];
  m->offcenter = (int)q[3];
  m->ruppert = (int)q[4];
}

/* checkseg4encroach: does apex p encroach subsegment (a,b)?
 * refine.cu:115-139, triangle.cpp:7138-7153 */
static int encroaches_xy(const double *A, const double *B, const double *p, int ruppert,
                         double lens) {
  double dot = (A[0] - p[0]) * (B[0] - p[0]) + (A[1] - p[1]) * (B[1] - p[1]);
  if (dot < 0.0) {
    if (ruppert) return 1;
    double la = (A[0] - p[0]) * (A[0] - p[0]) + (A[1] - p[1]) * (A[1] - p[1]);
    double lb = (B[0] - p[0]) * (B[0] - p[0]) + (B[1] - p[1]) * (B[1] - p[1]);
    if (dot * dot >= lens * la * lb) return 1;
  }
  return 0;
}

static int encroaches(const Mesh *m, int a, int b, const double *p) {
  return encroaches_xy(PT(m, a), PT(m, b), p, m->ruppert, m->lens);
}

/* cos^2 theta and the lens constant (2 cos^2 theta - 1)^2 of an angle in
 * degrees, as GPU_Refine_Quality derives them (refine.cu:6557-6567). */
static void angle_constants(double theta_deg, double *goodangle, double *lens) {
  double ga = cos(theta_deg * 3.141592653589793238462643383279502884197169399375105820974944592308 / 180.0);
  *goodangle = ga * ga;
  *lens = (2.0 * *goodangle - 1.0) * (2.0 * *goodangle - 1.0);
}

int orc_encroaches_xy(const double *a, const double *b, const double *p, double theta_deg,
                      int ruppert) {
  double goodangle, lens;
  angle_constants(theta_deg, &goodangle, &lens);
  return encroaches_xy(a, b, p, ruppert, lens);
}

/* testtriangle with Miller-Pav-Walkington exemption; triangle.cpp:7220-7384,
 * refine.cu:376-497.  org = v1, dest = v2, apex = v0 (orientation 0). */
/* The angle test of testtriangle: cos^2 of the angle opposite the shortest
 * edge against cos^2 theta; *e receives that edge (opposite corner e of
 * (apex, org, dest)). */
static int bad_angle_xy(const double *A, const double *O, const double *D, double goodangle,
                        int *e_out) {
  double dxod = O[0] - D[0], dyod = O[1] - D[1];
  double dxda = D[0] - A[0], dyda = D[1] - A[1];
  double dxao = A[0] - O[0], dyao = A[1] - O[1];
  double apexlen = dxod * dxod + dyod * dyod;
  double orglen = dxda * dxda + dyda * dyda;
  double destlen = dxao * dxao + dyao * dyao;
  double angle;
  int e;
  if ((apexlen < orglen) && (apexlen < destlen)) {
    angle = dxda * dxao + dyda * dyao;
    angle = angle * angle / (orglen * destlen);
    e = 0;
  } else if (orglen < destlen) {
    angle = dxod * dxao + dyod * dyao;
    angle = angle * angle / (apexlen * destlen);
    e = 1;
  } else {
    angle = dxod * dxda + dyod * dyda;
    angle = angle * angle / (apexlen * orglen);
    e = 2;
  }
  *e_out = e;
  return angle > goodangle;
}

int orc_bad_angle_xy(const double *apex, const double *org, const double *dest,
                     double theta_deg) {
  double goodangle, lens;
  int e;
  angle_constants(theta_deg, &goodangle, &lens);
  return bad_angle_xy(apex, org, dest, goodangle, &e);
}

static int tri_is_bad(const Mesh *m, int t) {
  const int *v = m->TV + 3 * (size_t)t;
  const double *O = PT(m, v[1]), *D = PT(m, v[2]), *A = PT(m, v[0]);
  int e;
  if (!bad_angle_xy(A, O, D, m->goodangle, &e)) return 0;
  /* the shortest edge's endpoints (edge e joins corners e+1 and e+2) */
  const int b1 = v[(e + 1) % 3], b2 = v[(e + 2) % 3];
  if ((m->pflag[b1] & PF_SEGV) && (m->pflag[b2] & PF_SEGV) &&
      m->TS[3 * (size_t)t + e] < 0) {
    int s1 = m->ppar[b1], s2 = m->ppar[b2];
    int o1 = m->SEG[2 * s1], d1 = m->SEG[2 * s1 + 1];
    int o2 = m->SEG[2 * s2], d2 = m->SEG[2 * s2 + 1];
    int common = -1;
    if (o1 == o2 || o1 == d2) common = o1;
    else if (d1 == o2 || d1 == d2) common = d1;
    if (common >= 0) {
      const double *J = PT(m, common), *B1 = PT(m, b1), *B2 = PT(m, b2);
      double dist1 = (B1[0] - J[0]) * (B1[0] - J[0]) + (B1[1] - J[1]) * (B1[1] - J[1]);
      double dist2 = (B2[0] - J[0]) * (B2[0] - J[0]) + (B2[1] - J[1]) * (B2[1] - J[1]);
      if ((dist1 < 1.001 * dist2) && (dist1 > 0.999 * dist2)) return 0;
    }
  }
  return 1;
}

/* ---- the VALUE of Triangle's adaptive orient2d ----------------------------
 * Triangle divides by counterclockwise(dest, apex, org) in findcircumcenter
 * (triangle.cpp:6556-6563); the reference's computeCircumcenter falls back to
 * cuda_ccw only when the raw determinant is 0 (refine.cu:3799-3804).  The
 * value (not only its sign) depends on which adaptive stage returns, so this
 * follows Shewchuk's orient2d / orient2dadapt stage by stage (CMU-CS-96-140,
 * section 4; Triangle's copy at triangle.cpp:5155-5291): filter A on the FP64
 * determinant, then B = the exact product expansion of the rounded
 * differences (returns its estimate), then C = that estimate corrected by the
 * difference tails, then D = the exact expansion, whose largest component is
 * returned.  Products are split with Dekker's splitter here (the device code
 * uses FMA; both give the exact error term). */
static const double ORC_EPS = 1.1102230246251565e-16; /* 2^-53 */
#define ORC_CCW_A ((3.0 + 16.0 * ORC_EPS) * ORC_EPS)
#define ORC_CCW_B ((2.0 + 12.0 * ORC_EPS) * ORC_EPS)
#define ORC_CCW_C ((9.0 + 64.0 * ORC_EPS) * ORC_EPS * ORC_EPS)
#define ORC_RESULT_ERR ((3.0 + 8.0 * ORC_EPS) * ORC_EPS)

static void orc_two_sum(double a, double b, double *x, double *y) {
  double s = a + b, bv = s - a, av = s - bv;
  *x = s;
  *y = (a - av) + (b - bv);
}
/* tail of x = fl(a - b) */
static double orc_diff_tail(double a, double b, double x) {
  double bv = a - x, av = x + bv;
  return (a - av) + (bv - b);
}
static void orc_split(double a, double *hi, double *lo) {
  double c = 134217729.0 * a; /* 2^27 + 1 */
  double big = c - a;
  *hi = c - big;
  *lo = a - *hi;
}
static void orc_two_product(double a, double b, double *x, double *y) {
  double p = a * b, ah, al, bh, bl;
  orc_split(a, &ah, &al);
  orc_split(b, &bh, &bl);
  double e1 = p - ah * bh, e2 = e1 - al * bh, e3 = e2 - ah * bl;
  *x = p;
  *y = al * bl - e3;
}
/* (a1 + a0) - (b1 + b0) as a 4-component expansion x[0..3] (increasing):
 * subtract b0 from (a1, a0), then b1 from the result, each by a
 * one-component difference followed by a two-sum. */
static void orc_two_two_diff(double a1, double a0, double b1, double b0, double *x) {
  double i, j, z;
  double d = a0 - b0;
  x[0] = orc_diff_tail(a0, b0, d);
  orc_two_sum(a1, d, &j, &z);
  double d2 = z - b1;
  x[1] = orc_diff_tail(z, b1, d2);
  orc_two_sum(j, d2, &i, &x[2]);
  x[3] = i;
}
/* Fast-Expansion-Sum with zero elimination (Shewchuk's Figure 6 variant). */
static int orc_fast_expansion_sum(int elen, const double *e, int flen, const double *f,
                                  double *h) {
  double Q, Qnew, hh, enow = e[0], fnow = f[0];
  int ei = 0, fi = 0, hi = 0;
  if ((fnow > enow) == (fnow > -enow)) { Q = enow; enow = ++ei < elen ? e[ei] : 0.0; }
  else { Q = fnow; fnow = ++fi < flen ? f[fi] : 0.0; }
  if (ei < elen && fi < flen) {
    if ((fnow > enow) == (fnow > -enow)) {
      Qnew = enow + Q; hh = Q - (Qnew - enow); enow = ++ei < elen ? e[ei] : 0.0;
    } else {
      Qnew = fnow + Q; hh = Q - (Qnew - fnow); fnow = ++fi < flen ? f[fi] : 0.0;
    }
    Q = Qnew;
    if (hh != 0.0) h[hi++] = hh;
    while (ei < elen && fi < flen) {
      if ((fnow > enow) == (fnow > -enow)) {
        orc_two_sum(Q, enow, &Qnew, &hh); enow = ++ei < elen ? e[ei] : 0.0;
      } else {
        orc_two_sum(Q, fnow, &Qnew, &hh); fnow = ++fi < flen ? f[fi] : 0.0;
      }
      Q = Qnew;
      if (hh != 0.0) h[hi++] = hh;
    }
  }
  while (ei < elen) {
    orc_two_sum(Q, enow, &Qnew, &hh); enow = ++ei < elen ? e[ei] : 0.0;
    Q = Qnew;
    if (hh != 0.0) h[hi++] = hh;
  }
  while (fi < flen) {
    orc_two_sum(Q, fnow, &Qnew, &hh); fnow = ++fi < flen ? f[fi] : 0.0;
    Q = Qnew;
    if (hh != 0.0) h[hi++] = hh;
  }
  if (Q != 0.0 || hi == 0) h[hi++] = Q;
  return hi;
}
/* (a1*b1 - a2*b2) of doubles as the 4-component expansion of the two exact
 * products' difference */
static void orc_prod_diff(double a1, double b1, double a2, double b2, double *u) {
  double s1, s0, t1, t0;
  orc_two_product(a1, b1, &s1, &s0);
  orc_two_product(a2, b2, &t1, &t0);
  orc_two_two_diff(s1, s0, t1, t0, u);
}
static double orc_ccw_adapt(const double *pa, const double *pb, const double *pc,
                            double detsum) {
  double acx = pa[0] - pc[0], bcx = pb[0] - pc[0];
  double acy = pa[1] - pc[1], bcy = pb[1] - pc[1];
  double B[4], C1[8], C2[12], D[16], u[4];
  orc_prod_diff(acx, bcy, acy, bcx, B);
  double det = B[0] + B[1] + B[2] + B[3];
  double err = ORC_CCW_B * detsum;
  if (det >= err || -det >= err) return det;
  double acxt = orc_diff_tail(pa[0], pc[0], acx), bcxt = orc_diff_tail(pb[0], pc[0], bcx);
  double acyt = orc_diff_tail(pa[1], pc[1], acy), bcyt = orc_diff_tail(pb[1], pc[1], bcy);
  if (acxt == 0.0 && acyt == 0.0 && bcxt == 0.0 && bcyt == 0.0) return det;
  err = ORC_CCW_C * detsum + ORC_RESULT_ERR * fabs(det);
  det += (acx * bcyt + bcy * acxt) - (acy * bcxt + bcx * acyt);
  if (det >= err || -det >= err) return det;
  orc_prod_diff(acxt, bcy, acyt, bcx, u);
  int n1 = orc_fast_expansion_sum(4, B, 4, u, C1);
  orc_prod_diff(acx, bcyt, acy, bcxt, u);
  int n2 = orc_fast_expansion_sum(n1, C1, 4, u, C2);
  orc_prod_diff(acxt, bcyt, acyt, bcxt, u);
  int n3 = orc_fast_expansion_sum(n2, C2, 4, u, D);
  return D[n3 - 1];
}
/* counterclockwise(pa, pb, pc) of Triangle, value included. */
double orc_ccw_value(const double *pa, const double *pb, const double *pc) {
  double dl = (pa[0] - pc[0]) * (pb[1] - pc[1]);
  double dr = (pa[1] - pc[1]) * (pb[0] - pc[0]);
  double det = dl - dr, detsum;
  if (dl > 0.0) {
    if (dr <= 0.0) return det;
    detsum = dl + dr;
  } else if (dl < 0.0) {
    if (dr >= 0.0) return det;
    detsum = -dl - dr;
  } else {
    return det;
  }
  double err = ORC_CCW_A * detsum;
  if (det >= err || -det >= err) return det;
  return orc_ccw_adapt(pa, pb, pc, detsum);
}

/* findcircumcenter / computeCircumcenter: triangle.cpp:6523-6618,
 * refine.cu:3780-3868.  The denominator is Triangle's adaptive
 * counterclockwise(dest, apex, org) (its sign is exact and its value
 * accurate for near-collinear triangles, where the raw determinant loses
 * every digit).  Returns 0 if the triangle is degenerate (collinear). */
static int steiner_point(const Mesh *m, int t, double *c) {
  const int *v = m->TV + 3 * (size_t)t;
  const double *A = PT(m, v[0]), *O = PT(m, v[1]), *D = PT(m, v[2]);
  double xdo = D[0] - O[0], ydo = D[1] - O[1];
  double xao = A[0] - O[0], yao = A[1] - O[1];
  double dodist = xdo * xdo + ydo * ydo;
  double aodist = xao * xao + yao * yao;
  double dadist = (D[0] - A[0]) * (D[0] - A[0]) + (D[1] - A[1]) * (D[1] - A[1]);
  double det = orc_ccw_value(D, A, O);
  if (det == 0.0) return 0;
  double denominator = 0.5 / det;
  double dx = (yao * dodist - ydo * aodist) * denominator;
  double dy = (xdo * aodist - xao * dodist) * denominator;
  if (m->offcenter && m->offconst > 0.0) {
    double dxoff, dyoff;
    if ((dodist < aodist) && (dodist < dadist)) {
      dxoff = 0.5 * xdo - m->offconst * ydo;
      dyoff = 0.5 * ydo + m->offconst * xdo;
      if (dxoff * dxoff + dyoff * dyoff < dx * dx + dy * dy) { dx = dxoff; dy = dyoff; }
    } else if (aodist < dadist) {
      dxoff = 0.5 * xao + m->offconst * yao;
      dyoff = 0.5 * yao - m->offconst * xao;
      if (dxoff * dxoff + dyoff * dyoff < dx * dx + dy * dy) { dx = dxoff; dy = dyoff; }
    } else {
      dxoff = 0.5 * (A[0] - D[0]) - m->offconst * (A[1] - D[1]);
      dyoff = 0.5 * (A[1] - D[1]) + m->offconst * (A[0] - D[0]);
      if (dxoff * dxoff + dyoff * dyoff <
          (dx - xdo) * (dx - xdo) + (dy - ydo) * (dy - ydo)) {
        dx = xdo + dxoff;
        dy = ydo + dyoff;
      }
    }
  }
  c[0] = O[0] + dx;
  c[1] = O[1] + dy;
  if ((c[0] == A[0] && c[1] == A[1]) || (c[0] == O[0] && c[1] == O[1]) ||
      (c[0] == D[0] && c[1] == D[1]))
    return 0;
  if (!(c[0] == c[0]) || !(c[1] == c[1])) return 0;
  return 1;
}

/* Locate the adjacent triangle(s) of subsegment s: t1 holds the directed
 * edge org->dest of the subsegment if it exists, t2 the reverse. */
static void seg_sides(const Mesh *m, int s, int *t1, int *j1, int *t2, int *j2) {
  int a = m->SV[2 * s];
  int ot = m->ST[s];
  int t = ot >> 2, j = ot & 3;
  int u = m->TV[3 * t + (j + 1) % 3];
  int n = m->TN[3 * t + j];
  int ta = t, ja = j, tb = -1, jb = -1;
  if (n >= 0) { tb = n >> 2; jb = n & 3; }
  if (u == a) { *t1 = ta; *j1 = ja; *t2 = tb; *j2 = jb; }
  else { *t1 = tb; *j1 = jb; *t2 = ta; *j2 = ja; }
}

static int split_point_xy(const double *A, const double *B, int acuteorg, int acutedest,
                          double *p);

/* splitencsegs split point: triangle.cpp:13290-13425, refine.cu:2686-2725.
 * Returns 0 if the point coincides with an endpoint (precision exhausted). */
static int split_point(const Mesh *m, int s, double *p) {
  int a = m->SV[2 * s], b = m->SV[2 * s + 1];
  int t1, j1, t2, j2;
  seg_sides(m, s, &t1, &j1, &t2, &j2);
  int acuteorg = 0, acutedest = 0;
  /* side where the edge runs a->b: the other edge at a is (j+2)%3, at b (j+1)%3 */
  if (t1 >= 0) {
    if (m->TS[3 * t1 + (j1 + 2) % 3] >= 0) acuteorg = 1;
    if (m->TS[3 * t1 + (j1 + 1) % 3] >= 0) acutedest = 1;
  }
  /* side where the edge runs b->a: the other edge at a is (j+1)%3, at b (j+2)%3 */
  if (t2 >= 0) {
    if (m->TS[3 * t2 + (j2 + 1) % 3] >= 0) acuteorg = 1;
    if (m->TS[3 * t2 + (j2 + 2) % 3] >= 0) acutedest = 1;
  }
  return split_point_xy(PT(m, a), PT(m, b), acuteorg, acutedest, p);
}

/* The split point of segment A-B: the midpoint, or, when an end is shared
 * with another subsegment (acuteorg / acutedest), a point at a power-of-two
 * distance from that end (concentric shells; the destination wins when both
 * are shared), then moved onto the segment's line by one rounding
 * correction (triangle.cpp:13368-13415). */
int orc_split_point_xy(const double *A, const double *B, int acuteorg, int acutedest,
                       double *p) {
  return split_point_xy(A, B, acuteorg, acutedest, p);
}

static int split_point_xy(const double *A, const double *B, int acuteorg, int acutedest,
                          double *p) {
  double split;
  if (acuteorg || acutedest) {
    double len = sqrt((B[0] - A[0]) * (B[0] - A[0]) + (B[1] - A[1]) * (B[1] - A[1]));
    double npt = 1.0;
    while (len > 3.0 * npt) npt *= 2.0;
    while (len < 1.5 * npt) npt *= 0.5;
    split = npt / len;
    if (acutedest) split = 1.0 - split;
  } else {
    split = 0.5;
  }
  p[0] = A[0] + split * (B[0] - A[0]);
  p[1] = A[1] + split * (B[1] - A[1]);
  double mult = (A[0] - p[0]) * (B[1] - p[1]) - (A[1] - p[1]) * (B[0] - p[0]);
  double div = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]);
  if (mult != 0.0 && div != 0.0) {
    mult = mult / div;
    if (mult == mult) {
      p[0] += mult * (B[1] - A[1]);
      p[1] += mult * (A[0] - B[0]);
    }
  }
  if ((p[0] == A[0] && p[1] == A[1]) || (p[0] == B[0] && p[1] == B[1])) return 0;
  return 1;
}

/* Half-octave bucket of the squared shortest edge: 2*e + (f >= sqrt(1/2))
 * with key = f * 2^e, 0.5 <= f < 1 (exact via frexp).  Shorter edges give
 * smaller buckets, i.e. higher priority, as in Triangle's queue. */
static int edge_bucket(double key) {
  int e;
  double f = frexp(key, &e);
  int b = 2 * e + (f >= 0.70710678118654752440 ? 1 : 0);
  b += 2200;
  if (b < 0) b = 0;
  if (b > 8191) b = 8191;
  return b;
}

static double shortest_edge_sq(const Mesh *m, int t) {
  const int *v = m->TV + 3 * (size_t)t;
  const double *A = PT(m, v[0]), *O = PT(m, v[1]), *D = PT(m, v[2]);
  double l0 = (O[0] - D[0]) * (O[0] - D[0]) + (O[1] - D[1]) * (O[1] - D[1]);
  double l1 = (D[0] - A[0]) * (D[0] - A[0]) + (D[1] - A[1]) * (D[1] - A[1]);
  double l2 = (A[0] - O[0]) * (A[0] - O[0]) + (A[1] - O[1]) * (A[1] - O[1]);
  double mn = l0 < l1 ? l0 : l1;
  if (l2 < mn) mn = l2;
  return mn;
}

static int shortest_edge_bucket(const Mesh *m, int t) {
  return edge_bucket(shortest_edge_sq(m, t));
}

static int cmp_int_c(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return x < y ? -1 : (x > y);
}

static int in_list(const int *L, int n, int x) {
  for (int i = 0; i < n; i++)
    if (L[i] == x) return 1;
  return 0;
}

/* Is edge j of cavity triangle t on the cavity boundary? */
static int is_bnd(const Mesh *m, int t, int j, int split_s, const int *L, int n) {
  int s = m->TS[3 * t + j];
  int nb = m->TN[3 * t + j];
  if (s >= 0 && s != split_s) return 1;
  if (nb < 0) return 1;
  return !in_list(L, n, nb >> 2);
}

/* Walk toward c.  Result codes: 0 inside t, 1 on edge *j, 2 on a vertex,
 * 3 blocked by subsegment *seg, 4 failed.  refine.cu:3908-4184 semantics. */
static int walk(const Mesh *m, int t, const double *c, int *tout, int *jout,
                int *seg) {
  for (long steps = 0; steps < WALK_LIMIT; steps++) {
    const int *v = m->TV + 3 * (size_t)t;
    int o[3];
    for (int j = 0; j < 3; j++)
      o[j] = orc_orient2d(PT(m, v[(j + 1) % 3]), PT(m, v[(j + 2) % 3]), c);
    int jn = -1;
    for (int j = 0; j < 3; j++)
      if (o[j] < 0) { jn = j; break; }
    if (jn < 0) {
      int zeros = (o[0] == 0) + (o[1] == 0) + (o[2] == 0);
      *tout = t;
      if (zeros == 0) return 0;
      if (zeros == 1) {
        *jout = (o[0] == 0) ? 0 : (o[1] == 0) ? 1 : 2;
        return 1;
      }
      return 2;
    }
    int s = m->TS[3 * t + jn];
    if (s >= 0) { *seg = s; *tout = t; return 3; }
    int nb = m->TN[3 * t + jn];
    if (nb < 0) return 4;
    t = nb >> 2;
  }
  return 4;
}

typedef struct {
  int id;       /* triangle (bad pass) or subsegment (segment pass) */
  int prio;     /* smaller wins; ties by candidate index */
  int status;
  double c[2];
  int n, r;     /* cavity size, ring size */
  int *L;       /* cavity then ring */
} Cand;

/* Grow the Bowyer-Watson cavity of c from the seeds; then collect the ring
 * of triangles across the boundary, and check that the cavity is a disk
 * whose star from c is positively oriented. */
static int build_cavity(const Mesh *m, Cand *cd, int seed0, int seed1,
                        int split_s, int K) {
  int *L = cd->L;
  int n = 0;
  L[n++] = seed0;
  if (seed1 >= 0) L[n++] = seed1;
  for (int idx = 0; idx < n; idx++) {
    int t = L[idx];
    for (int j = 0; j < 3; j++) {
      int s = m->TS[3 * t + j];
      if (s >= 0) continue; /* subsegments (and the split one) are never crossed */
      int nb = m->TN[3 * t + j];
      if (nb < 0) continue;
      int u = nb >> 2;
      if (in_list(L, n, u)) continue;
      const int *w = m->TV + 3 * (size_t)u;
      if (orc_incircle(PT(m, w[0]), PT(m, w[1]), PT(m, w[2]), cd->c) > 0) {
        if (n == K) return ST_OVERFLOW;
        L[n++] = u;
      }
    }
  }
  /* the cavity is a set: keep it sorted by triangle index (slot order) */
  qsort(L, (size_t)n, sizeof(int), cmp_int_c);
  /* disk check: boundary edges == n + 2; no subsegment strictly inside */
  int nbnd = 0;
  for (int idx = 0; idx < n; idx++) {
    int t = L[idx];
    for (int j = 0; j < 3; j++) {
      int s = m->TS[3 * t + j];
      int nb = m->TN[3 * t + j];
      if (s >= 0 && s != split_s && nb >= 0 && in_list(L, n, nb >> 2))
        return ST_GIVEUP;
      if (is_bnd(m, t, j, split_s, L, n)) {
        nbnd++;
        /* the star triangle (c, u, v) must be positively oriented, except
         * for the split subsegment's own edge on a one-sided split */
        if (!(s >= 0 && s == split_s)) {
          const int *v = m->TV + 3 * (size_t)t;
          if (orc_orient2d(cd->c, PT(m, v[(j + 1) % 3]), PT(m, v[(j + 2) % 3])) <= 0)
            return ST_GIVEUP;
        }
      }
    }
  }
  if (nbnd != n + 2) return ST_GIVEUP;
  /* ring: distinct triangles across boundary edges */
  int r = 0;
  for (int idx = 0; idx < n; idx++) {
    int t = L[idx];
    for (int j = 0; j < 3; j++) {
      if (!is_bnd(m, t, j, split_s, L, n)) continue;
      int nb = m->TN[3 * t + j];
      if (nb < 0) continue;
      int u = nb >> 2;
      if (in_list(L + n, r, u)) continue;
      L[n + r++] = u;
    }
  }
  cd->n = n;
  cd->r = r;
  return ST_VALID;
}

/* Next boundary edge (counter-clockwise around the cavity) after boundary
 * edge (t,j): pivot around its destination through cavity triangles. */
static void next_bnd(const Mesh *m, int t, int j, int split_s, const int *L,
                     int n, int *tn, int *jn) {
  int cur = t, k = (j + 1) % 3;
  for (;;) {
    if (is_bnd(m, cur, k, split_s, L, n)) { *tn = cur; *jn = k; return; }
    int nb = m->TN[3 * cur + k];
    cur = nb >> 2;
    k = ((nb & 3) + 1) % 3;
  }
}

typedef struct {
  int u, v, outer, seg;
} Bnd;

/* Retriangulate a winner's cavity as the star of its new vertex.
 * Star ordering: closed fans start at the boundary edge whose origin has the
 * smallest vertex id; the open fan of a one-sided segment split starts at the
 * boundary edge after the split subsegment's edge.  Star triangle i is
 * (p, u_i, v_i) and takes slot L[i] (cavity triangles in ascending index
 * order) for i < n, else new slot tbase + i - n. */
static void apply_star(Mesh *m, const Cand *cd, int split_s, int p, int tbase,
                       int snew, int is_seg) {
  const int *L = cd->L;
  int n = cd->n;
  int k = n + 2;
  Bnd *B = malloc(sizeof(Bnd) * (size_t)k);
  int t0 = -1, j0 = -1;
  int open = 0;
  const int inst = m->ipar ? m->tinst[L[0]] : 0;
  if (is_seg) {
    int t1, j1, t2, j2;
    seg_sides(m, split_s, &t1, &j1, &t2, &j2);
    if (t1 < 0 || t2 < 0) {
      open = 1;
      int ts = t1 >= 0 ? t1 : t2, js = t1 >= 0 ? j1 : j2;
      next_bnd(m, ts, js, split_s, L, n, &t0, &j0);
    }
  }
  if (!open) {
    int best = -1;
    for (int idx = 0; idx < n; idx++) {
      int t = L[idx];
      for (int j = 0; j < 3; j++) {
        if (!is_bnd(m, t, j, split_s, L, n)) continue;
        int u = m->TV[3 * t + (j + 1) % 3];
        if (best < 0 || u < best) { best = u; t0 = t; j0 = j; }
      }
    }
  }
  int nstar = open ? k - 1 : k;
  int t = t0, j = j0;
  for (int i = 0; i < nstar; i++) {
    B[i].u = m->TV[3 * t + (j + 1) % 3];
    B[i].v = m->TV[3 * t + (j + 2) % 3];
    B[i].outer = m->TN[3 * t + j];
    B[i].seg = m->TS[3 * t + j];
    int tn, jn;
    next_bnd(m, t, j, split_s, L, n, &tn, &jn);
    t = tn;
    j = jn;
  }
  int a = -1, b = -1, sA = -1, sB = -1;
  if (is_seg) {
    a = m->SV[2 * split_s];
    b = m->SV[2 * split_s + 1];
    sA = split_s;
    sB = snew;
  }
  int *slot = malloc(sizeof(int) * (size_t)nstar);
  for (int i = 0; i < nstar; i++) slot[i] = i < n ? L[i] : tbase + (i - n);
  for (int i = 0; i < nstar; i++) {
    int st = slot[i];
    int nx = (i + 1 < nstar) ? i + 1 : (open ? -1 : 0);
    int pv = (i > 0) ? i - 1 : (open ? -1 : nstar - 1);
    m->TV[3 * st + 0] = p;
    m->TV[3 * st + 1] = B[i].u;
    m->TV[3 * st + 2] = B[i].v;
    m->TN[3 * st + 0] = B[i].outer;
    m->TN[3 * st + 1] = nx >= 0 ? (slot[nx] << 2) | 2 : -1;
    m->TN[3 * st + 2] = pv >= 0 ? (slot[pv] << 2) | 1 : -1;
    m->TS[3 * st + 0] = B[i].seg;
    m->TS[3 * st + 1] = (B[i].v == a) ? sA : (B[i].v == b) ? sB : -1;
    m->TS[3 * st + 2] = (B[i].u == a) ? sA : (B[i].u == b) ? sB : -1;
    m->tflag[st] = TF_ALIVE | TF_CHECK;
    if (m->ipar) m->tinst[st] = inst;
  }
  if (open) { /* the second reserved slot stays unused */
    int st = tbase + 1;
    m->TV[3 * st] = m->TV[3 * st + 1] = m->TV[3 * st + 2] = -1;
    m->TN[3 * st] = m->TN[3 * st + 1] = m->TN[3 * st + 2] = -1;
    m->TS[3 * st] = m->TS[3 * st + 1] = m->TS[3 * st + 2] = -1;
    m->tflag[st] = 0;
  }
  for (int i = 0; i < nstar; i++) {
    int st = slot[i];
    if (B[i].outer >= 0) m->TN[3 * (B[i].outer >> 2) + (B[i].outer & 3)] = (st << 2) | 0;
    if (B[i].seg >= 0) {
      m->ST[B[i].seg] = (st << 2) | 0;
      if (is_seg && m->senc[B[i].seg] == 0 &&
          encroaches(m, m->SV[2 * B[i].seg], m->SV[2 * B[i].seg + 1], PT(m, p)))
        m->senc[B[i].seg] = 1;
    }
  }
  if (is_seg) {
    int par = m->SP[split_s];
    m->SV[2 * sA + 0] = a;
    m->SV[2 * sA + 1] = p;
    m->SV[2 * sB + 0] = p;
    m->SV[2 * sB + 1] = b;
    m->SP[sB] = par;
    int stA = -1, stB = -1, encA = 0, encB = 0;
    for (int i = 0; i < nstar; i++) {
      int st = slot[i];
      if (B[i].u == a) { if (stA < 0) stA = (st << 2) | 2; encA |= encroaches(m, a, p, PT(m, B[i].v)); }
      if (B[i].v == a) { if (stA < 0) stA = (st << 2) | 1; encA |= encroaches(m, a, p, PT(m, B[i].u)); }
      if (B[i].u == b) { if (stB < 0) stB = (st << 2) | 2; encB |= encroaches(m, p, b, PT(m, B[i].v)); }
      if (B[i].v == b) { if (stB < 0) stB = (st << 2) | 1; encB |= encroaches(m, p, b, PT(m, B[i].u)); }
    }
    m->ST[sA] = stA;
    m->ST[sB] = stB;
    m->senc[sA] = (uint8_t)encA;
    m->senc[sB] = (uint8_t)encB;
  }
  free(slot);
  free(B);
}

/* Mark subsegment s encroached unless it was given up on (senc == 2). */
static void mark_enc(Mesh *m, int s, long long *marks) {
  if (m->senc[s] == 0) {
    m->senc[s] = 1;
    (*marks)++;
  }
}

/* A bad triangle's Steiner point blocked by subsegment s (walk): mark s and
 * reject the candidate, or give the triangle up when s was given up. */
static int blocked_by(Mesh *m, int s, long long *marks) {
  if (m->senc[s] == 2) return ST_GIVEUP;
  mark_enc(m, s, marks);
  return ST_REJ_BLOCKED;
}

/* ---- vertex deletion ----------------------------------------------------
 * Chew's rule as Triangle applies it before splitting an encroached
 * subsegment (triangle.cpp:13319-13357; the reference marks the same
 * vertices, refine.cu:604-677): while the apex of the triangle on either side
 * of the subsegment is a free vertex (inserted at a circumcenter or
 * off-center, not on a segment) strictly inside the subsegment's diametral
 * circle, that vertex is deleted.  Bulk-synchronous form: in a pass with
 * subsegment candidates, every one of the first DEL_CANDS nominates its two
 * side apexes; a nominee whose star is closed, has at most DEL_MAX triangles
 * and no subsegment at the vertex, and whose link polygon the ear rule below
 * retriangulates, is a deletion candidate (index 2 i + side).  Candidates
 * claim their star and the outer triangles across the link; the smallest
 * index wins each triangle and a candidate holding all of its triangles
 * deletes its vertex.  The pass's cavities are built afterwards, so a
 * subsegment splits in the same pass as its apexes are deleted (Triangle
 * deletes, then splits); bad-triangle candidates whose triangle was
 * retriangulated skip the pass.  One apex per side and pass is deleted,
 * where Triangle repeats on the new apex until none qualifies.
 *
 * Retriangulation of the link polygon p[0..k) (counterclockwise): repeatedly
 * cut the first ear (p[q-1], p[q], p[q+1]) of the remaining polygon that
 * turns left and whose circumcircle holds no other remaining vertex strictly
 * inside, until a triangle is left.  Each cut triangle is a Delaunay triangle
 * of the link's vertices, which with the vertex gone are the constrained
 * Delaunay triangles inside the star (the mesh outside the star does not
 * change).  No ear found: the vertex is kept. */
#define DEL_MAX 32
#define DEL_CANDS (1 << 16)

typedef struct {
  int v, k;
  int star[DEL_MAX];          /* star[i] = (v, p[i], p[i+1]) */
  int p[DEL_MAX], o[DEL_MAX], s[DEL_MAX]; /* link, outer otri, subsegment of edge i */
  int nt[3 * (DEL_MAX - 2)];  /* new triangles, counterclockwise, cut order */
} DelCand;

/* Apex of triangle t (corner j) on subsegment s: a deletion candidate? */
static int del_candidate(const Mesh *m, int s, int t, int j, DelCand *d) {
  const int v = m->TV[3 * t + j];
  if ((m->pflag[v] & (PF_FREE | PF_DEAD)) != PF_FREE) return 0;
  const double *O = PT(m, m->SV[2 * s]), *D = PT(m, m->SV[2 * s + 1]), *V = PT(m, v);
  if (!((O[0] - V[0]) * (D[0] - V[0]) + (O[1] - V[1]) * (D[1] - V[1]) < 0.0)) return 0;
  /* star, counterclockwise around v from t */
  int cur = t, jc = j, k = 0;
  for (;;) {
    if (k == DEL_MAX) return 0;
    const int x = m->TV[3 * cur + (jc + 1) % 3], y = m->TV[3 * cur + (jc + 2) % 3];
    if (k == 0) d->p[0] = x;
    else if (d->p[k] != x) return 0;
    d->star[k] = cur;
    d->o[k] = m->TN[3 * cur + jc];
    d->s[k] = m->TS[3 * cur + jc];
    /* edge (v, y) is opposite x */
    if (m->TS[3 * cur + (jc + 1) % 3] >= 0) return 0;
    const int nb = m->TN[3 * cur + (jc + 1) % 3];
    if (nb < 0) return 0;
    k++;
    const int nx = nb >> 2;
    if (nx == t) {
      if (y != d->p[0]) return 0;
      break;
    }
    if (k < DEL_MAX) d->p[k] = y;
    int jn = -1;
    for (int q = 0; q < 3; q++)
      if (m->TV[3 * nx + q] == v) jn = q;
    if (jn < 0) return 0;
    cur = nx;
    jc = jn;
  }
  if (k < 3) return 0;
  /* ear cutting */
  int rem[DEL_MAX], r = k, ne = 0;
  for (int q = 0; q < k; q++) rem[q] = q;
  while (r > 3) {
    int found = -1;
    for (int q = 0; q < r && found < 0; q++) {
      const int qa = (q + r - 1) % r, qc = (q + 1) % r;
      const double *A = PT(m, d->p[rem[qa]]), *B = PT(m, d->p[rem[q]]), *C = PT(m, d->p[rem[qc]]);
      if (orc_orient2d(A, B, C) <= 0) continue;
      int ok = 1;
      for (int w = 0; w < r && ok; w++) {
        if (w == qa || w == q || w == qc) continue;
        if (orc_incircle(A, B, C, PT(m, d->p[rem[w]])) > 0) ok = 0;
      }
      if (ok) found = q;
    }
    if (found < 0) return 0;
    d->nt[3 * ne] = d->p[rem[(found + r - 1) % r]];
    d->nt[3 * ne + 1] = d->p[rem[found]];
    d->nt[3 * ne + 2] = d->p[rem[(found + 1) % r]];
    ne++;
    for (int q = found; q + 1 < r; q++) rem[q] = rem[q + 1];
    r--;
  }
  if (orc_orient2d(PT(m, d->p[rem[0]]), PT(m, d->p[rem[1]]), PT(m, d->p[rem[2]])) <= 0) return 0;
  d->nt[3 * ne] = d->p[rem[0]];
  d->nt[3 * ne + 1] = d->p[rem[1]];
  d->nt[3 * ne + 2] = d->p[rem[2]];
  d->v = v;
  d->k = k;
  return 1;
}

/* Replace a winner's star by the cut triangles: they take the star's
 * slots in ascending order, the two largest slots die. */
static void del_apply(Mesh *m, const DelCand *d) {
  const int k = d->k;
  int sl[DEL_MAX];
  for (int q = 0; q < k; q++) sl[q] = d->star[q];
  for (int q = 1; q < k; q++) {
    int x = sl[q], w = q - 1;
    while (w >= 0 && sl[w] > x) { sl[w + 1] = sl[w]; w--; }
    sl[w + 1] = x;
  }
  const int inst = m->ipar ? m->tinst[d->star[0]] : 0;
  for (int q = 0; q < k - 2; q++) {
    const int st = sl[q];
    const int *V = d->nt + 3 * q;
    for (int e = 0; e < 3; e++) {
      const int u = V[(e + 1) % 3], w = V[(e + 2) % 3];
      int nb = -2, sg = -1;
      for (int i = 0; i < k && nb == -2; i++)
        if (d->p[i] == u && d->p[(i + 1) % k] == w) {
          nb = d->o[i];
          sg = d->s[i];
          if (nb >= 0) m->TN[3 * (nb >> 2) + (nb & 3)] = (st << 2) | e;
          if (sg >= 0) m->ST[sg] = (st << 2) | e;
        }
      for (int q2 = 0; q2 < k - 2 && nb == -2; q2++) {
        if (q2 == q) continue;
        for (int e2 = 0; e2 < 3; e2++)
          if (d->nt[3 * q2 + (e2 + 1) % 3] == w && d->nt[3 * q2 + (e2 + 2) % 3] == u) {
            nb = (sl[q2] << 2) | e2;
            break;
          }
      }
      m->TN[3 * st + e] = nb;
      m->TS[3 * st + e] = sg;
    }
    m->TV[3 * st] = V[0];
    m->TV[3 * st + 1] = V[1];
    m->TV[3 * st + 2] = V[2];
    m->tflag[st] = TF_ALIVE | TF_CHECK;
    if (m->ipar) m->tinst[st] = inst;
  }
  for (int q = k - 2; q < k; q++) {
    const int st = sl[q];
    m->TV[3 * st] = m->TV[3 * st + 1] = m->TV[3 * st + 2] = -1;
    m->TN[3 * st] = m->TN[3 * st + 1] = m->TN[3 * st + 2] = -1;
    m->TS[3 * st] = m->TS[3 * st + 1] = m->TS[3 * st + 2] = -1;
    m->tflag[st] = 0;
  }
  m->pflag[d->v] |= PF_DEAD;
}

/* The deletion step of a pass over subsegment candidates ids[0..nseg);
 * returns the number of vertices deleted. */
static long long delete_step(Mesh *m, const int *ids, int nseg, int policy, orc_stats *st) {
  const int ns2 = nseg < DEL_CANDS ? nseg : DEL_CANDS;
  const int nd = 2 * ns2;
  if (nd == 0) return 0;
  DelCand *dc = malloc(sizeof(DelCand) * (size_t)nd);
  uint8_t *ok = calloc((size_t)nd, 1);
  for (int i = 0; i < ns2; i++) {
    const int s = ids[i];
    int t1, j1, t2, j2;
    seg_sides(m, s, &t1, &j1, &t2, &j2);
    for (int side = 0; side < 2; side++) {
      const int t = side ? t2 : t1, j = side ? j2 : j1;
      if (t < 0) continue;
      use_inst(m, t);
      if (!(policy == 2 || (policy == 1 && !m->ruppert))) continue;
      if (del_candidate(m, s, t, j, &dc[2 * i + side])) ok[2 * i + side] = 1;
    }
  }
  int *claim = malloc(sizeof(int) * (size_t)m->nt);
  for (int t = 0; t < m->nt; t++) claim[t] = INT32_MAX;
  for (int d = 0; d < nd; d++) {
    if (!ok[d]) continue;
    for (int q = 0; q < dc[d].k; q++) {
      int t = dc[d].star[q];
      if (claim[t] > d) claim[t] = d;
      if (dc[d].o[q] >= 0) {
        t = dc[d].o[q] >> 2;
        if (claim[t] > d) claim[t] = d;
      }
    }
  }
  long long ndel = 0;
  uint8_t *win = calloc((size_t)nd, 1);
  for (int d = 0; d < nd; d++) {
    if (!ok[d]) continue;
    int w = 1;
    for (int q = 0; q < dc[d].k && w; q++)
      w = claim[dc[d].star[q]] == d && (dc[d].o[q] < 0 || claim[dc[d].o[q] >> 2] == d);
    win[d] = (uint8_t)w;
  }
  for (int d = 0; d < nd; d++)
    if (win[d]) {
      del_apply(m, &dc[d]);
      ndel++;
      st->flips += dc[d].k - 3;
    }
  st->deleted += ndel;
  free(win);
  free(claim);
  free(ok);
  free(dc);
  return ndel;
}

static int cmp_ll(const void *a, const void *b) {
  long long x = *(const long long *)a, y = *(const long long *)b;
  return x < y ? -1 : (x > y);
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return x < y ? -1 : (x > y);
}

static int cands_per_pass(long long budget, int K) {
  long long c = budget / (6LL * K + 16);
  if (c < 1) c = 1;
  if (c > (1LL << 30)) c = 1LL << 30;
  return (int)c;
}

/* One conflict-resolution pass over candidates ids[0..nc): candidates below
 * nseg split subsegments, the rest are bad triangles.  Returns the number of
 * winners (*nwin_seg of them segment splits); sets *overflowed if any cavity
 * exceeded K. */
static int run_pass(Mesh *m, const int *ids, int nc, int nseg, int K,
                    int Kmax, int priority, int claim_iters, int shared_ring, int deletion,
                    orc_stats *st, int *overflowed, long long *progress, int *nwin_seg) {
  long long marks = 0, gave = st->giveups;
  Cand *cd = calloc((size_t)nc, sizeof(Cand));
  int *pool = malloc(sizeof(int) * (size_t)(2 * K + 2));
  const long long ndel = deletion ? delete_step(m, ids, nseg, deletion, st) : 0;
  *overflowed = 0;
  for (int i = 0; i < nc; i++) {
    Cand *c = &cd[i];
    const int is_seg = i < nseg;
    c->id = ids[i];
    c->L = pool; /* temporary; copied out below if valid */
    int seed0 = -1, seed1 = -1, split_s = -1;
    /* skipped: a bad triangle whose slot the deletion step rewrote */
    if (!is_seg && (m->tflag[ids[i]] & (TF_ALIVE | TF_BAD)) != (TF_ALIVE | TF_BAD)) {
      c->status = ST_DEFER;
      goto done;
    }
    c->prio = (priority == 1 && !is_seg) ? shortest_edge_bucket(m, ids[i]) : 0;
    if (is_seg) {
      int s = ids[i];
      split_s = s;
      int t1, j1, t2, j2;
      seg_sides(m, s, &t1, &j1, &t2, &j2);
      use_inst(m, t1 >= 0 ? t1 : t2);
      if (!split_point(m, s, c->c)) { c->status = ST_GIVEUP; goto done; }
      seed0 = t1 >= 0 ? t1 : t2;
      seed1 = t1 >= 0 ? t2 : -1;
    } else {
      int t0 = ids[i];
      use_inst(m, t0);
      if (!steiner_point(m, t0, c->c)) { c->status = ST_GIVEUP; goto done; }
      int tl = -1, jl = -1, sg = -1;
      int w = walk(m, t0, c->c, &tl, &jl, &sg);
      /* blocked by a subsegment: mark it, or give up when it was given up
       * (it will never be split, so the triangle could never be) */
      if (w == 3) { c->status = blocked_by(m, sg, &marks); goto done; }
      if (w == 2 || w == 4) { c->status = ST_GIVEUP; goto done; }
      seed0 = tl;
      if (w == 1) {
        int s = m->TS[3 * tl + jl];
        if (s >= 0) { c->status = blocked_by(m, s, &marks); goto done; }
        int nb = m->TN[3 * tl + jl];
        if (nb < 0) { c->status = ST_GIVEUP; goto done; }
        seed1 = nb >> 2;
      }
    }
    c->status = build_cavity(m, c, seed0, seed1, split_s, K);
    if (c->status == ST_OVERFLOW && K >= Kmax) c->status = ST_GIVEUP;
    if (c->status == ST_VALID && !is_seg) {
      /* A Steiner point within 1e-6 of the triangle's shortest edge of a
       * cavity vertex coincides with it up to rounding (in exact arithmetic
       * the empty circumcircle keeps every vertex at least half the shortest
       * edge away): the triangle is given up instead of seeding a cascade of
       * rounding-scale triangles (nearly cocircular inputs). */
      double near2 = 1e-12 * shortest_edge_sq(m, c->id);
      for (int idx = 0; idx < c->n && c->status == ST_VALID; idx++) {
        const int *v = m->TV + 3 * (size_t)c->L[idx];
        for (int j = 0; j < 3; j++) {
          const double *q = PT(m, v[j]);
          double dx = c->c[0] - q[0], dy = c->c[1] - q[1];
          if (dx * dx + dy * dy < near2) { c->status = ST_GIVEUP; break; }
        }
      }
    }
    if (c->status == ST_VALID && !is_seg) {
      /* Ruppert/Chew rule: a Steiner point that encroaches a subsegment on
       * its cavity boundary is not inserted; every such subsegment is marked.
       * When every subsegment it encroaches was given up, the triangle is
       * given up too. */
      int enc = 0, live = 0;
      for (int idx = 0; idx < c->n; idx++) {
        int t = c->L[idx];
        for (int j = 0; j < 3; j++) {
          int s = m->TS[3 * t + j];
          if (s < 0) continue;
          if (encroaches(m, m->SV[2 * s], m->SV[2 * s + 1], c->c)) {
            live |= m->senc[s] != 2;
            mark_enc(m, s, &marks);
            enc = 1;
          }
        }
      }
      if (enc) c->status = live ? ST_REJ_ENC : ST_GIVEUP;
    }
  done:
    if (c->status == ST_VALID) {
      int *Lc = malloc(sizeof(int) * (size_t)(c->n + c->r));
      memcpy(Lc, c->L, sizeof(int) * (size_t)(c->n + c->r));
      c->L = Lc;
    } else {
      c->L = NULL;
      if (c->status == ST_OVERFLOW) { *overflowed = 1; st->rej_overflow++; }
      else if (c->status == ST_REJ_ENC) st->rej_encroach++;
      else if (c->status == ST_REJ_BLOCKED) st->rej_blocked++;
      else if (c->status == ST_GIVEUP) {
        st->giveups++;
        if (is_seg) m->senc[c->id] = 2;
        else m->tflag[c->id] |= TF_GIVEUP;
      }
    }
  }
  /* claims: the smallest (priority, candidate index) wins each triangle.
   * Repeated claim_iters times: candidates that neither won nor touch a
   * winner's region compete again among themselves. */
  /* claim[t]: smallest key of a candidate holding t in its cavity or ring;
   * cavc[t] (shared rings): smallest key holding t in its cavity.
   * owner / cowner: t belongs to a winner's region / a winner's cavity. */
  long long *claim = malloc(sizeof(long long) * (size_t)m->nt);
  long long *cavc = shared_ring ? malloc(sizeof(long long) * (size_t)m->nt) : NULL;
  int *owner = malloc(sizeof(int) * (size_t)m->nt);
  int *cowner = malloc(sizeof(int) * (size_t)m->nt);
  for (int t = 0; t < m->nt; t++) owner[t] = cowner[t] = -1;
  int nwin = 0;
  int *win = calloc((size_t)nc, sizeof(int));
  int *active = malloc(sizeof(int) * (size_t)(nc > 0 ? nc : 1));
  for (int i = 0; i < nc; i++) active[i] = cd[i].status == ST_VALID;
  for (int it = 0; it < claim_iters; it++) {
    for (int t = 0; t < m->nt; t++) claim[t] = 0x7fffffffffffffffLL;
    if (cavc)
      for (int t = 0; t < m->nt; t++) cavc[t] = 0x7fffffffffffffffLL;
    for (int i = 0; i < nc; i++) {
      if (!active[i]) continue;
      long long key = ((long long)cd[i].prio << 32) | i;
      for (int q = 0; q < cd[i].n + cd[i].r; q++) {
        int t = cd[i].L[q];
        if (claim[t] > key) claim[t] = key;
        if (cavc && q < cd[i].n && cavc[t] > key) cavc[t] = key;
      }
    }
    int newwin = 0;
    for (int i = 0; i < nc; i++) {
      if (!active[i]) continue;
      long long key = ((long long)cd[i].prio << 32) | i;
      int ok = 1;
      for (int q = 0; q < cd[i].n + cd[i].r && ok; q++) {
        int t = cd[i].L[q];
        if (cavc && q >= cd[i].n) ok = cavc[t] >= key;  /* ring: no better cavity */
        else ok = claim[t] == key;
      }
      if (ok) { win[i] = 1; newwin++; }
    }
    for (int i = 0; i < nc; i++)
      if (win[i] && active[i])
        for (int q = 0; q < cd[i].n + cd[i].r; q++) {
          owner[cd[i].L[q]] = i;
          if (q < cd[i].n) cowner[cd[i].L[q]] = i;
        }
    for (int i = 0; i < nc; i++) {
      if (!active[i]) continue;
      if (win[i]) { active[i] = 0; continue; }
      for (int q = 0; q < cd[i].n + cd[i].r; q++) {
        int t = cd[i].L[q];
        int hit = (cavc && q >= cd[i].n) ? cowner[t] >= 0 : owner[t] >= 0;
        if (hit) { active[i] = 0; break; }
      }
    }
    nwin += newwin;
    if (newwin == 0) break;
  }
  for (int i = 0; i < nc; i++)
    if (cd[i].status == ST_VALID && !win[i]) st->lost++;
  free(owner);
  free(cowner);
  free(cavc);
  free(active);
  /* slot bases (winners take points/triangles/subsegments in candidate order) */
  int pbase = m->np, tbase = m->nt, sbase = m->ns;
  int nws = 0;
  for (int i = 0; i < nseg; i++) nws += win[i];
  grow_points(m, pbase + nwin);
  grow_tris(m, tbase + 2 * nwin);
  grow_segs(m, sbase + nws);
  int rank = 0;
  /* segment winners come first, so their rank is also their subsegment rank */
  for (int i = 0; i < nc; i++) {
    if (!win[i]) continue;
    const int is_seg = i < nseg;
    /* flips of the same insertion done by Lawson's algorithm: the first
     * triangle (two on a segment with both sides) split, then one flip per
     * further cavity triangle */
    int t1 = 0, j1, t2 = 0, j2;
    if (is_seg) seg_sides(m, cd[i].id, &t1, &j1, &t2, &j2);
    st->flips += cd[i].n - ((is_seg && t1 >= 0 && t2 >= 0) ? 2 : 1);
    int p = pbase + rank;
    m->P[2 * p] = cd[i].c[0];
    m->P[2 * p + 1] = cd[i].c[1];
    m->pflag[p] = is_seg ? PF_SEGV : PF_FREE;
    m->ppar[p] = is_seg ? m->SP[cd[i].id] : -1;
    use_inst(m, cd[i].L[0]);
    apply_star(m, &cd[i], is_seg ? cd[i].id : -1, p, tbase + 2 * rank,
               is_seg ? sbase + rank : -1, is_seg);
    rank++;
  }
  m->np = pbase + nwin;
  m->nt = tbase + 2 * nwin;
  m->ns = sbase + nws;
  st->seg_splits += nws;
  st->inserted += nwin - nws;
  *nwin_seg = nws;
  *progress = nwin + ndel + marks + (st->giveups - gave);
  if (getenv("ORC_TRACE")) {
    long long nval = 0, csum = 0, cmax = 0;
    for (int i = 0; i < nc; i++)
      if (cd[i].status == ST_VALID) {
        nval++;
        csum += cd[i].n;
        if (cd[i].n > cmax) cmax = cd[i].n;
      }
    fprintf(stderr, "pass nseg=%d nc=%d valid=%lld win=%d (seg %d) marks=%lld avgcav=%.2f maxcav=%lld K=%d\n",
            nseg, nc, nval, nwin, nws, marks, nval ? (double)csum / nval : 0.0, cmax, K);
  }
  for (int i = 0; i < nc; i++) free(cd[i].L);
  free(cd);
  free(pool);
  free(claim);
  free(win);
  return nwin;
}

static void mark_bad(Mesh *m) {
  for (int t = 0; t < m->nt; t++) {
    uint8_t f = m->tflag[t];
    if (!(f & TF_ALIVE) || !(f & TF_CHECK)) continue;
    f &= (uint8_t)~TF_CHECK;
    use_inst(m, t);
    if (tri_is_bad(m, t)) f |= TF_BAD;
    else f &= (uint8_t)~TF_BAD;
    m->tflag[t] = f;
  }
}

/* Working mesh from an input CDT: otri neighbours, subsegments attached to
 * their edges, initial encroachment (refine.cu:5557-5747, 3105-3178). */
static void init_mesh(Mesh *m, int np, const double *points, int nt,
                      const int *tris, const int *nbrs, int ns, const int *segs,
                      const int *tri_inst, const double *ipar,
                      const orc_params *prm) {
  memset(m, 0, sizeof(*m));
  m->ipar = ipar;
  double ga = cos(prm->theta_deg * 3.141592653589793238462643383279502884197169399375105820974944592308 / 180.0);
  m->offconst = (ga == 1.0) ? 0.0 : 0.475 * sqrt((1.0 + ga) / (1.0 - ga));
  m->goodangle = ga * ga;
  m->lens = (2.0 * m->goodangle - 1.0) * (2.0 * m->goodangle - 1.0);
  m->offcenter = prm->offcenter;
  m->ruppert = prm->mode == 1;
  grow_points(m, np);
  memcpy(m->P, points, sizeof(double) * 2 * (size_t)np);
  for (int i = 0; i < np; i++) { m->pflag[i] = PF_INPUT; m->ppar[i] = -1; }
  m->np = np;
  grow_tris(m, nt);
  memcpy(m->TV, tris, sizeof(int) * 3 * (size_t)nt);
  if (ipar) memcpy(m->tinst, tri_inst, sizeof(int) * (size_t)nt);
  for (int t = 0; t < nt; t++) {
    for (int j = 0; j < 3; j++) {
      int u = nbrs[3 * t + j];
      int code = -1;
      if (u >= 0) { /* find the neighbour's edge that matches (reversed) */
        int a = tris[3 * t + (j + 1) % 3], b = tris[3 * t + (j + 2) % 3];
        for (int k = 0; k < 3; k++)
          if (tris[3 * u + (k + 1) % 3] == b && tris[3 * u + (k + 2) % 3] == a) code = (u << 2) | k;
      }
      m->TN[3 * t + j] = code;
      m->TS[3 * t + j] = -1;
    }
    m->tflag[t] = TF_ALIVE | TF_CHECK;
  }
  m->nt = nt;
  /* input segments -> subsegments on triangle edges (the reference matches
   * segments to edges by brute force, refine.cu:5618-5747; a vertex-keyed
   * scan gives the same incidence) */
  m->nseg_in = ns;
  m->SEG = malloc(sizeof(int) * 2 * (size_t)(ns > 0 ? ns : 1));
  memcpy(m->SEG, segs, sizeof(int) * 2 * (size_t)ns);
  grow_segs(m, ns);
  for (int s = 0; s < ns; s++) {
    m->SV[2 * s] = segs[2 * s];
    m->SV[2 * s + 1] = segs[2 * s + 1];
    m->SP[s] = s;
    m->ST[s] = -1;
    m->senc[s] = 0;
  }
  m->ns = ns;
  {
    /* hash edges: for each segment find an adjacent triangle edge */
    size_t hc = 1;
    while (hc < (size_t)ns * 4 + 16) hc <<= 1;
    long long *hk = malloc(sizeof(long long) * hc);
    int *hv = malloc(sizeof(int) * hc);
    for (size_t i = 0; i < hc; i++) hk[i] = -1;
    for (int s = 0; s < ns; s++) {
      long long a = segs[2 * s], b = segs[2 * s + 1];
      long long key = a < b ? (a << 32) | b : (b << 32) | a;
      size_t h = (size_t)((unsigned long long)key * 0x9E3779B97F4A7C15ULL) & (hc - 1);
      while (hk[h] != -1) h = (h + 1) & (hc - 1);
      hk[h] = key;
      hv[h] = s;
    }
    for (int t = 0; t < nt; t++) {
      for (int j = 0; j < 3; j++) {
        long long a = tris[3 * t + (j + 1) % 3], b = tris[3 * t + (j + 2) % 3];
        long long key = a < b ? (a << 32) | b : (b << 32) | a;
        size_t h = (size_t)((unsigned long long)key * 0x9E3779B97F4A7C15ULL) & (hc - 1);
        while (hk[h] != -1) {
          if (hk[h] == key) {
            int s = hv[h];
            m->TS[3 * t + j] = s;
            /* ST: prefer the side where the edge runs org->dest */
            if (m->ST[s] < 0 || tris[3 * t + (j + 1) % 3] == segs[2 * s]) m->ST[s] = (t << 2) | j;
            break;
          }
          h = (h + 1) & (hc - 1);
        }
      }
    }
    free(hk);
    free(hv);
  }
  /* initial encroachment: every subsegment against both apexes */
  for (int t = 0; t < nt; t++)
    for (int j = 0; j < 3; j++) {
      int s = m->TS[3 * t + j];
      use_inst(m, t);
      if (s >= 0 && encroaches(m, m->SV[2 * s], m->SV[2 * s + 1], PT(m, m->TV[3 * t + j])))
        m->senc[s] = 1;
    }
}

static void free_mesh(Mesh *m) {
  free(m->P); free(m->pflag); free(m->ppar);
  free(m->TV); free(m->TN); free(m->TS); free(m->tflag);
  free(m->SV); free(m->SP); free(m->ST); free(m->senc); free(m->SEG);
  free(m->tinst);
}

/* Keep the cap smallest (priority, index) of the bad triangles ids[0..nb),
 * in index order (priority 1); otherwise the first cap.  Returns the count. */
static int select_bad(Mesh *m, int *ids, int nb, int cap, int priority) {
  if (nb <= cap) return nb;
  if (priority == 1) {
    long long *key = malloc(sizeof(long long) * (size_t)nb);
    for (int i = 0; i < nb; i++)
      key[i] = ((long long)shortest_edge_bucket(m, ids[i]) << 32) | ids[i];
    qsort(key, (size_t)nb, sizeof(long long), cmp_ll);
    for (int i = 0; i < cap; i++) ids[i] = (int)(key[i] & 0xffffffffLL);
    qsort(ids, (size_t)cap, sizeof(int), cmp_int);
    free(key);
  }
  return cap;
}

/* One round of the mixed form: the encroached subsegments (capped by the
 * workspace) followed by the selected bad triangles, in one pass.  Returns 1
 * when the refinement ends (st->status set). */
static int mixed_round(Mesh *m, const orc_params *prm, orc_stats *st, int *K, int *prev_win,
                       int **ids, int *idcap) {
  mark_bad(m);
  int ne = 0, nb = 0;
  for (int s = 0; s < m->ns; s++) ne += m->senc[s] == 1;
  for (int t = 0; t < m->nt; t++)
    nb += (m->tflag[t] & (TF_ALIVE | TF_BAD | TF_GIVEUP)) == (TF_ALIVE | TF_BAD);
  if (ne == 0 && nb == 0) { st->status = 0; return 1; }
  if (ne + nb > *idcap) {
    *idcap = ne + nb;
    *ids = xrealloc(*ids, sizeof(int) * (size_t)*idcap);
  }
  int *L = *ids;
  int cap = cands_per_pass(prm->cand_budget, *K);
  int k = 0;
  for (int s = 0; s < m->ns && k < cap; s++)
    if (m->senc[s] == 1) L[k++] = s;
  ne = k;
  int capb = cap - ne;
  if (prm->cand_factor > 0 && *prev_win >= 0) {
    long long lim = (long long)prm->cand_factor * *prev_win;
    if (lim < prm->cand_min) lim = prm->cand_min;
    if (lim < capb) capb = (int)lim;
  }
  int *B = L + ne;
  nb = 0;
  for (int t = 0; t < m->nt; t++)
    if ((m->tflag[t] & (TF_ALIVE | TF_BAD | TF_GIVEUP)) == (TF_ALIVE | TF_BAD)) B[nb++] = t;
  nb = select_bad(m, B, nb, capb, prm->priority);
  int of = 0, nws = 0;
  long long prog = 0;
  /* a pass that splits only subsegments takes seg_claim_iters iterations */
  const int ci = (nb == 0 && prm->seg_claim_iters > 0) ? prm->seg_claim_iters : prm->claim_iters;
  int nw = run_pass(m, L, ne + nb, ne, *K, prm->cav_cap_max, prm->priority, ci,
                    prm->shared_ring, prm->deletion, st, &of, &prog, &nws);
  if (ne > 0) st->seg_passes++;
  if (nb > 0) {
    st->bad_passes++;
    st->cand_evals += nb;
    *prev_win = nw - nws;
  }
  if (of && *K < prm->cav_cap_max) {
    *K = *K * 2 > prm->cav_cap_max ? prm->cav_cap_max : *K * 2;
    prog++;
  }
  if (prog == 0) { st->status = 2; return 1; }
  return 0;
}

static orc_result *refine_impl(int np, const double *points, int nt,
                               const int *tris, const int *nbrs, int ns,
                               const int *segs, const int *tri_inst,
                               const double *ipar, const orc_params *prm,
                               orc_stats *st) {
  Mesh M;
  memset(st, 0, sizeof(*st));
  Mesh *m = &M;
  init_mesh(m, np, points, nt, tris, nbrs, ns, segs, tri_inst, ipar, prm);
  int K = prm->cav_cap0;
  int prev_win = -1;
  int *ids = NULL;
  int idcap = 0;
  st->status = 1;
  int np_round = m->np;
  for (int round = 0; round < prm->max_rounds; round++) {
    st->rounds = round + 1;
    np_round = m->np;
    if (prm->mixed) {
      if (mixed_round(m, prm, st, &K, &prev_win, &ids, &idcap)) break;
      continue;
    }
    /* segment phase */
    for (;;) {
      int ne = 0;
      for (int s = 0; s < m->ns; s++) if (m->senc[s] == 1) ne++;
      if (ne <= prm->mix_below) break;
      if (ne > idcap) { idcap = ne; ids = xrealloc(ids, sizeof(int) * (size_t)idcap); }
      ne = 0;
      for (int s = 0; s < m->ns; s++) if (m->senc[s] == 1) ids[ne++] = s;
      int cap = cands_per_pass(prm->cand_budget, K);
      if (ne > cap) ne = cap;
      int of = 0, nws = 0;
      long long prog = 0;
      run_pass(m, ids, ne, ne, K, prm->cav_cap_max, prm->priority,
               prm->seg_claim_iters > 0 ? prm->seg_claim_iters : prm->claim_iters, prm->shared_ring,
               prm->deletion, st, &of, &prog, &nws);
      st->seg_passes++;
      if (of && K < prm->cav_cap_max) K = K * 2 > prm->cav_cap_max ? prm->cav_cap_max : K * 2;
    }
    /* bad-triangle phase (with the at most mix_below subsegments still
     * encroached, first in candidate order) */
    if (mixed_round(m, prm, st, &K, &prev_win, &ids, &idcap)) break;
  }
  st->final_cav_cap = K;
  free(ids);
  /* given-up elements still in the mesh: status 3 instead of converged */
  for (int t = 0; t < m->nt; t++)
    st->giveup_tris_left += (m->tflag[t] & (TF_ALIVE | TF_GIVEUP)) == (TF_ALIVE | TF_GIVEUP);
  for (int s = 0; s < m->ns; s++) st->giveup_segs_left += m->senc[s] == 2;
  if (st->status == 0 && (st->giveup_tris_left > 0 || st->giveup_segs_left > 0)) st->status = 3;

  for (int p = np_round; p < m->np; p++) m->pflag[p] |= PF_NEW;
  /* export: live points in index order (deleted vertices dropped, the
   * others renumbered), live triangles in slot order */
  orc_result *r = calloc(1, sizeof(orc_result));
  int *pnew = malloc(sizeof(int) * (size_t)(m->np > 0 ? m->np : 1));
  int npl = 0;
  for (int p = 0; p < m->np; p++) pnew[p] = (m->pflag[p] & PF_DEAD) ? -1 : npl++;
  int *newid = malloc(sizeof(int) * (size_t)(m->nt > 0 ? m->nt : 1));
  int nalive = 0;
  for (int t = 0; t < m->nt; t++) newid[t] = (m->tflag[t] & TF_ALIVE) ? nalive++ : -1;
  r->np = npl;
  r->nt = nalive;
  r->ns = m->ns;
  r->P = malloc(sizeof(double) * 2 * (size_t)(npl > 0 ? npl : 1));
  r->pflag = malloc((size_t)(npl > 0 ? npl : 1));
  for (int p = 0; p < m->np; p++) {
    if (pnew[p] < 0) continue;
    r->P[2 * (size_t)pnew[p]] = m->P[2 * (size_t)p];
    r->P[2 * (size_t)pnew[p] + 1] = m->P[2 * (size_t)p + 1];
    r->pflag[pnew[p]] = m->pflag[p];
  }
  r->T = malloc(sizeof(int) * 3 * (size_t)(nalive > 0 ? nalive : 1));
  r->I = m->ipar ? malloc(sizeof(int) * (size_t)(nalive > 0 ? nalive : 1)) : NULL;
  r->N = malloc(sizeof(int) * 3 * (size_t)(nalive > 0 ? nalive : 1));
  for (int t = 0; t < m->nt; t++) {
    int q = newid[t];
    if (q < 0) continue;
    for (int j = 0; j < 3; j++) {
      r->T[3 * q + j] = pnew[m->TV[3 * t + j]];
      int nb = m->TN[3 * t + j];
      r->N[3 * q + j] = nb < 0 ? -1 : newid[nb >> 2];
    }
    if (r->I) r->I[q] = m->tinst[t];
  }
  r->S = malloc(sizeof(int) * 2 * (size_t)(m->ns > 0 ? m->ns : 1));
  for (int q = 0; q < 2 * m->ns; q++) r->S[q] = pnew[m->SV[q]];
  free(newid);
  free(pnew);
  free_mesh(m);
  return r;
}

orc_result *orc_refine(int np, const double *points, int nt, const int *tris,
                       const int *nbrs, int ns, const int *segs,
                       const orc_params *prm, orc_stats *st) {
  return refine_impl(np, points, nt, tris, nbrs, ns, segs, NULL, NULL, prm, st);
}

/* Batched refinement of the union of independent meshes: tri_inst[nt] gives
 * every input triangle's instance, theta/offcenter/mode[n_inst] the
 * instances' parameters (prm's own theta, offcenter and mode are unused). */
orc_result *orc_refine_batch(int np, const double *points, int nt,
                             const int *tris, const int *nbrs, int ns,
                             const int *segs, const int *tri_inst, int n_inst,
                             const double *theta_deg, const int *offcenter,
                             const int *mode, const orc_params *prm,
                             orc_stats *st) {
  double *ip = malloc(sizeof(double) * 5 * (size_t)(n_inst > 0 ? n_inst : 1));
  for (int i = 0; i < n_inst; i++) {
    double ga = cos(theta_deg[i] * 3.141592653589793238462643383279502884197169399375105820974944592308 / 180.0);
    double g2 = ga * ga;
    ip[5 * i + 0] = g2;
    ip[5 * i + 1] = (2.0 * g2 - 1.0) * (2.0 * g2 - 1.0);
    ip[5 * i + 2] = (ga == 1.0) ? 0.0 : 0.475 * sqrt((1.0 + ga) / (1.0 - ga));
    ip[5 * i + 3] = offcenter[i];
    ip[5 * i + 4] = mode[i] == 1;
  }
  orc_result *r = refine_impl(np, points, nt, tris, nbrs, ns, segs, tri_inst, ip, prm, st);
  free(ip);
  return r;
}

/* The Bowyer-Watson cavity of point c in an input CDT, grown from the
 * triangle the walk from triangle t0 reaches (for tests against brute force):
 * writes the cavity's triangles (ascending) to out and returns their number,
 * or -1 if the walk or the cavity check fails, or the cavity exceeds cap. */
int orc_cavity(int np, const double *points, int nt, const int *tris,
               const int *nbrs, int ns, const int *segs, int t0,
               const double *c, int cap, int *out) {
  orc_params prm;
  memset(&prm, 0, sizeof(prm));
  prm.theta_deg = 20.0;
  Mesh M;
  init_mesh(&M, np, points, nt, tris, nbrs, ns, segs, NULL, NULL, &prm);
  Cand cd;
  memset(&cd, 0, sizeof(cd));
  cd.c[0] = c[0];
  cd.c[1] = c[1];
  cd.L = malloc(sizeof(int) * (size_t)(2 * cap + 2));
  int tl = -1, jl = -1, sg = -1, n = -1;
  int w = walk(&M, t0, cd.c, &tl, &jl, &sg);
  if (w == 0 || w == 1) {
    int seed1 = -1;
    if (w == 1 && M.TS[3 * tl + jl] < 0 && M.TN[3 * tl + jl] >= 0) seed1 = M.TN[3 * tl + jl] >> 2;
    if (w == 0 || seed1 >= 0) {
      int stt = build_cavity(&M, &cd, tl, seed1, -1, cap);
      if (stt == ST_VALID) {
        n = cd.n;
        memcpy(out, cd.L, sizeof(int) * (size_t)n);
      }
    }
  }
  free(cd.L);
  free_mesh(&M);
  return n;
}

void orc_result_inst(const orc_result *r, int *tri_inst) {
  if (r->I) memcpy(tri_inst, r->I, sizeof(int) * (size_t)r->nt);
}

void orc_result_counts(const orc_result *r, int *np, int *nt, int *ns) {
  *np = r->np;
  *nt = r->nt;
  *ns = r->ns;
}

void orc_result_copy(const orc_result *r, double *points, int *tris, int *nbrs,
                     int *segs, uint8_t *pflag) {
  if (points) memcpy(points, r->P, sizeof(double) * 2 * (size_t)r->np);
  if (tris) memcpy(tris, r->T, sizeof(int) * 3 * (size_t)r->nt);
  if (nbrs) memcpy(nbrs, r->N, sizeof(int) * 3 * (size_t)r->nt);
  if (segs) memcpy(segs, r->S, sizeof(int) * 2 * (size_t)r->ns);
  if (pflag) memcpy(pflag, r->pflag, (size_t)r->np);
}

void orc_result_free(orc_result *r) {
  if (!r) return;
  free(r->P); free(r->T); free(r->N); free(r->S); free(r->pflag); free(r->I);
  free(r);
}

/* Closed-form constructions exposed for tests. */
int orc_steiner_point(const double *apex, const double *org, const double *dest,
                      double theta_deg, int offcenter, double *out) {
  Mesh m;
  memset(&m, 0, sizeof(m));
  double P[6] = {apex[0], apex[1], org[0], org[1], dest[0], dest[1]};
  int TV[3] = {0, 1, 2};
  double ga = cos(theta_deg * 3.141592653589793238462643383279502884197169399375105820974944592308 / 180.0);
  m.offconst = (ga == 1.0) ? 0.0 : 0.475 * sqrt((1.0 + ga) / (1.0 - ga));
  m.offcenter = offcenter;
  m.P = P;
  m.TV = TV;
  return steiner_point(&m, 0, out);
}
