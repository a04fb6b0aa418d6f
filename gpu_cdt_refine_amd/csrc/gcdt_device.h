// Device-side mesh model and per-element refinement steps for gfx950.
//
// Mesh layout in HBM.  Triangles are 128-byte records, one L2 line each:
// every kernel that visits a triangle needs its vertices, neighbours,
// subsegments and vertex coordinates together, and the visits are random
// gathers, so one line per triangle replaces six (three 16-byte arrays and
// three point lookups).  Per-triangle flags and claims, which are scanned or
// updated alone, stay in their own arrays.
//   T     TriRec[capT]    V (v0, v1, v2, instance) counter-clockwise,
//                         N neighbour across edge j as otri = tri<<2 | edge
//                         or -1, S subsegment on edge j or -1, X[3] the
//                         vertices' coordinates (copies of P)
//   P     double2[capP]   point coordinates
//   pflag uint8[capP]     PF_* bits (PF_DEAD: deleted, dropped at export);
//                         ppar int[capP]: parent input segment
//   tflag uint8[capT]     TF_* bits
//   SV    int2[capS]      subsegment endpoints (org, dest)
//   SP    int[capS]       parent input segment;  ST int[capS]: an adjacent otri
//   senc  int[capS]       0 free, 1 encroached, 2 given up
//   claim uint64[capT]    epoch-tagged conflict claims (atomicMax)
// Edge j of triangle t is (v[(j+1)%3], v[(j+2)%3]); its apex is v[j]
// (otri encoding as the reference's encode_tri, refine.cu:95-113).
//
// Every step here has a line-for-line counterpart in the CPU oracle
// (oracle/refine_oracle.c) so that both produce bit-identical meshes; the
// comments name the reference code each step stands in for.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "predicates.h"

namespace gcdt {

enum : uint8_t { TF_ALIVE = 1, TF_CHECK = 2, TF_BAD = 4, TF_GIVEUP = 8 };
// PF_NEW: inserted in the run's last round (set at export)
enum : uint8_t { PF_INPUT = 1, PF_SEGV = 2, PF_FREE = 4, PF_DEAD = 8, PF_NEW = 16 };
enum : int {
  ST_VALID = 0,
  ST_REJ_ENC = 1,
  ST_REJ_BLOCKED = 2,
  ST_OVERFLOW = 3,
  ST_GIVEUP = 4,
  ST_BIG = 5  // cavity larger than the LDS list; recomputed in global memory
};
enum : int {
  CNT_MARKS = 0,
  CNT_GIVEUPS,
  CNT_OVERFLOW,
  CNT_REJ_ENC,
  CNT_REJ_BLOCKED,
  CNT_LOST,
  CNT_FP32,
  CNT_FP64,
  CNT_EXACT,
  CNT_ERRORS,
  CNT_VALID,
  CNT_BIG,
  CNT_DD,
  CNT_SLOW,    // candidates deferred to the exact-predicate kernel
  CNT_SEGWIN,  // winners that split a subsegment
  CNT_EXAM_SEG,  // triangle records examined by the segment cavity kernels
  CNT_EXAM_BAD,  // ... by the bad-triangle cavity kernels
  CNT_WS_SEG,    // ints written to candidate workspaces (segment kernels)
  CNT_WS_BAD,    // ... (bad-triangle kernels)
  CNT_NWIN,      // winners of the pass
  CNT_LEFT_TRI,  // end of the run: alive triangles flagged TF_GIVEUP
  CNT_LEFT_SEG,  // end of the run: subsegments with senc == 2
  CNT_DEL,       // vertices deleted
  CNT_FLIPS,     // flips Lawson's algorithm would make for the pass's changes
  CNT_NENC,      // encroached subsegments after the pass (listed behind its apply)
  CNT_N = 25
};

#define GCDT_WALK_LIMIT (1 << 22)

struct MeshD {
  double2 *P;
  uint8_t *pflag;
  int *ppar;
  struct TriRec *T;
  uint8_t *tflag;
  uint16_t *tbucket;  // shortest-edge bucket, written when a triangle is marked bad
  int2 *SV;
  int *SP, *ST, *senc;
  const int2 *SEG;
  unsigned long long *claim;  // per triangle t: [2t] claims (shared rings:
                              // cavity claims), [2t+1] ring claims (shared rings)
  int batched;  // 0: one instance, parameters from ParamsD; 1: T[t].V.w
};

struct __align__(128) TriRec {
  int4 V;          // vertices; w = instance (batched refinement), else 0
  int4 N;          // neighbour otri across edge j (w unused)
  int4 S;          // subsegment on edge j (w unused)
  double2 X[3];    // coordinates of V.x, V.y, V.z
  int4 pad[2];
};
static_assert(sizeof(TriRec) == 128, "one L2 line per triangle");

// Quality parameters of one PSLG instance (the reference's per-call theta,
// insertpolicy and mode; refine.cu:6557-6567 derives them from theta).
struct InstPar {
  double goodangle;  // cos^2(theta)
  double lens;       // (2 cos^2(theta) - 1)^2
  double offconst;   // 0.475 sqrt((1+cos)/(1-cos)), kept in double
  int offcenter, ruppert;
};

struct ParamsD {
  double goodangle;  // cos^2(theta)
  double lens;       // (2 cos^2(theta) - 1)^2
  double offconst;   // 0.475 sqrt((1+cos)/(1-cos)), kept in double
  int offcenter, ruppert, use_fp32, priority;
  const InstPar *ipar;  // per-instance parameters when MeshD::batched is set
};

// Parameters that apply to triangle t (its instance's, in a batch).
__device__ __forceinline__ ParamsD params_at(const ParamsD &pr, const MeshD &m,
                                             int t) {
  if (!m.batched) return pr;
  InstPar q = pr.ipar[m.T[t].V.w];
  ParamsD r = pr;
  r.goodangle = q.goodangle;
  r.lens = q.lens;
  r.offconst = q.offconst;
  r.offcenter = q.offcenter;
  r.ruppert = q.ruppert;
  return r;
}

// Device-resident control block of the graph-driven round loop: live
// counts, loop state and statistics, read and written only by kernels
// between two graph launches (the host reads it once per launch).
enum : int {
  C_GO = 0,      // the current pass runs (1) or its kernels return at once (0)
  C_DONE,        // refinement finished
  C_SUSPEND,     // 1: the next pass needs more mesh capacity (host grows)
  C_STATUS,      // gcdt_stats.status
  C_NP, C_NT, C_NS,  // live slot counts
  C_K,           // cavity capacity wanted by the loop
  C_KCAP,        // cavity capacity the captured graph was built for
  C_NE, C_NB, C_NC, C_NWIN, C_PREVWIN,
  C_ROUNDS, C_ROUND_OPEN, C_SEG_DONE,
  C_SEL,         // bad pass: priority selection needed
  C_USESEL,      // candidates are read from the selected list
  C_PASS_INDEX, C_SEG_PASSES, C_BAD_PASSES,
  C_EPOCH,       // last claim epoch used
  C_BCUT, C_TAKE, C_ICUT,
  C_CAPP, C_CAPT, C_CAPS,
  C_MAXROUNDS, C_KMAX, C_CANDMIN, C_CANDFACTOR, C_ITERS, C_PRIORITY,
  C_NTOUT,
  C_NP_ROUND,    // live points when the current round opened
  C_NI = 64
};
enum : int {
  A_REJ_OVF = 0, A_REJ_ENC, A_REJ_BLK, A_GIVEUPS, A_LOST, A_FP32, A_FP64,
  A_EXACT, A_DD, A_BIG, A_SEGSPLITS, A_INSERTED, A_CANDEVALS, A_BUDGET,
  A_EXAM_SEG, A_EXAM_BAD, A_WS_SEG, A_WS_BAD, A_DELETED, A_FLIPS,
  A_NA = 20
};
struct Ctl {
  int i[C_NI];
  unsigned long long cnt[CNT_N];  // per-pass counters (CNT_*)
  unsigned long long acc[A_NA];   // run totals
};

// One conflict-resolution pass over nc candidates.
struct PassD {
  const int *ids;
  int nc;
  int nseg;  // candidates [0, nseg) split subsegments, the rest are bad
             // triangles; -1: every candidate of the pass is a subsegment
  int K;   // cavity capacity (triangles)
  int Kf;  // cavity capacity of the filtered kernel's LDS fast path
  int R;  // workspace ints per candidate = 6K+16
  unsigned long long epoch;
  double2 *cpt;
  int *cstatus, *cn, *cr, *cprio, *cwin, *crank;
  int *ws;
  double *pws;  // exact-predicate arena, GCDT_INCIRCLE_WS doubles per thread
  unsigned long long *cnt;
  unsigned long long epoch0;  // first claim epoch of this pass
  int iter;                   // claim iteration
  int iters;                  // claim iterations of the pass
  int shared_ring;            // winners' rings may overlap (cavities may not
                              // meet another winner's cavity or ring)
  float *feat;                // optional: FEAT features per candidate
  int pass_index;             // bad-pass counter (feature)
  // graph-driven loop: counts, epochs and the candidate list come from ctl
  Ctl *ctl;
  const int *ids_sel;  // selected candidates when ctl->i[C_USESEL]
  int *slow;           // candidates whose filters could not decide a sign
  int eskip;  // graph loop: the pass's vertex deletion step took an epoch
  // learned priority folded in by the first claim iteration (host loop):
  // the network's scores, and the bucket coarsening shift; nullptr: cprio
  // is final
  const float *score;
  int lshift;
};

// Learned claim priority of a bad-triangle candidate: its shortest-edge
// bucket coarsened by 2^shift half-octaves, then higher predicted conflict
// first (the score quantised to 8 bits).
__device__ __forceinline__ int learned_prio(int bucket, float score, int shift) {
  const float q = score * 32.f;
  const int qi = q != q ? 0 : (int)fminf(fmaxf(q, 0.f), 255.f);
  return ((bucket >> shift) << 8) | (255 - qi);
}

__device__ __forceinline__ int pass_nc(const PassD &ps) {
  if (!ps.ctl) return ps.nc;
  return ps.ctl->i[C_GO] ? ps.ctl->i[C_NC] : 0;
}
__device__ __forceinline__ int pass_nseg(const PassD &ps) {
  return ps.nseg < 0 ? pass_nc(ps) : ps.nseg;
}
__device__ __forceinline__ const int *pass_ids(const PassD &ps) {
  return (ps.ctl && ps.ctl->i[C_USESEL]) ? ps.ids_sel : ps.ids;
}
__device__ __forceinline__ unsigned long long pass_epoch0(const PassD &ps) {
  return ps.ctl ? (unsigned long long)ps.ctl->i[C_EPOCH] + 1ull + (unsigned long long)ps.eskip
                : ps.epoch0;
}
__device__ __forceinline__ unsigned long long pass_epoch(const PassD &ps) {
  return ps.ctl ? pass_epoch0(ps) + (unsigned long long)ps.iter : ps.epoch;
}
// Winners' regions are stamped with the epoch after the pass's last claim
// epoch: no claim of the pass can displace them.
__device__ __forceinline__ unsigned long long pass_sticky(const PassD &ps) {
  return (pass_epoch0(ps) + (unsigned long long)ps.iters) & 2047ull;
}
__device__ __forceinline__ int pass_index_of(const PassD &ps) {
  return ps.ctl ? ps.ctl->i[C_PASS_INDEX] : ps.pass_index;
}


__device__ __forceinline__ int comp(const int4 &v, int j) {
  return j == 0 ? v.x : (j == 1 ? v.y : v.z);
}
// Component j of a triangle's neighbour / subsegment vector, for writes.
__device__ __forceinline__ int *nbr_slot(TriRec *T, int t, int j) {
  return &T[t].N.x + j;
}
__device__ __forceinline__ int *seg_slot(TriRec *T, int t, int j) {
  return &T[t].S.x + j;
}

// checkseg4encroach (refine.cu:115-139, triangle.cpp:7138-7153): does
// (px, py) encroach the subsegment with endpoints A, B?
__device__ __forceinline__ bool encroaches_xy(const ParamsD &pr, double2 A, double2 B,
                                              double px, double py) {
  double dot = (A.x - px) * (B.x - px) + (A.y - py) * (B.y - py);
  if (dot < 0.0) {
    if (pr.ruppert) return true;
    double la = (A.x - px) * (A.x - px) + (A.y - py) * (A.y - py);
    double lb = (B.x - px) * (B.x - px) + (B.y - py) * (B.y - py);
    if (dot * dot >= pr.lens * la * lb) return true;
  }
  return false;
}
__device__ __forceinline__ bool encroaches(const MeshD &m, const ParamsD &pr,
                                           int a, int b, double px,
                                           double py) {
  return encroaches_xy(pr, m.P[a], m.P[b], px, py);
}

// testtriangle + Miller-Pav-Walkington (refine.cu:376-497,
// triangle.cpp:7220-7384); org = v1, dest = v2, apex = v0.
__device__ __forceinline__ bool tri_is_bad(const MeshD &m, const ParamsD &pr,
                                           const TriRec &r) {
  const int4 v = r.V;
  double2 O = r.X[1], D = r.X[2], A = r.X[0];
  double dxod = O.x - D.x, dyod = O.y - D.y;
  double dxda = D.x - A.x, dyda = D.y - A.y;
  double dxao = A.x - O.x, dyao = A.y - O.y;
  double apexlen = dxod * dxod + dyod * dyod;
  double orglen = dxda * dxda + dyda * dyda;
  double destlen = dxao * dxao + dyao * dyao;
  double angle;
  int e, b1, b2;
  double2 B1, B2;
  if ((apexlen < orglen) && (apexlen < destlen)) {
    angle = dxda * dxao + dyda * dyao;
    angle = angle * angle / (orglen * destlen);
    e = 0; b1 = v.y; b2 = v.z; B1 = O; B2 = D;
  } else if (orglen < destlen) {
    angle = dxod * dxao + dyod * dyao;
    angle = angle * angle / (apexlen * destlen);
    e = 1; b1 = v.z; b2 = v.x; B1 = D; B2 = A;
  } else {
    angle = dxod * dxda + dyod * dyda;
    angle = angle * angle / (apexlen * orglen);
    e = 2; b1 = v.x; b2 = v.y; B1 = A; B2 = O;
  }
  if (!(angle > pr.goodangle)) return false;
  if ((m.pflag[b1] & PF_SEGV) && (m.pflag[b2] & PF_SEGV) &&
      comp(r.S, e) < 0) {
    int s1 = m.ppar[b1], s2 = m.ppar[b2];
    int2 g1 = m.SEG[s1], g2 = m.SEG[s2];
    int common = -1;
    if (g1.x == g2.x || g1.x == g2.y) common = g1.x;
    else if (g1.y == g2.x || g1.y == g2.y) common = g1.y;
    if (common >= 0) {
      double2 J = m.P[common];
      double dist1 = (B1.x - J.x) * (B1.x - J.x) + (B1.y - J.y) * (B1.y - J.y);
      double dist2 = (B2.x - J.x) * (B2.x - J.x) + (B2.y - J.y) * (B2.y - J.y);
      if ((dist1 < 1.001 * dist2) && (dist1 > 0.999 * dist2)) return false;
    }
  }
  return true;
}

// Circumcenter / Ungor off-center (refine.cu:3780-3868,
// triangle.cpp:6523-6618).  The denominator is Triangle's adaptive
// counterclockwise(dest, apex, org) value (triangle.cpp:6556-6563): the plain
// FP64 determinant when its filter certifies it, else the adaptive stages
// (orient2d_adapt_value), whose value stays accurate for near-collinear
// triangles.  Returns 1, 0 if the triangle is degenerate (collinear or the
// point coincides with a vertex), or STEINER_DEFER (FAST only) when the
// adaptive stages are needed: the candidate then goes to the exact kernel.
constexpr int STEINER_DEFER = 2;
template <bool FAST>
__host__ __device__ __forceinline__ int steiner_point(const ParamsD &pr, const TriRec &r, double &cx,
                                             double &cy) {
  double2 A = r.X[0], O = r.X[1], D = r.X[2];
  double xdo = D.x - O.x, ydo = D.y - O.y;
  double xao = A.x - O.x, yao = A.y - O.y;
  double dodist = xdo * xdo + ydo * ydo;
  double aodist = xao * xao + yao * yao;
  double dadist = (D.x - A.x) * (D.x - A.x) + (D.y - A.y) * (D.y - A.y);
  double det, detsum = 0.0;
  if (!orient2d_value_filter(D.x, D.y, A.x, A.y, O.x, O.y, &det, &detsum)) {
    if (FAST) return STEINER_DEFER;
    det = orient2d_adapt_value(D.x, D.y, A.x, A.y, O.x, O.y, detsum);
  }
  if (det == 0.0) return 0;
  double denominator = 0.5 / det;
  double dx = (yao * dodist - ydo * aodist) * denominator;
  double dy = (xdo * aodist - xao * dodist) * denominator;
  if (pr.offcenter && pr.offconst > 0.0) {
    double dxoff, dyoff;
    if ((dodist < aodist) && (dodist < dadist)) {
      dxoff = 0.5 * xdo - pr.offconst * ydo;
      dyoff = 0.5 * ydo + pr.offconst * xdo;
      if (dxoff * dxoff + dyoff * dyoff < dx * dx + dy * dy) { dx = dxoff; dy = dyoff; }
    } else if (aodist < dadist) {
      dxoff = 0.5 * xao + pr.offconst * yao;
      dyoff = 0.5 * yao - pr.offconst * xao;
      if (dxoff * dxoff + dyoff * dyoff < dx * dx + dy * dy) { dx = dxoff; dy = dyoff; }
    } else {
      dxoff = 0.5 * (A.x - D.x) - pr.offconst * (A.y - D.y);
      dyoff = 0.5 * (A.y - D.y) + pr.offconst * (A.x - D.x);
      if (dxoff * dxoff + dyoff * dyoff <
          (dx - xdo) * (dx - xdo) + (dy - ydo) * (dy - ydo)) {
        dx = xdo + dxoff;
        dy = ydo + dyoff;
      }
    }
  }
  cx = O.x + dx;
  cy = O.y + dy;
  if ((cx == A.x && cy == A.y) || (cx == O.x && cy == O.y) ||
      (cx == D.x && cy == D.y))
    return 0;
  if (!(cx == cx) || !(cy == cy)) return 0;
  return 1;
}

// Half-octave bucket of a squared edge length (Triangle's queue order,
// triangle.cpp:6932-6968): shorter edges -> smaller bucket -> higher priority.
__device__ __forceinline__ int edge_bucket(double key) {
  int e;
  double f = frexp(key, &e);
  int b = 2 * e + (f >= 0.70710678118654752440 ? 1 : 0);
  b += 2200;
  b = b < 0 ? 0 : (b > 8191 ? 8191 : b);
  return b;
}

__device__ __forceinline__ double shortest_edge_sq(const TriRec &r) {
  double2 A = r.X[0], O = r.X[1], D = r.X[2];
  double l0 = (O.x - D.x) * (O.x - D.x) + (O.y - D.y) * (O.y - D.y);
  double l1 = (D.x - A.x) * (D.x - A.x) + (D.y - A.y) * (D.y - A.y);
  double l2 = (A.x - O.x) * (A.x - O.x) + (A.y - O.y) * (A.y - O.y);
  double mn = l0 < l1 ? l0 : l1;
  if (l2 < mn) mn = l2;
  return mn;
}

__device__ __forceinline__ int shortest_edge_bucket(const TriRec &r) {
  return edge_bucket(shortest_edge_sq(r));
}

// Sides of subsegment s: (t1,j1) holds org->dest, (t2,j2) dest->org.
__device__ __forceinline__ void seg_sides(const MeshD &m, int s, int &t1,
                                          int &j1, int &t2, int &j2) {
  int a = m.SV[s].x;
  int ot = m.ST[s];
  int t = ot >> 2, j = ot & 3;
  int u = comp(m.T[t].V, (j + 1) % 3);
  int n = comp(m.T[t].N, j);
  int tb = -1, jb = -1;
  if (n >= 0) { tb = n >> 2; jb = n & 3; }
  if (u == a) { t1 = t; j1 = j; t2 = tb; j2 = jb; }
  else { t1 = tb; j1 = jb; t2 = t; j2 = j; }
}

// Segment split point with concentric shells + one collinearity step
// (refine.cu:2686-2725, triangle.cpp:13290-13425).
__device__ __forceinline__ bool split_point(const MeshD &m, int s, double &px,
                                            double &py) {
  int2 ab = m.SV[s];
  int t1, j1, t2, j2;
  seg_sides(m, s, t1, j1, t2, j2);
  bool acuteorg = false, acutedest = false;
  if (t1 >= 0) {
    int4 ts = m.T[t1].S;
    if (comp(ts, (j1 + 2) % 3) >= 0) acuteorg = true;
    if (comp(ts, (j1 + 1) % 3) >= 0) acutedest = true;
  }
  if (t2 >= 0) {
    int4 ts = m.T[t2].S;
    if (comp(ts, (j2 + 1) % 3) >= 0) acuteorg = true;
    if (comp(ts, (j2 + 2) % 3) >= 0) acutedest = true;
  }
  double2 A = m.P[ab.x], B = m.P[ab.y];
  double split;
  if (acuteorg || acutedest) {
    double len = sqrt((B.x - A.x) * (B.x - A.x) + (B.y - A.y) * (B.y - A.y));
    double npt = 1.0;
    while (len > 3.0 * npt) npt *= 2.0;
    while (len < 1.5 * npt) npt *= 0.5;
    split = npt / len;
    if (acutedest) split = 1.0 - split;
  } else {
    split = 0.5;
  }
  px = A.x + split * (B.x - A.x);
  py = A.y + split * (B.y - A.y);
  double mult = (A.x - px) * (B.y - py) - (A.y - py) * (B.x - px);
  double div = (A.x - B.x) * (A.x - B.x) + (A.y - B.y) * (A.y - B.y);
  if (mult != 0.0 && div != 0.0) {
    mult = mult / div;
    if (mult == mult) {
      px += mult * (B.y - A.y);
      py += mult * (A.x - B.x);
    }
  }
  if ((px == A.x && py == A.y) || (px == B.x && py == B.y)) return false;
  return true;
}

// Orientation walk toward (cx,cy), stopping at subsegments
// (refine.cu:3908-4184).  0 inside, 1 on edge jout, 2 on a vertex,
// 3 blocked by subsegment seg, 4 failed, 5 (FAST only) a sign the filters
// could not decide.
template <bool FAST>
__device__ __forceinline__ int walk(const MeshD &m, int t, double cx,
                                    double cy, bool fp32, PredStats *st,
                                    int &tout, int &jout, int &seg, int &nsteps) {
  nsteps = 0;
  for (int steps = 0; steps < GCDT_WALK_LIMIT; steps++) {
    nsteps = steps + 1;  // records read
    const TriRec &r = m.T[t];
    double2 p0 = r.X[0], p1 = r.X[1], p2 = r.X[2];
    int o0, o1, o2;
    if (FAST) {
      o0 = orient2d_filtered(p1.x, p1.y, p2.x, p2.y, cx, cy, fp32, st);
      o1 = orient2d_filtered(p2.x, p2.y, p0.x, p0.y, cx, cy, fp32, st);
      o2 = orient2d_filtered(p0.x, p0.y, p1.x, p1.y, cx, cy, fp32, st);
      if (o0 == PRED_UNSURE || o1 == PRED_UNSURE || o2 == PRED_UNSURE) return 5;
    } else {
      o0 = orient2d_sign(p1.x, p1.y, p2.x, p2.y, cx, cy, fp32, st);
      o1 = orient2d_sign(p2.x, p2.y, p0.x, p0.y, cx, cy, fp32, st);
      o2 = orient2d_sign(p0.x, p0.y, p1.x, p1.y, cx, cy, fp32, st);
    }
    int jn = o0 < 0 ? 0 : (o1 < 0 ? 1 : (o2 < 0 ? 2 : -1));
    if (jn < 0) {
      int zeros = (o0 == 0) + (o1 == 0) + (o2 == 0);
      tout = t;
      if (zeros == 0) return 0;
      if (zeros == 1) {
        jout = (o0 == 0) ? 0 : (o1 == 0) ? 1 : 2;
        return 1;
      }
      return 2;
    }
    int s = comp(r.S, jn);
    if (s >= 0) { seg = s; tout = t; return 3; }
    int nb = comp(r.N, jn);
    if (nb < 0) return 4;
    t = nb >> 2;
  }
  return 4;
}

}  // namespace gcdt
