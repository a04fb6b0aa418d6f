/*
 * CPU oracle for the MI355X CDT quality-refinement path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (gpu_cdt_refine_amd/,
 * include/) includes, links or calls this code; only tests/, the smoke check
 * and bench.py's cpu_baseline leg load oracle/liboracle.so, and only as the
 * checker.
 *
 * Contents:
 *   predicates_oracle.c  exact orient2d / incircle by big-integer arithmetic
 *   refine_oracle.c      sequential restatement of the bulk-synchronous
 *                        refinement rounds (reference refine.cu:6500-7139 and
 *                        Triangle's quality rules, triangle.cpp:6523-6618,
 *                        7106-7384, 13256-13632); the GPU path must reproduce
 *                        its output bit for bit
 *   check_oracle.c       mesh invariant checks (reference main.cpp:79-380,
 *                        with exact predicates and an O(T) segment lookup)
 */
#ifndef GCDT_ORACLE_H
#define GCDT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int orc_orient2d(const double *a, const double *b, const double *c);
int orc_incircle(const double *a, const double *b, const double *c,
                 const double *d);
int orc_orient2d_exact_sign(const double *a, const double *b, const double *c);
int orc_incircle_exact_sign(const double *a, const double *b, const double *c,
                            const double *d);

/* Parameters that determine the refined mesh.  The GPU path takes the same
 * values; with equal parameters both produce identical output. */
typedef struct {
  double theta_deg;   /* minimum angle */
  int offcenter;      /* 1 = off-center (Ungor), 0 = circumcenter */
  int mode;           /* 1 = Ruppert (diametral circle), else Chew (lens) */
  int cav_cap0;       /* initial cavity capacity K (triangles) */
  int cav_cap_max;    /* largest K */
  long long cand_budget; /* workspace ints; candidates per pass = budget / (6K+16) */
  int max_rounds;
  int priority;       /* 0 = lowest candidate index wins (refine.cu:4159);
                         1 = shortest-edge half-octave bucket first, then
                         index (Triangle's queue order, triangle.cpp:6932-6968,
                         and the reference's commented-out rule) */
  int cand_factor;    /* bad passes evaluate at most max(cand_min,
                         cand_factor * winners of the previous bad pass)
                         candidates, chosen by priority (0 = no limit) */
  int cand_min;
  int claim_iters;    /* conflict-resolution rounds per pass (>= 1): after the
                         first, valid candidates whose region touches no winner
                         claim again among themselves */
  int mixed;          /* 0: a round is segment passes until no subsegment is
                         encroached, then one bad-triangle pass (refine.cu:
                         6700-7100 round structure); 1: a round is ONE pass in
                         which the encroached subsegments (first in candidate
                         order, so they win every tie) and the selected bad
                         triangles compete together */
  int shared_ring;    /* 0: winners' cavity + ring regions are disjoint;
                         1: only a cavity may not meet another winner's cavity
                         or ring (rings may overlap) */
  int mix_below;      /* classic rounds: the segment phase ends once at most
                         this many subsegments are encroached; they join the
                         round's bad pass (0: the reference's round) */
  int seg_claim_iters; /* claim iterations of passes that split only
                          subsegments (0: claim_iters) */
  int deletion;       /* free vertices inside an encroached subsegment's
                         diametral circle are deleted before it splits:
                         0 never, 1 in Chew mode (Triangle's rule,
                         triangle.cpp:13319-13357), 2 in every mode (the
                         reference's, refine.cu:604-677) */
} orc_params;

typedef struct {
  int rounds, seg_passes, bad_passes;
  long long inserted, seg_splits, rej_encroach, rej_blocked, rej_overflow,
      lost, giveups;
  int final_cav_cap;
  int status; /* 0 ok, 1 round cap hit, 2 stalled, 3 converged except for
                 given-up elements (counted below) */
  long long cand_evals; /* bad-pass candidates evaluated */
  long long giveup_tris_left, giveup_segs_left; /* alive triangles still
                 flagged given-up; subsegments whose split was given up */
  long long deleted;    /* free vertices deleted */
  long long flips;      /* edge flips Lawson's algorithm would make for the
                 same insertions and deletions (triangle.h:283 numberofflips) */
} orc_stats;

typedef struct orc_result orc_result;

orc_result *orc_refine(int np, const double *points, int nt, const int *tris,
                       const int *nbrs, int ns, const int *segs,
                       const orc_params *prm, orc_stats *st);
orc_result *orc_refine_batch(int np, const double *points, int nt,
                             const int *tris, const int *nbrs, int ns,
                             const int *segs, const int *tri_inst, int n_inst,
                             const double *theta_deg, const int *offcenter,
                             const int *mode, const orc_params *prm,
                             orc_stats *st);
void orc_result_inst(const orc_result *r, int *tri_inst);
int orc_cavity(int np, const double *points, int nt, const int *tris,
               const int *nbrs, int ns, const int *segs, int t0,
               const double *c, int cap, int *out);
void orc_result_counts(const orc_result *r, int *np, int *nt, int *ns);
void orc_result_copy(const orc_result *r, double *points, int *tris, int *nbrs,
                     int *segs, uint8_t *pflag);
void orc_result_free(orc_result *r);

/* Circumcenter / off-center of triangle (apex, org, dest) as the refinement
 * computes it; returns 0 if degenerate. */
int orc_steiner_point(const double *apex, const double *org, const double *dest,
                      double theta_deg, int offcenter, double *out);

/* Value of Triangle's adaptive orient2d counterclockwise(pa, pb, pc)
 * (triangle.cpp:5155-5291): the sign is exact, the value is the first
 * adaptive stage's that certifies it. */
double orc_ccw_value(const double *pa, const double *pb, const double *pc);

/* The refinement's per-element rules on plain coordinates (for closed-form
 * tests): encroachment of segment a-b by p at theta degrees (ruppert = 1:
 * diametral circle, 0: Chew's lens); the split point of segment A-B given
 * which ends are shared with other subsegments (0 when it coincides with an
 * endpoint); the quality test of triangle (apex, org, dest) without the
 * Miller-Pav-Walkington exemption. */
int orc_encroaches_xy(const double *a, const double *b, const double *p, double theta_deg,
                      int ruppert);
int orc_split_point_xy(const double *A, const double *B, int acuteorg, int acutedest,
                       double *p);
int orc_bad_angle_xy(const double *apex, const double *org, const double *dest,
                     double theta_deg);

/* Invariant checker.  Returns 0 when every check passes; otherwise the
 * number of violations, with per-check counts in viol[] (see check_oracle.c
 * for the index meaning). */
long long orc_check_mesh(int np, const double *points, int nt, const int *tris,
                         const int *nbrs, int ns, const int *segs, int np_in,
                         const double *points_in, int ns_in,
                         const int *segs_in, double theta_deg, double seg_tol,
                         long long *viol, double *min_angle_deg,
                         long long *n_exempt);

#ifdef __cplusplus
}
#endif
#endif
