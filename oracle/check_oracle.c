/*
 * Mesh invariant checker (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * Restates the reference's end-of-run contract checks with exact predicates
 * and linear-time segment lookup:
 *   - input points preserved by index          main.cpp:336-344 (checkResult)
 *   - triangle / neighbour indices in range    main.cpp:348-362
 *   - neighbour links symmetric                main.cpp:79-113 (checkNeighbors)
 *   - constrained Delaunay: exact incircle <= 0 across every non-segment
 *     edge                                    main.cpp:153-242 (checkIncircle,
 *     which the reference leaves commented out and evaluates with
 *     uninitialised host predicates)
 *   - minimum angle >= theta                   main.cpp:280-328 (checkQuality),
 *     here with the Miller-Pav-Walkington exemption the refinement applies
 *     (triangle.cpp:7331-7378), which the reference checker omits
 * plus checks the reference does not make: every triangle positively
 * oriented, every boundary edge a subsegment, every output subsegment a mesh
 * edge, and every input segment covered end-to-end by a chain of collinear
 * output subsegments.
 *
 * viol[] indices: 0 index range, 1 neighbour symmetry, 2 orientation,
 * 3 unsegmented boundary edge, 4 Delaunay, 5 subsegment not an edge,
 * 6 input segment not covered, 7 input point moved, 8 quality, 9 non-manifold
 * edge.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NV 10

typedef struct {
  long long *k;
  int *v;
  size_t cap;
} EdgeMap;

static unsigned long long mixh(unsigned long long x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

static long long ekey(int a, int b) {
  long long lo = a < b ? a : b, hi = a < b ? b : a;
  return (lo << 32) | hi;
}

static void em_init(EdgeMap *m, size_t n) {
  m->cap = 1;
  while (m->cap < 2 * n + 16) m->cap <<= 1;
  m->k = malloc(sizeof(long long) * m->cap);
  m->v = malloc(sizeof(int) * m->cap);
  for (size_t i = 0; i < m->cap; i++) m->k[i] = -1;
}

static int *em_slot(EdgeMap *m, long long key, int create) {
  size_t h = (size_t)mixh((unsigned long long)key) & (m->cap - 1);
  while (m->k[h] != -1) {
    if (m->k[h] == key) return &m->v[h];
    h = (h + 1) & (m->cap - 1);
  }
  if (!create) return NULL;
  m->k[h] = key;
  m->v[h] = 0;
  return &m->v[h];
}

static void em_free(EdgeMap *m) {
  free(m->k);
  free(m->v);
}

long long orc_check_mesh(int np, const double *P, int nt, const int *T,
                         const int *N, int ns, const int *S, int np_in,
                         const double *P_in, int ns_in, const int *S_in,
                         double theta_deg, double seg_tol, long long *viol,
                         double *min_angle_deg, long long *n_exempt) {
  long long v[NV];
  memset(v, 0, sizeof(v));
  *n_exempt = 0;
  *min_angle_deg = 180.0;

  /* 7: input points preserved bit for bit */
  if (np < np_in) v[7] += np_in - np;
  for (int i = 0; i < np_in && i < np; i++)
    if (memcmp(P + 2 * (size_t)i, P_in + 2 * (size_t)i, 2 * sizeof(double)) != 0) v[7]++;

  /* 0: ranges */
  for (size_t i = 0; i < 3 * (size_t)nt; i++) {
    if (T[i] < 0 || T[i] >= np) v[0]++;
    if (N[i] < -1 || N[i] >= nt) v[0]++;
  }
  for (size_t i = 0; i < 2 * (size_t)ns; i++)
    if (S[i] < 0 || S[i] >= np) v[0]++;
  if (v[0]) goto out;

  /* subsegment set */
  EdgeMap seg;
  em_init(&seg, (size_t)ns);
  for (int s = 0; s < ns; s++) *em_slot(&seg, ekey(S[2 * s], S[2 * s + 1]), 1) += 1;

  /* edge multiplicity (9) */
  EdgeMap edg;
  em_init(&edg, 3 * (size_t)nt);
  for (int t = 0; t < nt; t++)
    for (int j = 0; j < 3; j++) {
      int a = T[3 * t + (j + 1) % 3], b = T[3 * t + (j + 2) % 3];
      int *c = em_slot(&edg, ekey(a, b), 1);
      *c += 1;
      if (*c == 3) v[9]++;
    }

  /* 5: every subsegment is an edge */
  for (int s = 0; s < ns; s++)
    if (!em_slot(&edg, ekey(S[2 * s], S[2 * s + 1]), 0)) v[5]++;

  for (int t = 0; t < nt; t++) {
    const int *tv = T + 3 * (size_t)t;
    /* 2: orientation */
    if (orc_orient2d(P + 2 * (size_t)tv[0], P + 2 * (size_t)tv[1], P + 2 * (size_t)tv[2]) <= 0) v[2]++;
    for (int j = 0; j < 3; j++) {
      int a = tv[(j + 1) % 3], b = tv[(j + 2) % 3];
      int u = N[3 * t + j];
      int is_seg = em_slot(&seg, ekey(a, b), 0) != NULL;
      if (u < 0) {
        if (!is_seg) v[3]++;
        continue;
      }
      /* 1: symmetry -- u must hold the reversed edge with t as neighbour */
      int found = -1;
      for (int k = 0; k < 3; k++)
        if (T[3 * u + (k + 1) % 3] == b && T[3 * u + (k + 2) % 3] == a) found = k;
      if (found < 0 || N[3 * u + found] != t) {
        v[1]++;
        continue;
      }
      /* 4: Delaunay across non-segment edges (checked once per edge) */
      if (!is_seg && t < u) {
        int d = T[3 * u + found];
        if (orc_incircle(P + 2 * (size_t)tv[0], P + 2 * (size_t)tv[1], P + 2 * (size_t)tv[2],
                         P + 2 * (size_t)d) > 0)
          v[4]++;
      }
    }
  }
  em_free(&edg);

  /* 6: segment conformity.  Adjacency of output subsegments; walk every
   * input segment from its origin to its destination along subsegments whose
   * vertices lie on the segment (distance <= seg_tol * length). */
  int *deg = calloc((size_t)np + 1, sizeof(int));
  for (int s = 0; s < ns; s++) { deg[S[2 * s]]++; deg[S[2 * s + 1]]++; }
  int *off = malloc(sizeof(int) * ((size_t)np + 1));
  off[0] = 0;
  for (int i = 0; i < np; i++) off[i + 1] = off[i] + deg[i];
  int *adj = malloc(sizeof(int) * (size_t)(2 * (size_t)ns + 1));
  int *fill = calloc((size_t)np + 1, sizeof(int));
  for (int s = 0; s < ns; s++) {
    int a = S[2 * s], b = S[2 * s + 1];
    adj[off[a] + fill[a]++] = b;
    adj[off[b] + fill[b]++] = a;
  }
  int *onseg = malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
  for (int i = 0; i < np; i++) onseg[i] = -1;
  long long used = 0;
  for (int q = 0; q < ns_in; q++) {
    int A = S_in[2 * q], B = S_in[2 * q + 1];
    const double *pa = P + 2 * (size_t)A, *pb = P + 2 * (size_t)B;
    double dx = pb[0] - pa[0], dy = pb[1] - pa[1];
    double len2 = dx * dx + dy * dy, len = sqrt(len2);
    int cur = A;
    double tcur = 0.0;
    int steps = 0, ok = 0;
    while (steps++ <= np) {
      if (cur == B) { ok = 1; break; }
      int best = -1;
      double tb = 2.0;
      for (int e = off[cur]; e < off[cur + 1]; e++) {
        int w = adj[e];
        const double *pw = P + 2 * (size_t)w;
        double tw = ((pw[0] - pa[0]) * dx + (pw[1] - pa[1]) * dy) / len2;
        double dist = fabs((pw[0] - pa[0]) * dy - (pw[1] - pa[1]) * dx) / len;
        if (dist <= seg_tol * len && tw > tcur && tw < tb) { tb = tw; best = w; }
      }
      if (best < 0) break;
      if (best != B) onseg[best] = q;
      cur = best;
      tcur = tb;
      used++;
    }
    if (!ok) v[6]++;
  }
  if (used != ns) v[6] += (used > ns ? used - ns : ns - used);
  free(deg);
  free(off);
  free(adj);
  free(fill);

  /* 8: quality with the MPW exemption (shortest edge joins two segment
   * interior vertices on different input segments that share an endpoint,
   * equidistant from it within 0.1%) */
  {
    long double ga = cosl((long double)theta_deg * 3.14159265358979323846264338327950288L / 180.0L);
    ga *= ga;
    double minang = 180.0;
    for (int t = 0; t < nt; t++) {
      const int *tv = T + 3 * (size_t)t;
      long double x[3], y[3], l[3];
      for (int k = 0; k < 3; k++) { x[k] = P[2 * (size_t)tv[k]]; y[k] = P[2 * (size_t)tv[k] + 1]; }
      for (int k = 0; k < 3; k++) {
        int a = (k + 1) % 3, b = (k + 2) % 3;
        l[k] = (x[a] - x[b]) * (x[a] - x[b]) + (y[a] - y[b]) * (y[a] - y[b]);
      }
      int e = 0;
      if (l[1] < l[e]) e = 1;
      if (l[2] < l[e]) e = 2;
      int a = (e + 1) % 3, b = (e + 2) % 3;
      long double dot = (x[a] - x[e]) * (x[b] - x[e]) + (y[a] - y[e]) * (y[b] - y[e]);
      long double cos2 = dot * dot / (l[a] * l[b]);
      double ang = (double)(acosl(sqrtl(cos2 > 1 ? 1 : cos2)) * 180.0L / 3.14159265358979323846264338327950288L);
      if (cos2 > ga * (1.0L + 1e-9L)) {
        int va = tv[a], vb = tv[b];
        int exempt = 0;
        if (va >= np_in && vb >= np_in && onseg[va] >= 0 && onseg[vb] >= 0 &&
            !em_slot(&seg, ekey(va, vb), 0)) {
          int s1 = onseg[va], s2 = onseg[vb];
          int o1 = S_in[2 * s1], d1 = S_in[2 * s1 + 1], o2 = S_in[2 * s2], d2 = S_in[2 * s2 + 1];
          int common = -1;
          if (o1 == o2 || o1 == d2) common = o1;
          else if (d1 == o2 || d1 == d2) common = d1;
          if (common >= 0) {
            const double *J = P + 2 * (size_t)common, *A = P + 2 * (size_t)va, *B = P + 2 * (size_t)vb;
            double d1s = (A[0] - J[0]) * (A[0] - J[0]) + (A[1] - J[1]) * (A[1] - J[1]);
            double d2s = (B[0] - J[0]) * (B[0] - J[0]) + (B[1] - J[1]) * (B[1] - J[1]);
            if (d1s < 1.0011 * d2s && d1s > 0.9989 * d2s) exempt = 1;
          }
        }
        if (exempt) (*n_exempt)++;
        else {
          v[8]++;
          if (ang < minang) minang = ang;
        }
      } else if (ang < minang) {
        minang = ang;
      }
    }
    *min_angle_deg = minang;
  }
  free(onseg);
  em_free(&seg);
out:;
  long long total = 0;
  for (int i = 0; i < NV; i++) {
    if (viol) viol[i] = v[i];
    total += v[i];
  }
  return total;
}
