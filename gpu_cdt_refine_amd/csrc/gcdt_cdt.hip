// Constrained Delaunay triangulation of a PSLG on the GPU: the step before
// the refinement, which the reference leaves to Triangle on the host
// (triangulate("pzQnc"), mesh.cpp:397, main.cpp:454).  Output convention as
// Triangle's -pzQnc: the input points unchanged, CCW triangles, neighbour j
// opposite corner j (-1 on the hull), and the segments: the input segments
// (split where they pass through an input point) followed by the convex
// hull edges that are not already segments.
//
// Phase 1, Delaunay triangulation by rounds of parallel Bowyer-Watson
// insertion.  A ghost vertex G closes the triangulation into a sphere: every
// hull edge (a, b) has a ghost triangle (b, a, G), so the convex hull is
// exact and no bounding triangle is needed.  The rounds start from the
// Delaunay triangulation of a seed (the exact convex hull's vertices and a
// fixed pseudo-random 1/128 of the points, triangulated on the host), so
// every other point lies inside the hull.  Every uninserted point knows a
// triangle containing it (or, on a hull edge, a ghost triangle).  A round:
// one candidate point per triangle (the one nearest its centroid), cavities
// grown from the candidates' triangles, claims with the shared-ring rule in
// several iterations (64-bit atomicMax of epoch-tagged keys), star insertion
// of the winners, relocation of the points whose triangle was rewritten, by
// a walk from that slot.  A triangulation of n >= 3 points has exactly
// 2n - 2 triangles with the ghost ones, so the slots are allocated once.
//
// Phase 2, constraint recovery: every segment missing from the
// triangulation takes the triangles its interior crosses, claims them and
// their ring, and the winners retriangulate the two pseudo-polygons on either
// side with the constrained-Delaunay rule (Anglada's: the chain vertex whose
// circle with the base edge holds no other chain vertex), which keeps the
// triangulation constrained Delaunay.  A segment through an input point is
// split there.
//
// Predicates: exact orient2d everywhere; incircle through the filters and
// the double-double stage, candidates left undecided are redone by a small
// grid with the exact expansion arena.  Every index read from memory is
// range-checked before use: a broken invariant becomes an error code, never
// an out-of-bounds access, and every walk has a step cap.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "gcdt_internal.h"
#include "gpu_cdt_refine.h"
#include "predicates.h"

namespace gcdt {
namespace cdtgpu {

#define CC(x)                                                                   \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess)                                                       \
      throw std::runtime_error(std::string("HIP error in CDT construction: ") + \
                               hipGetErrorString(e_) + " at " #x);              \
  } while (0)

// error codes (bits of the device error word)
enum : unsigned {
  E_INDEX = 1u,     // an index read from memory was out of range
  E_WALK = 2u,      // a walk exceeded its step cap
  E_CAVITY = 4u,    // a cavity or segment region exceeded its capacity
  E_TOPOLOGY = 8u,  // inconsistent neighbour links
  E_DUPLICATE = 16u // two input points coincide
};

constexpr int BLK = 256;
constexpr int KCAV = 64;          // cavity capacity (triangles) of an insertion
constexpr int WS_INS = 4 * KCAV + 8;  // per-candidate ints: cavity | boundary u, v, outer
constexpr int EXACT_GRID = 512;   // blocks of 64 of the exact-stage kernels
constexpr long long WALK_CAP = 1LL << 16;

struct MeshG {
  const double2 *P;
  int np;   // input points; the ghost vertex is np
  int4 *V;  // triangle vertices (w unused)
  int4 *N;  // neighbour otri (t << 2 | edge in the neighbour), -1 none
  int cap;  // triangle slots (2 np - 2)
  unsigned long long *claim;
  unsigned *err;
};

__host__ __device__ inline int comp4(const int4 &v, int j) {
  return j == 0 ? v.x : (j == 1 ? v.y : v.z);
}
__device__ inline void set4(int4 &v, int j, int x) {
  if (j == 0) v.x = x;
  else if (j == 1) v.y = x;
  else v.z = x;
}
__device__ inline void flag(const MeshG &m, unsigned e) { atomicOr(m.err, e); }
__device__ inline bool tri_ok(const MeshG &m, int t) { return t >= 0 && t < m.cap; }
__device__ inline bool vert_ok(const MeshG &m, int v) { return v >= 0 && v <= m.np; }

// Orientation of (a, b, p) with a, b real vertices.
__device__ inline int orient_v(const MeshG &m, int a, int b, double2 p) {
  const double2 A = m.P[a], B = m.P[b];
  return orient2d_sign(A.x, A.y, B.x, B.y, p.x, p.y, false, nullptr);
}

// p strictly inside segment (a, b), p collinear with it
__device__ inline bool strictly_between(const MeshG &m, int a, int b, double2 p) {
  const double2 A = m.P[a], B = m.P[b];
  // the dot products of (p - a, b - a) and (p - b, a - b) are both positive;
  // p is on the line, so their signs are exact when both are far from 0 and
  // exactly decidable by comparing coordinates otherwise
  if (A.x != B.x) return (p.x > fmin(A.x, B.x)) && (p.x < fmax(A.x, B.x));
  return (p.y > fmin(A.y, B.y)) && (p.y < fmax(A.y, B.y));
}

// Conflict of point p with triangle t: 1 yes, 0 no, PRED_UNSURE undecided by
// the filters (FAST) -- with ws given, always decided.
template <bool FAST>
__device__ int conflict(const MeshG &m, int t, double2 p, double *ws) {
  const int4 v = m.V[t];
  int g = -1;
  if (v.x == m.np) g = 0;
  else if (v.y == m.np) g = 1;
  else if (v.z == m.np) g = 2;
  if (g >= 0) {  // ghost triangle (a, b, G): the open half-plane right of b -> a ...
    const int a = comp4(v, (g + 1) % 3), b = comp4(v, (g + 2) % 3);
    const int o = orient_v(m, a, b, p);
    if (o > 0) return 1;  // p strictly outside the hull edge (b, a)
    if (o < 0) return 0;
    return strictly_between(m, a, b, p) ? 1 : 0;
  }
  const double2 A = m.P[v.x], B = m.P[v.y], C = m.P[v.z];
  int s = incircle_filtered(A.x, A.y, B.x, B.y, C.x, C.y, p.x, p.y, false, nullptr);
  if (s == PRED_UNSURE) {
    s = incircle_dd_sign(A.x, A.y, B.x, B.y, C.x, C.y, p.x, p.y);
    if (s == 0) {
      if (FAST) return PRED_UNSURE;
      double e = incircle_exact(A.x, A.y, B.x, B.y, C.x, C.y, p.x, p.y, ws);
      s = e > 0.0 ? 1 : (e < 0.0 ? -1 : 0);
    }
  }
  return s > 0 ? 1 : 0;
}

// Visibility walk from triangle t toward p: returns a triangle in conflict
// with p that contains it (real) or sees it (ghost), -1 on an error; *on_vertex
// is set when p coincides with a vertex.
__device__ int locate(const MeshG &m, int t, double2 p, bool *on_vertex) {
  *on_vertex = false;
  for (long long step = 0; step < WALK_CAP; step++) {
    if (!tri_ok(m, t)) {
      flag(m, E_INDEX);
      return -1;
    }
    const int4 v = m.V[t];
    const int4 n = m.N[t];
    int g = -1;
    if (v.x == m.np) g = 0;
    else if (v.y == m.np) g = 1;
    else if (v.z == m.np) g = 2;
    if (g >= 0) {
      const int a = comp4(v, (g + 1) % 3), b = comp4(v, (g + 2) % 3);
      if (!vert_ok(m, a) || !vert_ok(m, b) || a == m.np || b == m.np) {
        flag(m, E_INDEX);
        return -1;
      }
      const int o = orient_v(m, a, b, p);
      if (o > 0 || (o == 0 && strictly_between(m, a, b, p))) return t;
      const double2 A = m.P[a], B = m.P[b];
      if ((A.x == p.x && A.y == p.y) || (B.x == p.x && B.y == p.y)) {
        *on_vertex = true;
        return t;
      }
      const int nb = comp4(n, g);  // the real triangle across (a, b)
      if (nb < 0) {
        flag(m, E_TOPOLOGY);
        return -1;
      }
      t = nb >> 2;
      continue;
    }
    int jn = -1, zeros = 0;
    for (int j = 0; j < 3; j++) {
      const int a = comp4(v, (j + 1) % 3), b = comp4(v, (j + 2) % 3);
      if (!vert_ok(m, a) || !vert_ok(m, b)) {
        flag(m, E_INDEX);
        return -1;
      }
      const int o = orient_v(m, a, b, p);
      if (o < 0 && jn < 0) jn = j;
      zeros += o == 0;
    }
    if (jn < 0) {
      if (zeros >= 2) *on_vertex = true;
      return t;
    }
    const int nb = comp4(n, jn);
    if (nb < 0) {
      flag(m, E_TOPOLOGY);
      return -1;
    }
    t = nb >> 2;
  }
  flag(m, E_WALK);
  return -1;
}

__device__ inline bool in_list(const int *L, int n, int x) {
  for (int i = 0; i < n; i++)
    if (L[i] == x) return true;
  return false;
}


// ---------------------------------------------------------------------------
// Phase 1: rounds of parallel point insertion
// ---------------------------------------------------------------------------
__device__ inline unsigned hash16(unsigned i) {
  unsigned h = i * 0x9E3779B1u;
  h ^= h >> 15;
  h *= 0x85EBCA77u;
  h ^= h >> 13;
  return h & 0xffffu;
}
// One candidate per triangle: the point nearest the triangle's centroid
// (ghost triangles: the edge's midpoint), so that a triangle holding many
// points is split into parts of similar size; ties by the smaller index.
// Key (largest wins): round epoch | inverted float bits of the squared
// distance | inverted index.
__device__ inline unsigned long long pick_key(const MeshG &m, unsigned long long ep, int i, int t) {
  const int4 v = m.V[t];
  const double2 p = m.P[i];
  double cx = 0.0, cy = 0.0;
  int k = 0;
  for (int j = 0; j < 3; j++) {
    const int x = comp4(v, j);
    if (x >= 0 && x < m.np) {
      cx += m.P[x].x;
      cy += m.P[x].y;
      k++;
    }
  }
  cx /= k > 0 ? k : 1;
  cy /= k > 0 ? k : 1;
  const float d = (float)((p.x - cx) * (p.x - cx) + (p.y - cy) * (p.y - cy));
  const unsigned db = __float_as_uint(d) >> 12;  // non-negative: bits order like values
  return (ep << 52) | ((unsigned long long)(0xfffffu - db) << 32) |
         (unsigned long long)(0xffffffffu - (unsigned)i);
}

__global__ void k_pick(MeshG m, const int *U, int nu, const int *ptri,
                       unsigned long long *pick, unsigned long long ep) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nu; k += gridDim.x * blockDim.x) {
    const int i = U[k];
    const int t = ptri[i];
    if (!tri_ok(m, t)) {
      flag(m, E_INDEX);
      continue;
    }
    atomicMax(&pick[t], pick_key(m, ep, i, t));
  }
}

__global__ void k_cand_flags(MeshG m, const int *U, int nu, const int *ptri,
                             const unsigned long long *pick, unsigned long long ep, int *flag_out) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nu; k += gridDim.x * blockDim.x) {
    const int i = U[k];
    const int t = ptri[i];
    flag_out[k] = tri_ok(m, t) && pick[t] == pick_key(m, ep, i, t);
  }
}

enum { CS_NONE = 0, CS_VALID = 1, CS_DEFER = 2, CS_OVERFLOW = 3 };

// Cavity of candidate point c seeded at its triangle: conflicting triangles
// reachable across edges (ghost triangles by the half-plane rule), sorted by
// slot; the boundary in star order (starting at the edge with the smallest
// origin vertex), written to the candidate's workspace as cavity | u | v |
// outer.  FAST: undecided incircle signs defer the candidate.
template <bool FAST>
__device__ void cavity_one(const MeshG &m, int c, int seed, int *W, int *status, int *ncav,
                           double *ws) {
  const double2 p = m.P[c];
  int *L = W;
  int *Bu = W + KCAV, *Bv = Bu + (KCAV + 2), *Bo = Bv + (KCAV + 2);
  if (!tri_ok(m, seed)) {
    flag(m, E_INDEX);
    *status = CS_NONE;
    return;
  }
  int c0 = conflict<FAST>(m, seed, p, ws);
  if (c0 == PRED_UNSURE) {
    *status = CS_DEFER;
    return;
  }
  if (c0 != 1) {  // the stored triangle must contain (or see) the point
    flag(m, E_TOPOLOGY);
    *status = CS_NONE;
    return;
  }
  int n = 0;
  L[n++] = seed;
  for (int idx = 0; idx < n; idx++) {
    const int4 nb4 = m.N[L[idx]];
    for (int j = 0; j < 3; j++) {
      const int nb = comp4(nb4, j);
      if (nb < 0) continue;
      const int u = nb >> 2;
      if (!tri_ok(m, u)) {
        flag(m, E_INDEX);
        *status = CS_NONE;
        return;
      }
      if (in_list(L, n, u)) continue;
      const int cf = conflict<FAST>(m, u, p, ws);
      if (cf == PRED_UNSURE) {
        *status = CS_DEFER;
        return;
      }
      if (cf == 1) {
        if (n == KCAV) {
          *status = CS_OVERFLOW;
          return;
        }
        L[n++] = u;
      }
    }
  }
  // slot order (insertion sort; n is small)
  for (int a = 1; a < n; a++) {
    const int x = L[a];
    int b = a - 1;
    while (b >= 0 && L[b] > x) {
      L[b + 1] = L[b];
      b--;
    }
    L[b + 1] = x;
  }
  // boundary edges
  int nb_cnt = 0;
  int tu[KCAV + 2], tv[KCAV + 2], to[KCAV + 2];
  for (int idx = 0; idx < n; idx++) {
    const int t = L[idx];
    const int4 v = m.V[t], nn = m.N[t];
    for (int j = 0; j < 3; j++) {
      const int nb = comp4(nn, j);
      if (nb >= 0 && in_list(L, n, nb >> 2)) continue;
      if (nb_cnt == KCAV + 2) {
        flag(m, E_TOPOLOGY);
        *status = CS_NONE;
        return;
      }
      tu[nb_cnt] = comp4(v, (j + 1) % 3);
      tv[nb_cnt] = comp4(v, (j + 2) % 3);
      to[nb_cnt] = nb;
      nb_cnt++;
    }
  }
  if (nb_cnt != n + 2) {  // a Delaunay cavity is a disk
    flag(m, E_TOPOLOGY);
    *status = CS_NONE;
    return;
  }
  int start = 0;
  for (int e = 1; e < nb_cnt; e++)
    if (tu[e] < tu[start]) start = e;
  int e = start;
  for (int q = 0; q < nb_cnt; q++) {
    Bu[q] = tu[e];
    Bv[q] = tv[e];
    Bo[q] = to[e];
    int f = 0;
    while (f < nb_cnt && tu[f] != tv[e]) f++;
    if (f == nb_cnt) {
      flag(m, E_TOPOLOGY);
      *status = CS_NONE;
      return;
    }
    e = f;
  }
  if (e != start) {
    flag(m, E_TOPOLOGY);
    *status = CS_NONE;
    return;
  }
  *ncav = n;
  *status = CS_VALID;
}

__global__ void k_ins_cavity_fast(MeshG m, const int *C, int nc, const int *ptri, int *W,
                                  int *cstat, int *cn, int *slow, int *nslow) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nc; k += gridDim.x * blockDim.x) {
    const int c = C[k];
    int st = CS_NONE, n = 0;
    cavity_one<true>(m, c, ptri[c], W + (size_t)k * WS_INS, &st, &n, nullptr);
    cstat[k] = st;
    cn[k] = n;
    if (st == CS_DEFER) slow[atomicAdd(nslow, 1)] = k;
  }
}

// the deferred candidates with the exact stage (GCDT_INCIRCLE_WS doubles of
// arena per lane of a grid of at most EXACT_GRID blocks of 64)
__global__ void k_ins_cavity_slow(MeshG m, const int *C, const int *ptri, int *W, int *cstat,
                                  int *cn, const int *slow, const int *nslow, double *arena) {
  const int lane = blockIdx.x * blockDim.x + threadIdx.x;
  double *ws = arena + (size_t)lane * GCDT_INCIRCLE_WS;
  const int ns = *nslow;
  for (int q = lane; q < ns; q += gridDim.x * blockDim.x) {
    const int k = slow[q];
    const int c = C[k];
    int st = CS_NONE, n = 0;
    cavity_one<false>(m, c, ptri[c], W + (size_t)k * WS_INS, &st, &n, ws);
    cstat[k] = st;
    cn[k] = n;
  }
}

__device__ inline unsigned long long claim_key(unsigned long long ep, int k) {
  return (ep << 32) | (unsigned long long)(0xffffffffu - (unsigned)k);
}

// Claims, shared-ring rule (two words per triangle): cavity triangles claim
// word 0, ring triangles word 1.  A candidate wins when it holds word 0 of
// every cavity triangle and no better candidate holds one of them as ring
// (word 1), and no better candidate holds one of its ring triangles as
// cavity: winners' cavities are disjoint and never meet another winner's
// ring, so their stars are independent; rings may be shared (each winner
// repairs its own edge of a shared ring triangle).  The smallest candidate
// index is best.  Several iterations per round: a winner stamps its region
// with the round's sticky epoch; a loser whose region meets a stamp drops
// out, the others claim again.
__global__ void k_ins_claim(MeshG m, int nc, const int *W, const int *cstat, const int *cn,
                            int *active, int iter, unsigned long long ep,
                            unsigned long long sticky) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nc; k += gridDim.x * blockDim.x) {
    const int *L = W + (size_t)k * WS_INS;
    const int *Bo = L + KCAV + 2 * (KCAV + 2);
    const int n = cn[k];
    if (iter == 0) {
      active[k] = cstat[k] == CS_VALID;
    } else if (active[k]) {
      bool hit = false;
      for (int q = 0; q < n && !hit; q++)
        hit = (m.claim[2 * (size_t)L[q]] >> 32) == sticky ||
              (m.claim[2 * (size_t)L[q] + 1] >> 32) == sticky;
      for (int q = 0; q < n + 2 && !hit; q++)
        if (Bo[q] >= 0) hit = (m.claim[2 * (size_t)(Bo[q] >> 2)] >> 32) == sticky;
      if (hit) active[k] = 0;
    }
    if (!active[k]) continue;
    const unsigned long long key = claim_key(ep, k);
    for (int q = 0; q < n; q++) atomicMax(&m.claim[2 * (size_t)L[q]], key);
    for (int q = 0; q < n + 2; q++)
      if (Bo[q] >= 0) atomicMax(&m.claim[2 * (size_t)(Bo[q] >> 2) + 1], key);
  }
}

__global__ void k_ins_win(MeshG m, int nc, const int *W, const int *cn, int *active, int iter,
                          unsigned long long ep, unsigned long long sticky, int *win) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nc; k += gridDim.x * blockDim.x) {
    if (iter == 0) win[k] = 0;
    if (!active[k]) continue;
    const int *L = W + (size_t)k * WS_INS;
    const int *Bo = L + KCAV + 2 * (KCAV + 2);
    const int n = cn[k];
    const unsigned long long key = claim_key(ep, k);
    bool w = true;
    for (int q = 0; q < n && w; q++)
      w = m.claim[2 * (size_t)L[q]] == key && m.claim[2 * (size_t)L[q] + 1] <= key;
    for (int q = 0; q < n + 2 && w; q++)
      if (Bo[q] >= 0) w = m.claim[2 * (size_t)(Bo[q] >> 2)] <= key;
    if (!w) continue;
    win[k] = 1;
    active[k] = 0;
    const unsigned long long stamp = (sticky << 32) | 0xffffffffull;
    for (int q = 0; q < n; q++) m.claim[2 * (size_t)L[q]] = stamp;
    for (int q = 0; q < n + 2; q++)
      if (Bo[q] >= 0) m.claim[2 * (size_t)(Bo[q] >> 2) + 1] = stamp;
  }
}

// Star insertion of a winner: star triangle q = (p, u_q, v_q) takes cavity
// slot L[q] (q < n) or new slot tbase + 2 rank + (q - n); the outer
// neighbours' links are repaired; the star's slots are stamped so that the
// points they held are relocated.
__global__ void k_ins_apply(MeshG m, const int *C, int nc, const int *W, const int *cn,
                            const int *win, const int *rank, int tbase, int *ptri,
                            unsigned long long *dirty, unsigned long long ep) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nc; k += gridDim.x * blockDim.x) {
    if (!win[k]) continue;
    const int c = C[k];
    const int *L = W + (size_t)k * WS_INS;
    const int *Bu = L + KCAV, *Bv = Bu + (KCAV + 2), *Bo = Bv + (KCAV + 2);
    const int n = cn[k], ns = n + 2;
    const int tb = tbase + 2 * rank[k];
    if (tb < 0 || tb + 1 >= m.cap) {
      flag(m, E_INDEX);
      continue;
    }
    for (int q = 0; q < ns; q++) {
      const int st = q < n ? L[q] : tb + (q - n);
      const int qn = q + 1 < ns ? q + 1 : 0, qp = q > 0 ? q - 1 : ns - 1;
      const int sn = qn < n ? L[qn] : tb + (qn - n), sp = qp < n ? L[qp] : tb + (qp - n);
      m.V[st] = make_int4(c, Bu[q], Bv[q], 0);
      m.N[st] = make_int4(Bo[q], (sn << 2) | 2, (sp << 2) | 1, 0);
      dirty[st] = ep;
      const int o = Bo[q];
      if (o >= 0) {
        const int r = o >> 2, j = o & 3;
        int *slot = reinterpret_cast<int *>(&m.N[r]) + j;
        *slot = (st << 2) | 0;
      }
    }
    ptri[c] = -1;
  }
}

// Points whose triangle was rewritten this round walk from that slot (now a
// triangle of the star that replaced it) to their new triangle.
__global__ void k_relocate(MeshG m, const int *U, int nu, int *ptri,
                           const unsigned long long *dirty, unsigned long long ep) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nu; k += gridDim.x * blockDim.x) {
    const int i = U[k];
    const int t = ptri[i];
    if (t < 0) continue;
    if (!tri_ok(m, t)) {
      flag(m, E_INDEX);
      continue;
    }
    if (dirty[t] != ep) continue;
    bool on_vertex = false;
    const int t2 = locate(m, t, m.P[i], &on_vertex);
    if (on_vertex) {
      flag(m, E_DUPLICATE);
      ptri[i] = -2;
      continue;
    }
    ptri[i] = t2 < 0 ? -2 : t2;
  }
}

__global__ void k_still_out(const int *U, int nu, const int *ptri, int *flag_out) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nu; k += gridDim.x * blockDim.x)
    flag_out[k] = ptri[U[k]] >= 0;
}

// ---------------------------------------------------------------------------
// Ordered compaction (exclusive scan of 0/1 flags): out[rank] = src[k] for
// flagged k, in ascending k; rank_out (optional) receives each k's rank.
// ---------------------------------------------------------------------------
constexpr int SCAN_CHUNK = 1024;

__device__ inline int block_exclusive_scan(int x, int *tmp, int *total) {
  const int t = threadIdx.x;
  tmp[t] = x;
  __syncthreads();
  for (int off = 1; off < SCAN_CHUNK; off <<= 1) {
    int y = t >= off ? tmp[t - off] : 0;
    __syncthreads();
    tmp[t] += y;
    __syncthreads();
  }
  const int incl = tmp[t];
  if (total) *total = tmp[SCAN_CHUNK - 1];
  __syncthreads();
  return incl - x;
}
__global__ __launch_bounds__(SCAN_CHUNK) void k_scan_sums(const int *f, int n, int *bsum) {
  __shared__ int tmp[SCAN_CHUNK];
  const long long i = (long long)blockIdx.x * SCAN_CHUNK + threadIdx.x;
  int tot = 0;
  block_exclusive_scan(i < n ? f[i] : 0, tmp, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}
__global__ __launch_bounds__(SCAN_CHUNK) void k_scan_top(int *bsum, int nb, int *total) {
  __shared__ int tmp[SCAN_CHUNK];
  int carry = 0;
  for (int base = 0; base < nb; base += SCAN_CHUNK) {
    const int i = base + threadIdx.x;
    int tot = 0;
    const int ex = block_exclusive_scan(i < nb ? bsum[i] : 0, tmp, &tot);
    if (i < nb) bsum[i] = ex + carry;
    carry += tot;
  }
  if (threadIdx.x == 0) *total = carry;
}
__global__ __launch_bounds__(SCAN_CHUNK) void k_scan_scatter(const int *f, int n, const int *bsum,
                                                              const int *src, int *out,
                                                              int *rank_out) {
  __shared__ int tmp[SCAN_CHUNK];
  const long long i = (long long)blockIdx.x * SCAN_CHUNK + threadIdx.x;
  const int fl = i < n ? f[i] : 0;
  const int ex = block_exclusive_scan(fl, tmp, nullptr) + bsum[blockIdx.x];
  if (i < n) {
    if (rank_out) rank_out[i] = ex;
    if (fl && out) out[ex] = src ? src[i] : (int)i;
  }
}

// ---------------------------------------------------------------------------
// Phase 2: constraint recovery
// ---------------------------------------------------------------------------
// Region capacity: crossed triangles of one segment region.  Passes start
// at KSEG0; a region that does not fit reports SS_OVER, and once a pass
// makes no progress the host grows the capacity (fewer segments a pass).
constexpr int KSEG0 = 512;
// per-segment workspace for capacity k: region | left chain | right chain |
// boundary u, v, outer | new triangles (3 ints) | Anglada stack (4 ints)
__host__ __device__ constexpr size_t ws_seg(int k) {
  return (size_t)k + 2 * ((size_t)k + 2) + 3 * ((size_t)k + 2) + 3 * (size_t)k + 4 * ((size_t)k + 2);
}
__host__ __device__ constexpr size_t bo_off(int k) { return (size_t)k + 4 * ((size_t)k + 2); }

enum { SS_PENDING = 0, SS_EXISTS = 1, SS_REGION = 2, SS_SPLIT_EDGE = 3, SS_ERROR = 4,
       SS_DEFER = 5, SS_OVER = 6 };
constexpr int SLIT = -3;  // boundary entry whose far side is in the region

// Triangles incident to each vertex: the smallest slot (deterministic).
__global__ void k_vtri(MeshG m, int nt, int *vtri) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
    const int4 v = m.V[t];
    for (int j = 0; j < 3; j++) {
      const int x = comp4(v, j);
      if (x >= 0 && x < m.np) atomicMin(&vtri[x], t);
    }
  }
}

__device__ inline int index_of(const int4 &v, int x) {
  return v.x == x ? 0 : (v.y == x ? 1 : (v.z == x ? 2 : -1));
}

struct SegOut {
  int status;
  int z;     // the vertex where the segment is split (SS_SPLIT_EDGE / region end), or b
  int n;     // crossed triangles
  int nl, nr;  // chain lengths (left, right; endpoints excluded)
  int nbnd;
  int t_edge, j_edge;  // SS_EXISTS / SS_SPLIT_EDGE: a triangle holding edge (a, z) and its index
};

// Segment a -> b: find whether (a, b) is an edge; otherwise collect the
// triangles its interior crosses up to b or up to the first vertex z lying on
// it.  Chains: left (vertices left of a->b) and right, in order from a.
// Boundary edges of the region with their outer otri.
__device__ SegOut seg_scan(const MeshG &m, int a, int b, const int *vtri, int *W, int kseg) {
  SegOut o{SS_ERROR, -1, 0, 0, 0, 0, -1, -1};
  int *R = W;
  int *CL = R + kseg, *CR = CL + (kseg + 2);
  int *Bu = CR + (kseg + 2), *Bv = Bu + (kseg + 2), *Bo = Bv + (kseg + 2);
  const SegOut over{SS_OVER, -1, 0, 0, 0, 0, -1, -1};
  if (a < 0 || a >= m.np || b < 0 || b >= m.np || a == b) {
    flag(m, E_INDEX);
    return o;
  }
  const double2 pb = m.P[b], pa = m.P[a];
  int t = vtri[a];
  if (!tri_ok(m, t)) {
    flag(m, E_INDEX);
    return o;
  }
  // rotate counterclockwise around a until b lies in the wedge of a triangle
  int x = -1, y = -1, t0 = -1;
  for (int guard = 0; guard < 65536; guard++) {
    const int4 v = m.V[t];
    const int ia = index_of(v, a);
    if (ia < 0) {
      flag(m, E_TOPOLOGY);
      return o;
    }
    const int xv = comp4(v, (ia + 1) % 3), yv = comp4(v, (ia + 2) % 3);
    if (xv == b || yv == b) {
      o.status = SS_EXISTS;
      o.z = b;
      o.t_edge = t;
      o.j_edge = xv == b ? (ia + 2) % 3 : (ia + 1) % 3;  // edge (a, b) is opposite the other vertex
      return o;
    }
    if (xv != m.np && yv != m.np) {
      if (!vert_ok(m, xv) || !vert_ok(m, yv)) {
        flag(m, E_INDEX);
        return o;
      }
      const double2 X = m.P[xv], Y = m.P[yv];
      const int o1 = orient2d_sign(pa.x, pa.y, X.x, X.y, pb.x, pb.y, false, nullptr);
      const int o2 = orient2d_sign(pa.x, pa.y, Y.x, Y.y, pb.x, pb.y, false, nullptr);
      // a vertex on the ray a -> b: the segment is split there (it lies
      // between a and b: b is a vertex, so it cannot be beyond b on an edge)
      if (o1 == 0 && (X.x - pa.x) * (pb.x - pa.x) + (X.y - pa.y) * (pb.y - pa.y) > 0) {
        o.status = SS_SPLIT_EDGE;
        o.z = xv;
        o.t_edge = t;
        o.j_edge = (ia + 2) % 3;
        return o;
      }
      if (o2 == 0 && (Y.x - pa.x) * (pb.x - pa.x) + (Y.y - pa.y) * (pb.y - pa.y) > 0) {
        o.status = SS_SPLIT_EDGE;
        o.z = yv;
        o.t_edge = t;
        o.j_edge = (ia + 1) % 3;
        return o;
      }
      if (o1 > 0 && o2 < 0) {
        x = xv;
        y = yv;
        t0 = t;
        break;
      }
    }
    const int nb = comp4(m.N[t], (ia + 1) % 3);  // across (a, y): the next triangle ccw
    if (nb < 0 || !tri_ok(m, nb >> 2)) {
      flag(m, E_TOPOLOGY);
      return o;
    }
    t = nb >> 2;
  }
  if (t0 < 0) {
    flag(m, E_WALK);
    return o;
  }
  // walk across the triangles the segment crosses; x right of a->b, y left
  int n = 0, nl = 0, nr = 0, nbnd = 0;
  auto add_bnd = [&](int tt, int j) -> bool {
    if (nbnd == kseg + 2) return false;
    const int ob = comp4(m.N[tt], j);
    if (ob >= 0 && !tri_ok(m, ob >> 2)) {
      flag(m, E_INDEX);
      return false;
    }
    const int4 v = m.V[tt];
    Bu[nbnd] = comp4(v, (j + 1) % 3);
    Bv[nbnd] = comp4(v, (j + 2) % 3);
    Bo[nbnd] = comp4(m.N[tt], j);
    nbnd++;
    return true;
  };
  {
    const int ia = index_of(m.V[t0], a);
    R[n++] = t0;
    CR[nr++] = x;
    CL[nl++] = y;
    // edges (a, x) (opposite y) and (y, a) (opposite x)
    if (!add_bnd(t0, (ia + 2) % 3) || !add_bnd(t0, (ia + 1) % 3)) {
      flag(m, E_CAVITY);
      return o;
    }
  }
  int cur = t0, jx = (index_of(m.V[t0], a));  // crossed edge of cur: opposite a
  int L = y, Rr = x;
  for (int guard = 0; guard < kseg + 8; guard++) {
    const int nb = comp4(m.N[cur], jx);
    if (nb < 0 || !tri_ok(m, nb >> 2)) {
      flag(m, E_TOPOLOGY);
      return o;
    }
    const int u = nb >> 2, ju = nb & 3;  // in u, the crossed edge is opposite ju
    const int4 vu = m.V[u];
    const int z = comp4(vu, ju);
    if (z == m.np || !vert_ok(m, z)) {  // the segment cannot leave the hull
      flag(m, E_TOPOLOGY);
      return o;
    }
    if (n == kseg) return over;
    R[n++] = u;
    // u = (z, ., .): the edge (L, z) and the edge (z, Rr) are its other edges
    const int iL = index_of(vu, L), iR = index_of(vu, Rr);
    if (iL < 0 || iR < 0) {
      flag(m, E_TOPOLOGY);
      return o;
    }
    // edge (z, Rr) is opposite L; edge (L, z) is opposite Rr
    if (z == b) {
      if (!add_bnd(u, iL) || !add_bnd(u, iR)) return over;
      o.status = SS_REGION;
      o.z = b;
      break;
    }
    const double2 Z = m.P[z];
    const int oz = orient2d_sign(pa.x, pa.y, pb.x, pb.y, Z.x, Z.y, false, nullptr);
    if (oz == 0) {  // the segment passes through z: the region ends here
      if (!add_bnd(u, iL) || !add_bnd(u, iR)) return over;
      o.status = SS_REGION;
      o.z = z;
      break;
    }
    if (oz > 0) {  // z left: (L, z) is a boundary edge, the segment leaves through (z, Rr)
      if (!add_bnd(u, iR) || nl == kseg + 2) return over;
      CL[nl++] = z;
      L = z;
      jx = iL;
    } else {  // z right: (z, Rr) is a boundary edge, the segment leaves through (L, z)
      if (!add_bnd(u, iL) || nr == kseg + 2) return over;
      CR[nr++] = z;
      Rr = z;
      jx = iR;
    }
    cur = u;
  }
  if (o.status != SS_REGION) {
    flag(m, E_CAVITY);
    return o;
  }
  // An uncrossed edge between two region triangles (the segment crosses all
  // triangles around a chain vertex of degree 3, so that vertex's chain
  // neighbour appears on both sides of it) is a slit of the pseudo-polygon,
  // not a boundary edge: its partner is found among the new triangles.
  for (int e = 0; e < nbnd; e++)
    if (Bo[e] >= 0 && in_list(R, n, Bo[e] >> 2)) Bo[e] = SLIT;
  o.n = n;
  o.nl = nl;
  o.nr = nr;
  o.nbnd = nbnd;
  return o;
}

// Per pending segment: scan (above) and record the outcome.
__global__ void k_seg_scan(MeshG m, const int2 *seg, const int *pend, int npend, const int *vtri,
                           int *W, int kseg, int4 *info, int4 *info2, int *nover) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < npend; k += gridDim.x * blockDim.x) {
    const int s = pend[k];
    const int2 e = seg[s];
    const SegOut o = seg_scan(m, e.x, e.y, vtri, W + (size_t)k * ws_seg(kseg), kseg);
    if (o.status == SS_OVER) atomicAdd(nover, 1);
    info[k] = make_int4(o.status, o.z, o.n, o.nbnd);
    info2[k] = make_int4(o.nl, o.nr, o.t_edge, o.j_edge);
  }
}

// Claims on the region and its ring; existing edges need none.
__global__ void k_seg_claim(MeshG m, int npend, const int *W, int kseg, const int4 *info,
                            unsigned long long ep) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < npend; k += gridDim.x * blockDim.x) {
    const int4 in = info[k];
    if (in.x != SS_REGION) continue;
    const int *Wk = W + (size_t)k * ws_seg(kseg);
    const int *R = Wk, *Bo = Wk + bo_off(kseg);
    const unsigned long long key = claim_key(ep, k);
    for (int q = 0; q < in.z; q++) atomicMax(&m.claim[2 * (size_t)R[q]], key);
    for (int q = 0; q < in.w; q++)
      if (Bo[q] >= 0) atomicMax(&m.claim[2 * (size_t)(Bo[q] >> 2)], key);
  }
}

__global__ void k_seg_win(MeshG m, int npend, const int *W, int kseg, const int4 *info,
                          unsigned long long ep, int *win) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < npend; k += gridDim.x * blockDim.x) {
    const int4 in = info[k];
    int w = 0;
    if (in.x == SS_REGION) {
      const int *Wk = W + (size_t)k * ws_seg(kseg);
      const int *R = Wk, *Bo = Wk + bo_off(kseg);
      const unsigned long long key = claim_key(ep, k);
      w = 1;
      for (int q = 0; q < in.z && w; q++) w = m.claim[2 * (size_t)R[q]] == key;
      for (int q = 0; q < in.w && w; q++)
        if (Bo[q] >= 0) w = m.claim[2 * (size_t)(Bo[q] >> 2)] == key;
    }
    win[k] = w;
  }
}

// exact-sign incircle of real vertices (filters, double-double, expansion)
__device__ inline int incircle_v(const MeshG &m, int a, int b, int c, int d, double *ws) {
  const double2 A = m.P[a], B = m.P[b], C = m.P[c], D = m.P[d];
  return incircle_sign(A.x, A.y, B.x, B.y, C.x, C.y, D.x, D.y, ws, false, nullptr);
}

// The segment pieces recovered so far: (input segment, from, to).
struct Piece {
  int s, a, b, pad;
};

// Winners: the two pseudo-polygons on either side of the new edge (a, z)
// are triangulated by Anglada's rule -- for a base edge (x, y) with the chain
// c_lo..c_hi on its left, the chain vertex whose circle with x, y holds no
// other (scan: a vertex inside the current best's circle replaces it) gives
// the triangle (x, y, c) and the sub-problems (x, c) and (c, y).  The new
// triangles take the region's slots in ascending order; their links are
// matched edge by edge, inside the region or to the boundary's outer
// triangles (whose back-links are repaired).
__global__ void k_seg_apply(MeshG m, const int2 *seg_cur, const int *pend, int npend, int *W,
                            int kseg, const int4 *info, const int4 *info2, const int *win, double *arena,
                            Piece *pieces, int *npieces, int pcap, int2 *seg_next, int *done) {
  const int lane = blockIdx.x * blockDim.x + threadIdx.x;
  double *ws = arena + (size_t)lane * GCDT_INCIRCLE_WS;
  for (int k = lane; k < npend; k += gridDim.x * blockDim.x) {
    if (!win[k]) continue;
    const int s = pend[k];
    const int a = seg_cur[s].x, b = seg_cur[s].y;
    const int4 in = info[k], in2 = info2[k];
    const int z = in.y, n = in.z, nbnd = in.w, nl = in2.x, nr = in2.y;
    int *R = W + (size_t)k * ws_seg(kseg);
    int *CL = R + kseg, *CR = CL + (kseg + 2);
    int *Bu = CR + (kseg + 2), *Bv = Bu + (kseg + 2), *Bo = Bv + (kseg + 2);
    int *NT = Bo + (kseg + 2);
    int *STK = NT + 3 * (size_t)kseg;
    if (nl + nr != n || n > kseg || nl < 1 || nr < 1) {
      flag(m, E_TOPOLOGY);
      continue;
    }
    int ne = 0;
    bool bad = false;
    for (int side = 0; side < 2 && !bad; side++) {
      // side 0: base (a, z), chain CL in order; side 1: base (z, a), chain CR reversed
      const int cnt = side == 0 ? nl : nr;
      int sp = 0;
      STK[0] = side == 0 ? a : z;
      STK[1] = side == 0 ? z : a;
      STK[2] = 0;
      STK[3] = cnt - 1;
      sp = 1;
      while (sp > 0) {
        sp--;
        const int x = STK[4 * sp], y = STK[4 * sp + 1], lo = STK[4 * sp + 2], hi = STK[4 * sp + 3];
        if (lo > hi) continue;
        int best = lo;
        for (int i = lo + 1; i <= hi; i++) {
          const int cb = side == 0 ? CL[best] : CR[nr - 1 - best];
          const int ci = side == 0 ? CL[i] : CR[nr - 1 - i];
          if (incircle_v(m, x, y, cb, ci, ws) > 0) best = i;
        }
        const int c = side == 0 ? CL[best] : CR[nr - 1 - best];
        if (ne == n || sp + 2 > kseg + 2) {
          bad = true;
          break;
        }
        NT[3 * ne] = x;
        NT[3 * ne + 1] = y;
        NT[3 * ne + 2] = c;
        ne++;
        STK[4 * sp] = x;
        STK[4 * sp + 1] = c;
        STK[4 * sp + 2] = lo;
        STK[4 * sp + 3] = best - 1;
        sp++;
        STK[4 * sp] = c;
        STK[4 * sp + 1] = y;
        STK[4 * sp + 2] = best + 1;
        STK[4 * sp + 3] = hi;
        sp++;
      }
    }
    if (bad || ne != n) {
      flag(m, E_TOPOLOGY);
      continue;
    }
    // region slots in ascending order
    for (int i = 1; i < n; i++) {
      const int x = R[i];
      int j = i - 1;
      while (j >= 0 && R[j] > x) {
        R[j + 1] = R[j];
        j--;
      }
      R[j + 1] = x;
    }
    // links: first find every edge's partner, then write
    bool ok = true;
    for (int q = 0; q < n && ok; q++) {
      const int st = R[q];
      int nn[3];
      for (int j = 0; j < 3 && ok; j++) {
        const int u = NT[3 * q + (j + 1) % 3], v = NT[3 * q + (j + 2) % 3];
        int link = -2;
        for (int q2 = 0; q2 < n && link == -2; q2++) {
          if (q2 == q) continue;
          for (int j2 = 0; j2 < 3; j2++)
            if (NT[3 * q2 + (j2 + 1) % 3] == v && NT[3 * q2 + (j2 + 2) % 3] == u) {
              link = (R[q2] << 2) | j2;
              break;
            }
        }
        if (link == -2)
          for (int e = 0; e < nbnd; e++)
            if (Bu[e] == u && Bv[e] == v && Bo[e] != SLIT) {
              link = Bo[e];
              break;
            }
        if (link == -2) ok = false;
        nn[j] = link;
      }
      if (!ok) break;
      m.V[st] = make_int4(NT[3 * q], NT[3 * q + 1], NT[3 * q + 2], 0);
      m.N[st] = make_int4(nn[0], nn[1], nn[2], 0);
    }
    if (!ok) {
      flag(m, E_TOPOLOGY);
      continue;
    }
    // back-links of the outer triangles
    for (int q = 0; q < n; q++) {
      const int st = R[q];
      for (int j = 0; j < 3; j++) {
        const int u = NT[3 * q + (j + 1) % 3], v = NT[3 * q + (j + 2) % 3];
        for (int e = 0; e < nbnd; e++)
          if (Bu[e] == u && Bv[e] == v && Bo[e] >= 0) {
            const int o = Bo[e];
            int *slot = reinterpret_cast<int *>(&m.N[o >> 2]) + (o & 3);
            *slot = (st << 2) | j;
            break;
          }
      }
    }
    const int pi = atomicAdd(npieces, 1);
    if (pi >= pcap) {
      flag(m, E_CAVITY);
      continue;
    }
    pieces[pi] = Piece{s, a, z, 0};
    seg_next[s] = make_int2(z, b);
    done[s] = z == b;
  }
}

// Segments that are (or begin with) an existing edge: record the piece and
// continue from its end.
__global__ void k_seg_edges(MeshG m, const int2 *seg_cur, const int *pend, int npend,
                            const int4 *info, Piece *pieces, int *npieces, int pcap,
                            int2 *seg_next, int *done) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < npend; k += gridDim.x * blockDim.x) {
    const int4 in = info[k];
    if (in.x != SS_EXISTS && in.x != SS_SPLIT_EDGE) continue;
    const int s = pend[k];
    const int a = seg_cur[s].x, b = seg_cur[s].y, z = in.y;
    const int pi = atomicAdd(npieces, 1);
    if (pi >= pcap) {
      flag(m, E_CAVITY);
      continue;
    }
    pieces[pi] = Piece{s, a, z, 0};
    seg_next[s] = make_int2(z, b);
    done[s] = z == b;
  }
}

__global__ void k_seg_pending_flags(const int *done, int ns, int *flag_out) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += gridDim.x * blockDim.x)
    flag_out[s] = !done[s];
}

// ---------------------------------------------------------------------------
// Export: real triangles in slot order, neighbours as triangle indices
// ---------------------------------------------------------------------------
__global__ void k_real_flags(MeshG m, int nt, int *flag_out) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
    const int4 v = m.V[t];
    flag_out[t] = v.x != m.np && v.y != m.np && v.z != m.np;
  }
}
__global__ void k_export_cdt(MeshG m, int nt, const int *flag_in, const int *newid, int *T,
                             int *Nb) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
    if (!flag_in[t]) continue;
    const int o = newid[t];
    const int4 v = m.V[t], n = m.N[t];
    T[3 * o] = v.x;
    T[3 * o + 1] = v.y;
    T[3 * o + 2] = v.z;
    for (int j = 0; j < 3; j++) {
      const int nb = comp4(n, j);
      int r = -1;
      if (nb >= 0 && tri_ok(m, nb >> 2) && flag_in[nb >> 2]) r = newid[nb >> 2];
      Nb[3 * o + j] = r;
    }
  }
}

// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------
template <class T>
void dev_alloc(T **p, size_t n) {
  CC(hipMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1)));
}

struct Builder {
  hipStream_t st = nullptr;
  std::vector<void *> owned;
  int *scan_bsum = nullptr, *scan_total = nullptr, *h_total = nullptr;
  int bsum_cap = 0;
  unsigned *d_err = nullptr;

  template <class T>
  T *alloc(size_t n) {
    T *p = nullptr;
    dev_alloc(&p, n);
    owned.push_back(p);
    return p;
  }
  Builder() {
    CC(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    CC(hipHostMalloc((void **)&h_total, 4 * sizeof(int), hipHostMallocDefault));
    d_err = alloc<unsigned>(1);
    scan_total = alloc<int>(1);
  }
  ~Builder() {
    for (void *p : owned) (void)hipFree(p);
    if (h_total) (void)hipHostFree(h_total);
    if (st) (void)hipStreamDestroy(st);
  }
  static int grid(long long n, int b = BLK) {
    return (int)std::max<long long>(1, std::min<long long>((n + b - 1) / b, 65536));
  }
  // ordered compaction of flags[0..n): out[r] = src[k] (or k); rank (opt.)
  int compact(const int *flags, int n, const int *src, int *out, int *rank) {
    if (n <= 0) return 0;
    const int nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    if (nb > bsum_cap) {
      bsum_cap = std::max(nb, 2 * bsum_cap);
      scan_bsum = alloc<int>((size_t)bsum_cap);
    }
    k_scan_sums<<<nb, SCAN_CHUNK, 0, st>>>(flags, n, scan_bsum);
    k_scan_top<<<1, SCAN_CHUNK, 0, st>>>(scan_bsum, nb, scan_total);
    k_scan_scatter<<<nb, SCAN_CHUNK, 0, st>>>(flags, n, scan_bsum, src, out, rank);
    CC(hipGetLastError());
    CC(hipMemcpyAsync(h_total, scan_total, sizeof(int), hipMemcpyDeviceToHost, st));
    CC(hipStreamSynchronize(st));
    return h_total[0];
  }
  void check_err(const char *phase) {
    unsigned e = 0;
    CC(hipMemcpyAsync(&e, d_err, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    CC(hipStreamSynchronize(st));
    if (!e) return;
    std::string msg = std::string("GPU CDT construction failed in ") + phase + ":";
    if (e & E_INDEX) msg += " index out of range;";
    if (e & E_WALK) msg += " walk step cap;";
    if (e & E_CAVITY) msg += " region capacity exceeded;";
    if (e & E_TOPOLOGY) msg += " inconsistent triangulation;";
    if (e & E_DUPLICATE) msg += " duplicate input points;";
    throw std::runtime_error(msg);
  }
};

int host_orient(const double *P, int a, int b, int c) {
  return orient2d_sign(P[2 * a], P[2 * a + 1], P[2 * b], P[2 * b + 1], P[2 * c], P[2 * c + 1],
                       false, nullptr);
}

// Vertices of the convex hull, counterclockwise, collinear points left out
// (they join the hull edges later, by insertion).  Points strictly inside
// the octagon of the extreme points in eight directions cannot be hull
// vertices and are dropped first; Andrew's monotone chain with exact
// orientation on the rest.
std::vector<int> convex_hull(const double *P, int np) {
  int ext[8];
  for (int k = 0; k < 8; k++) ext[k] = 0;
  auto key = [&](int i, int k) {
    const double x = P[2 * i], y = P[2 * i + 1];
    switch (k) {
      case 0: return x;
      case 1: return -x;
      case 2: return y;
      case 3: return -y;
      case 4: return x + y;
      case 5: return -(x + y);
      case 6: return x - y;
      default: return y - x;
    }
  };
  for (int i = 1; i < np; i++)
    for (int k = 0; k < 8; k++)
      if (key(i, k) > key(ext[k], k)) ext[k] = i;
  auto lex = [&](int a, int b) {
    return P[2 * a] < P[2 * b] || (P[2 * a] == P[2 * b] && P[2 * a + 1] < P[2 * b + 1]);
  };
  auto chain = [&](std::vector<int> pts) {
    std::sort(pts.begin(), pts.end(), lex);
    pts.erase(std::unique(pts.begin(), pts.end(), [&](int a, int b) {
                return P[2 * a] == P[2 * b] && P[2 * a + 1] == P[2 * b + 1];
              }), pts.end());
    if (pts.size() < 3) return pts;
    std::vector<int> h(2 * pts.size());
    size_t k = 0;
    for (size_t i = 0; i < pts.size(); i++) {
      while (k >= 2 && host_orient(P, h[k - 2], h[k - 1], pts[i]) <= 0) k--;
      h[k++] = pts[i];
    }
    for (size_t i = pts.size() - 1, t = k + 1; i-- > 0;) {
      while (k >= t && host_orient(P, h[k - 2], h[k - 1], pts[i]) <= 0) k--;
      h[k++] = pts[i];
    }
    h.resize(k - 1);
    return h;
  };
  std::vector<int> oct = chain(std::vector<int>(ext, ext + 8));
  std::vector<int> cand;
  if (oct.size() >= 3) {
    for (int i = 0; i < np; i++) {
      bool inside = true;
      for (size_t e = 0; e < oct.size() && inside; e++)
        inside = host_orient(P, oct[e], oct[(e + 1) % oct.size()], i) > 0;
      if (!inside) cand.push_back(i);
    }
  } else {
    cand.resize(np);
    for (int i = 0; i < np; i++) cand[i] = i;
  }
  return chain(cand);
}

}  // namespace cdtgpu

// Builds the CDT of (points, segments) on the device: see the file header.
// Fills T (3 nt), N (3 nt) and S_out (2 ns_out); stats[0..3] = insertion
// rounds, segment passes, segment regions retriangulated, device ms.
void build_cdt_device(const double *points, int np, const int *segs, int ns,
                      std::vector<int> &T_out, std::vector<int> &N_out, std::vector<int> &S_out,
                      double *stats) {
  using namespace cdtgpu;
  if (np < 3) throw std::runtime_error("CDT needs at least 3 points");
  if ((long long)np * 2 >= (1LL << 29)) throw std::runtime_error("too many points");
  for (long long i = 0; i < 2LL * ns; i++)
    if (segs[i] < 0 || segs[i] >= np) throw std::runtime_error("segment vertex index out of range");
  for (int s = 0; s < ns; s++)
    if (segs[2 * s] == segs[2 * s + 1]) throw std::runtime_error("degenerate segment");
  const std::vector<int> hull = convex_hull(points, np);
  if (hull.size() < 3) throw std::runtime_error("all input points are collinear");
  Builder B;
  const int cap = 2 * np - 2;
  const int G = np;
  hipEvent_t e0, e1;
  CC(hipEventCreate(&e0));
  CC(hipEventCreate(&e1));
  CC(hipEventRecord(e0, B.st));
  double2 *dP = B.alloc<double2>(np);
  CC(hipMemcpyAsync(dP, points, sizeof(double2) * np, hipMemcpyHostToDevice, B.st));
  MeshG m;
  m.P = dP;
  m.np = np;
  m.cap = cap;
  m.V = B.alloc<int4>(cap);
  m.N = B.alloc<int4>(cap);
  m.claim = B.alloc<unsigned long long>(2 * (size_t)cap);
  m.err = B.d_err;
  CC(hipMemsetAsync(m.claim, 0, sizeof(unsigned long long) * 2 * cap, B.st));
  CC(hipMemsetAsync(m.err, 0, sizeof(unsigned), B.st));
  // The initial triangulation: the Delaunay triangulation of a seed -- the
  // hull vertices plus a fixed pseudo-random 1/128 of the points (host
  // builder, a few milliseconds at 10^6 points) -- closed by a ghost
  // triangle per hull edge, links matched edge by edge.  Every other point
  // lies inside the hull or on a hull edge, and the rounds start with
  // thousands of triangles instead of a handful.
  std::vector<char> seeded(np, 0);
  for (int i : hull) seeded[i] = 1;
  for (int i = 0; i < np; i++) {
    unsigned x = (unsigned)i * 0x9E3779B1u;
    x ^= x >> 16;
    x *= 0x85EBCA6Bu;
    x ^= x >> 13;
    if ((x & 127u) == 0u) seeded[i] = 1;
  }
  std::vector<int> seed;
  for (int i = 0; i < np; i++)
    if (seeded[i]) seed.push_back(i);
  const int h = (int)seed.size();
  std::vector<int> tv;  // 3 per triangle
  {
    std::vector<double> hp(2 * (size_t)h);
    for (int i = 0; i < h; i++) {
      hp[2 * i] = points[2 * seed[i]];
      hp[2 * i + 1] = points[2 * seed[i] + 1];
    }
    gcdt_cdt *hc = nullptr;
    if (gcdt_build_cdt(hp.data(), h, nullptr, 0, &hc) != 0 || !hc)
      throw std::runtime_error("triangulating the seed points failed");
    for (size_t k = 0; k < hc->T.size(); k++) tv.push_back(seed[hc->T[k]]);
    for (size_t t = 0; t < hc->T.size() / 3; t++)
      for (int j = 0; j < 3; j++)
        if (hc->N[3 * t + j] < 0) {  // hull edge (u, v) -> ghost (v, u, G)
          const int u = seed[hc->T[3 * t + (j + 1) % 3]], v = seed[hc->T[3 * t + (j + 2) % 3]];
          tv.push_back(v);
          tv.push_back(u);
          tv.push_back(G);
        }
    gcdt_cdt_free(hc);
  }
  const int t_init = (int)(tv.size() / 3);
  if (t_init != 2 * h - 2) throw std::runtime_error("seed triangulation has the wrong size");
  {
    std::vector<int> tn(tv.size(), -1);
    std::vector<std::pair<unsigned long long, int>> edges;
    edges.reserve(tv.size());
    for (int t = 0; t < t_init; t++)
      for (int j = 0; j < 3; j++) {
        const unsigned u = (unsigned)tv[3 * t + (j + 1) % 3], v = (unsigned)tv[3 * t + (j + 2) % 3];
        edges.push_back({((unsigned long long)u << 32) | v, (t << 2) | j});
      }
    std::sort(edges.begin(), edges.end());
    for (int t = 0; t < t_init; t++)
      for (int j = 0; j < 3; j++) {
        const unsigned u = (unsigned)tv[3 * t + (j + 1) % 3], v = (unsigned)tv[3 * t + (j + 2) % 3];
        const unsigned long long rk = ((unsigned long long)v << 32) | u;
        auto it = std::lower_bound(edges.begin(), edges.end(), std::make_pair(rk, INT_MIN));
        if (it == edges.end() || it->first != rk) throw std::runtime_error("initial triangulation links");
        tn[3 * t + j] = it->second;
      }
    std::vector<int4> hv(t_init), hn(t_init);
    for (int t = 0; t < t_init; t++) {
      hv[t] = make_int4(tv[3 * t], tv[3 * t + 1], tv[3 * t + 2], 0);
      hn[t] = make_int4(tn[3 * t], tn[3 * t + 1], tn[3 * t + 2], 0);
    }
    CC(hipMemcpyAsync(m.V, hv.data(), sizeof(int4) * t_init, hipMemcpyHostToDevice, B.st));
    CC(hipMemcpyAsync(m.N, hn.data(), sizeof(int4) * t_init, hipMemcpyHostToDevice, B.st));
    CC(hipStreamSynchronize(B.st));
  }
  // uninserted points, all located from triangle 0 first
  std::vector<int> hU;
  hU.reserve(np);
  std::vector<int> hptri(np, 0);
  for (int i = 0; i < np; i++) {
    if (seeded[i]) hptri[i] = -1;
    else hU.push_back(i);
  }
  int nu = (int)hU.size();
  int *ptri = B.alloc<int>(np);
  int *U = B.alloc<int>(np), *U2 = B.alloc<int>(np);
  unsigned long long *dirty = B.alloc<unsigned long long>(cap);
  unsigned long long *pick = B.alloc<unsigned long long>(cap);
  int *flags = B.alloc<int>(std::max(np, cap));
  CC(hipMemcpyAsync(ptri, hptri.data(), sizeof(int) * np, hipMemcpyHostToDevice, B.st));
  if (nu) CC(hipMemcpyAsync(U, hU.data(), sizeof(int) * nu, hipMemcpyHostToDevice, B.st));
  CC(hipMemsetAsync(dirty, 0, sizeof(unsigned long long) * cap, B.st));
  CC(hipMemsetAsync(pick, 0, sizeof(unsigned long long) * cap, B.st));
  unsigned long long ep = 1;
  {
    std::vector<unsigned long long> d0((size_t)t_init, ep);
    CC(hipMemcpyAsync(dirty, d0.data(), sizeof(unsigned long long) * t_init, hipMemcpyHostToDevice,
                      B.st));
    k_relocate<<<Builder::grid(nu), BLK, 0, B.st>>>(m, U, nu, ptri, dirty, ep);
    CC(hipGetLastError());
    CC(hipStreamSynchronize(B.st));
  }
  B.check_err("initial location");
  // ---- phase 1: rounds of insertion ----
  const int CMAX = 1 << 19;
  int *C = B.alloc<int>(np);  // every flagged point is compacted into it
  int *cstat = B.alloc<int>(std::min(np, CMAX) + 1);
  int *cn = B.alloc<int>(std::min(np, CMAX) + 1);
  int *slow = B.alloc<int>(std::min(np, CMAX) + 1);
  int *nslow = B.alloc<int>(1);
  int *win = B.alloc<int>(std::min(np, CMAX) + 1);
  int *active = B.alloc<int>(std::min(np, CMAX) + 1);
  const int ITERS = 4;
  int *rank = B.alloc<int>(std::min(np, CMAX) + 1);
  double *arena = B.alloc<double>((size_t)EXACT_GRID * 64 * GCDT_INCIRCLE_WS);
  int *W = nullptr;
  long long w_cap = 0;
  int tbase = t_init;
  unsigned long long cep = 0;  // claim epoch
  int rounds = 0;
  while (nu > 0) {
    if (++rounds > 100000) throw std::runtime_error("CDT insertion did not finish");
    ep++;
    k_pick<<<Builder::grid(nu), BLK, 0, B.st>>>(m, U, nu, ptri, pick, ep);
    k_cand_flags<<<Builder::grid(nu), BLK, 0, B.st>>>(m, U, nu, ptri, pick, ep, flags);
    CC(hipGetLastError());
    int nc = B.compact(flags, nu, U, C, nullptr);
    nc = std::min(nc, CMAX);
    if (nc == 0) {
      B.check_err("candidate selection");
      throw std::runtime_error("CDT insertion: no candidate");
    }
    if ((long long)nc * WS_INS > w_cap) {
      w_cap = std::max<long long>((long long)nc * WS_INS, 2 * w_cap);
      W = B.alloc<int>((size_t)w_cap);
    }
    CC(hipMemsetAsync(nslow, 0, sizeof(int), B.st));
    k_ins_cavity_fast<<<Builder::grid(nc, 64), 64, 0, B.st>>>(m, C, nc, ptri, W, cstat, cn, slow,
                                                              nslow);
    k_ins_cavity_slow<<<EXACT_GRID, 64, 0, B.st>>>(m, C, ptri, W, cstat, cn, slow, nslow, arena);
    // claim iterations; the round's sticky epoch is above all its claims
    const unsigned long long sticky = cep + ITERS + 1;
    for (int it = 0; it < ITERS; it++) {
      cep++;
      k_ins_claim<<<Builder::grid(nc), BLK, 0, B.st>>>(m, nc, W, cstat, cn, active, it, cep,
                                                        sticky);
      k_ins_win<<<Builder::grid(nc), BLK, 0, B.st>>>(m, nc, W, cn, active, it, cep, sticky, win);
    }
    cep = sticky;
    CC(hipGetLastError());
    const int nw = B.compact(win, nc, nullptr, nullptr, rank);
    B.check_err("insertion cavities");
    if (nw == 0) throw std::runtime_error("CDT insertion made no progress (cavity capacity)");
    if (tbase + 2LL * nw > cap) throw std::runtime_error("CDT triangle slots exhausted");
    k_ins_apply<<<Builder::grid(nc), BLK, 0, B.st>>>(m, C, nc, W, cn, win, rank, tbase, ptri,
                                                      dirty, ep);
    tbase += 2 * nw;
    k_relocate<<<Builder::grid(nu), BLK, 0, B.st>>>(m, U, nu, ptri, dirty, ep);
    k_still_out<<<Builder::grid(nu), BLK, 0, B.st>>>(U, nu, ptri, flags);
    CC(hipGetLastError());
    nu = B.compact(flags, nu, U, U2, nullptr);
    std::swap(U, U2);
    B.check_err("insertion");
  }
  if (tbase != cap) throw std::runtime_error("CDT insertion left unused slots (duplicate points?)");
  // ---- phase 2: constraint recovery ----
  int2 *seg_cur = B.alloc<int2>(ns), *seg_next = B.alloc<int2>(ns);
  int *done = B.alloc<int>(ns);
  int *pend = B.alloc<int>(ns), *pend2 = B.alloc<int>(ns);
  int *vtri = B.alloc<int>(np);
  const int PCAP = 4 * ns + 16;
  Piece *pieces = B.alloc<Piece>(PCAP);
  int *npieces = B.alloc<int>(1);
  int4 *info = nullptr, *info2 = nullptr;
  int *swin = nullptr, *W2 = nullptr;
  int seg_passes = 0, npieces_total = 0;
  if (ns > 0) {
    CC(hipMemcpyAsync(seg_cur, segs, sizeof(int2) * ns, hipMemcpyHostToDevice, B.st));
    CC(hipMemsetAsync(done, 0, sizeof(int) * ns, B.st));
    CC(hipMemsetAsync(npieces, 0, sizeof(int), B.st));
    std::vector<int> hp(ns);
    for (int s = 0; s < ns; s++) hp[s] = s;
    CC(hipMemcpyAsync(pend, hp.data(), sizeof(int) * ns, hipMemcpyHostToDevice, B.st));
    const int SBATCH = std::min(ns, 1 << 15);
    info = B.alloc<int4>(SBATCH);
    info2 = B.alloc<int4>(SBATCH);
    swin = B.alloc<int>(SBATCH);
    int kseg = KSEG0;
    size_t w2_ints = (size_t)SBATCH * ws_seg(kseg);
    W2 = B.alloc<int>(w2_ints);
    int *nover = B.alloc<int>(1);
    int npend = ns;
    int last_pieces = -1;
    while (npend > 0) {
      if (++seg_passes > 1000000) throw std::runtime_error("segment recovery did not finish");
      const int batch = (int)std::min<size_t>(SBATCH, w2_ints / ws_seg(kseg));
      const int nb = std::min(npend, batch);
      CC(hipMemsetAsync(nover, 0, sizeof(int), B.st));
      CC(hipMemsetAsync(vtri, 0x7f, sizeof(int) * np, B.st));
      k_vtri<<<Builder::grid(cap), BLK, 0, B.st>>>(m, cap, vtri);
      k_seg_scan<<<Builder::grid(nb, 64), 64, 0, B.st>>>(m, seg_cur, pend, nb, vtri, W2, kseg,
                                                          info, info2, nover);
      CC(hipMemcpyAsync(seg_next, seg_cur, sizeof(int2) * ns, hipMemcpyDeviceToDevice, B.st));
      k_seg_edges<<<Builder::grid(nb), BLK, 0, B.st>>>(m, seg_cur, pend, nb, info, pieces,
                                                       npieces, PCAP, seg_next, done);
      cep++;
      k_seg_claim<<<Builder::grid(nb), BLK, 0, B.st>>>(m, nb, W2, kseg, info, cep);
      k_seg_win<<<Builder::grid(nb), BLK, 0, B.st>>>(m, nb, W2, kseg, info, cep, swin);
      k_seg_apply<<<EXACT_GRID, 64, 0, B.st>>>(m, seg_cur, pend, nb, W2, kseg, info, info2, swin,
                                               arena, pieces, npieces, PCAP, seg_next, done);
      CC(hipGetLastError());
      CC(hipMemcpyAsync(seg_cur, seg_next, sizeof(int2) * ns, hipMemcpyDeviceToDevice, B.st));
      k_seg_pending_flags<<<Builder::grid(ns), BLK, 0, B.st>>>(done, ns, flags);
      CC(hipGetLastError());
      npend = B.compact(flags, ns, nullptr, pend2, nullptr);
      std::swap(pend, pend2);
      int hpieces = 0, hover = 0;
      CC(hipMemcpyAsync(&hpieces, npieces, sizeof(int), hipMemcpyDeviceToHost, B.st));
      CC(hipMemcpyAsync(&hover, nover, sizeof(int), hipMemcpyDeviceToHost, B.st));
      CC(hipStreamSynchronize(B.st));
      B.check_err("segment recovery");
      if (hpieces == last_pieces) {
        // no progress: only regions over capacity can be left
        if (hover == 0) throw std::runtime_error("segment recovery made no progress");
        if (kseg >= cap) throw std::runtime_error("segment region larger than the triangulation");
        kseg = (int)std::min<long long>(4LL * kseg, cap);
        if (ws_seg(kseg) > w2_ints) {
          w2_ints = ws_seg(kseg);
          W2 = B.alloc<int>(w2_ints);
        }
        continue;
      }
      if (hpieces > PCAP) throw std::runtime_error("too many segment pieces");
      last_pieces = hpieces;
    }
    npieces_total = last_pieces;
  }
  // ---- export ----
  k_real_flags<<<Builder::grid(cap), BLK, 0, B.st>>>(m, cap, flags);
  int *newid = B.alloc<int>(cap);
  const int nreal = B.compact(flags, cap, nullptr, nullptr, newid);
  int *dT = B.alloc<int>(3 * (size_t)nreal), *dN = B.alloc<int>(3 * (size_t)nreal);
  k_export_cdt<<<Builder::grid(cap), BLK, 0, B.st>>>(m, cap, flags, newid, dT, dN);
  CC(hipGetLastError());
  T_out.resize(3 * (size_t)nreal);
  N_out.resize(3 * (size_t)nreal);
  CC(hipMemcpyAsync(T_out.data(), dT, sizeof(int) * 3 * nreal, hipMemcpyDeviceToHost, B.st));
  CC(hipMemcpyAsync(N_out.data(), dN, sizeof(int) * 3 * nreal, hipMemcpyDeviceToHost, B.st));
  int hpieces = 0;
  CC(hipMemcpyAsync(&hpieces, npieces, sizeof(int), hipMemcpyDeviceToHost, B.st));
  CC(hipStreamSynchronize(B.st));
  std::vector<Piece> hp((size_t)hpieces);
  if (hpieces)
    CC(hipMemcpyAsync(hp.data(), pieces, sizeof(Piece) * hpieces, hipMemcpyDeviceToHost, B.st));
  CC(hipEventRecord(e1, B.st));
  CC(hipStreamSynchronize(B.st));
  float ms = 0.f;
  CC(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  B.check_err("export");
  // segments: each input segment's pieces in order from its first endpoint,
  // then the hull edges that are not segments
  std::stable_sort(hp.begin(), hp.end(), [](const Piece &x, const Piece &y) { return x.s < y.s; });
  S_out.clear();
  std::vector<unsigned long long> keys;
  keys.reserve(hp.size());
  for (size_t lo = 0; lo < hp.size();) {
    size_t hi = lo;
    while (hi < hp.size() && hp[hi].s == hp[lo].s) hi++;
    int cur = segs[2 * hp[lo].s];
    for (size_t c = lo; c < hi; c++) {
      size_t f = lo;
      while (f < hi && hp[f].a != cur) f++;
      if (f == hi) throw std::runtime_error("segment pieces do not chain");
      S_out.push_back(hp[f].a);
      S_out.push_back(hp[f].b);
      const unsigned lo_v = (unsigned)std::min(hp[f].a, hp[f].b), hi_v = (unsigned)std::max(hp[f].a, hp[f].b);
      keys.push_back(((unsigned long long)lo_v << 32) | hi_v);
      cur = hp[f].b;
    }
    lo = hi;
  }
  std::sort(keys.begin(), keys.end());
  for (int t = 0; t < nreal; t++)
    for (int j = 0; j < 3; j++) {
      if (N_out[3 * (size_t)t + j] != -1) continue;
      const int u = T_out[3 * (size_t)t + (j + 1) % 3], v = T_out[3 * (size_t)t + (j + 2) % 3];
      const unsigned long long k =
          ((unsigned long long)(unsigned)std::min(u, v) << 32) | (unsigned)std::max(u, v);
      if (std::binary_search(keys.begin(), keys.end(), k)) continue;
      S_out.push_back(u);
      S_out.push_back(v);
    }
  if (stats) {
    stats[0] = rounds;
    stats[1] = seg_passes;
    stats[2] = npieces_total;
    stats[3] = ms;
  }
}

}  // namespace gcdt
