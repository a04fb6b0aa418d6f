// Device-side invariant checks of a refined constrained Delaunay
// triangulation, with exact predicates: the reference's end-of-run host
// checks (checkNeighbors main.cpp:79-113, checkIncircle main.cpp:153-242,
// checkQuality main.cpp:280-328, the index ranges of checkResult
// main.cpp:336-362) plus orientation, the boundary / subsegment edge rules,
// segment coverage and edge multiplicity.  Counts per category
// (gcdt_check_mesh in include/gpu_cdt_refine.h):
//   0 index range          1 neighbour symmetry     2 orientation (<= 0)
//   3 boundary edge that is not a subsegment        4 Delaunay: exact
//     incircle > 0 across a non-subsegment edge     5 subsegment that is not
//     a mesh edge          6 input segment not covered by a chain of output
//     subsegments          7 input point moved      8 quality: smallest angle
//     below theta, outside the Miller-Pav-Walkington exemption
//   9 edge in >= 3 triangles
//
// Layout: one thread per triangle / subsegment / edge / input segment;
// subsegments and edges go into open-addressing hash tables of 64-bit keys
// (lo << 32 | hi vertex); the subsegment graph is held in CSR form (degree
// count, exclusive scan, fill) so that one thread per input segment walks
// its chain of subsegments.  Signs are decided by the FP64 filter and the
// double-double stage in the triangle kernel; the few that stay undecided are
// listed and re-evaluated with exact expansions by a small grid holding a
// per-lane arena.
//
// Every kernel is enqueued without a host round trip: a range violation is
// read on the device (the later kernels return when cnt[V_RANGE] != 0), the
// exact grid reads its work count from its list, and the results land in a
// device array.  All scratch comes from one region carved by ck_layout, so
// the same kernels serve the allocating entry points and the one that works
// in a static device array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include "gcdt_internal.h"
#include "gpu_cdt_refine.h"
#include "predicates.h"

namespace gcdt {

namespace {

enum { V_RANGE = 0, V_SYM = 1, V_ORIENT = 2, V_BOUNDARY = 3, V_DELAUNAY = 4, V_SUBSEG = 5,
       V_COVER = 6, V_MOVED = 7, V_QUALITY = 8, V_MANIFOLD = 9, V_N = 10 };
// counter words after the ten categories
enum { C_EXEMPT = V_N, C_USED, C_MINANG, C_WORDS = 16 };
constexpr unsigned long long EMPTY_KEY = ~0ull;
constexpr int SCAN_CHUNK = 1024;

#define CK(x)                                                                   \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess)                                                       \
      throw std::runtime_error(std::string("HIP error in mesh check: ") +       \
                               hipGetErrorString(e_));                          \
  } while (0)

struct View {
  const double2 *P;
  int np;
  const int *T, *N;
  int nt;
  const int2 *S;
  int ns;
  const double2 *P_in;
  int np_in;
  const int2 *S_in;  // input segments (categories 6 and 8); null: not checked
  int ns_in;
};

// Scratch of one check, carved from a single region (ck_layout).
struct Scratch {
  unsigned long long *cnt;  // C_WORDS counters
  unsigned long long *skeys, *ekeys;
  unsigned long long scap, ecap;
  int *ecnt;
  int *defer;  // [0] count, then edge ids (3t + j) left to the exact stage
  int *deg;    // subsegment degree per vertex, then (in place) CSR offsets
  int *bsum;   // scan block sums
  int *adj;    // CSR: incident subsegment ids
  int *fill;   // CSR fill cursors
  int *onseg;  // input segment whose chain passes through the vertex, or -1
  double *ws;  // exact-stage arena, GCDT_INCIRCLE_WS doubles per lane
  int lanes;   // lanes of the exact-stage grid
};

__device__ __forceinline__ unsigned long long edge_key64(int a, int b) {
  unsigned lo = (unsigned)min(a, b), hi = (unsigned)max(a, b);
  return ((unsigned long long)lo << 32) | hi;
}
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}
// slot of key (inserted if absent); cap is a power of two above the key count
__device__ __forceinline__ unsigned long long hash_insert(unsigned long long *keys,
                                                          unsigned long long cap,
                                                          unsigned long long k) {
  unsigned long long h = mix64(k) & (cap - 1);
  for (;;) {
    unsigned long long prev = atomicCAS(&keys[h], EMPTY_KEY, k);
    if (prev == EMPTY_KEY || prev == k) return h;
    h = (h + 1) & (cap - 1);
  }
}
__device__ __forceinline__ bool hash_has(const unsigned long long *keys, unsigned long long cap,
                                         unsigned long long k) {
  unsigned long long h = mix64(k) & (cap - 1);
  for (;;) {
    unsigned long long v = keys[h];
    if (v == k) return true;
    if (v == EMPTY_KEY) return false;
    h = (h + 1) & (cap - 1);
  }
}
// the checks after the range pass gather through the mesh's indices: they
// run only when every index is in range
__device__ __forceinline__ bool range_failed(const unsigned long long *cnt) {
  return *(volatile const unsigned long long *)&cnt[V_RANGE] != 0ull;
}

__global__ void k_ck_fill64(unsigned long long *a, unsigned long long n, unsigned long long v) {
  for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * blockDim.x)
    a[i] = v;
}
__global__ void k_ck_fill32(int *a, unsigned long long n, int v) {
  for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * blockDim.x)
    a[i] = v;
}

__global__ void k_ck_range(View v, unsigned long long *cnt) {
  unsigned long long bad = 0;
  const long long nT = 3LL * v.nt, nS = 2LL * v.ns, nI = v.S_in ? 2LL * v.ns_in : 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nT + nS + nI;
       i += (long long)gridDim.x * blockDim.x) {
    if (i < nT) {
      int a = v.T[i], u = v.N[i];
      bad += (a < 0 || a >= v.np) + (u < -1 || u >= v.nt);
    } else if (i < nT + nS) {
      long long k = i - nT;
      int2 s = v.S[k >> 1];
      int a = (k & 1) ? s.y : s.x;
      bad += (a < 0 || a >= v.np);
    } else {
      long long k = i - nT - nS;
      int2 s = v.S_in[k >> 1];
      int a = (k & 1) ? s.y : s.x;
      bad += (a < 0 || a >= v.np);
    }
  }
  if (bad) atomicAdd(&cnt[V_RANGE], bad);
}

__global__ void k_ck_moved(View v, unsigned long long *cnt) {
  const int n = min(v.np, v.np_in);
  unsigned long long bad = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double2 a = v.P[i], b = v.P_in[i];
    bad += (__double_as_longlong(a.x) != __double_as_longlong(b.x)) ||
           (__double_as_longlong(a.y) != __double_as_longlong(b.y));
  }
  if (bad) atomicAdd(&cnt[V_MOVED], bad);
}

__global__ void k_ck_seg_insert(View v, Scratch s) {
  if (range_failed(s.cnt)) return;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < v.ns; q += gridDim.x * blockDim.x) {
    int2 e = v.S[q];
    hash_insert(s.skeys, s.scap, edge_key64(e.x, e.y));
  }
}

__global__ void k_ck_edge_insert(View v, Scratch s) {
  if (range_failed(s.cnt)) return;
  const long long n = 3LL * v.nt;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long t = i / 3;
    const int j = (int)(i % 3);
    int a = v.T[3 * t + (j + 1) % 3], b = v.T[3 * t + (j + 2) % 3];
    atomicAdd(&s.ecnt[hash_insert(s.ekeys, s.ecap, edge_key64(a, b))], 1);
  }
}

__global__ void k_ck_edge_count(Scratch s) {
  if (range_failed(s.cnt)) return;
  unsigned long long bad = 0;
  for (unsigned long long h = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; h < s.ecap;
       h += (unsigned long long)gridDim.x * blockDim.x)
    bad += s.ekeys[h] != EMPTY_KEY && s.ecnt[h] >= 3;
  if (bad) atomicAdd(&s.cnt[V_MANIFOLD], bad);
}

__global__ void k_ck_subseg(View v, Scratch s) {
  if (range_failed(s.cnt)) return;
  unsigned long long bad = 0;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < v.ns; q += gridDim.x * blockDim.x) {
    int2 e = v.S[q];
    bad += !hash_has(s.ekeys, s.ecap, edge_key64(e.x, e.y));
  }
  if (bad) atomicAdd(&s.cnt[V_SUBSEG], bad);
}

// Per triangle: orientation, then per edge the boundary rule, neighbour
// symmetry and (once per edge, from the lower index) the Delaunay test.
__global__ void k_ck_tri(View v, Scratch s) {
  if (range_failed(s.cnt)) return;
  unsigned long long bo = 0, bb = 0, bs = 0, bd = 0;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < v.nt; t += gridDim.x * blockDim.x) {
    const int t0 = v.T[3 * t], t1 = v.T[3 * t + 1], t2 = v.T[3 * t + 2];
    const double2 p0 = v.P[t0], p1 = v.P[t1], p2 = v.P[t2];
    bo += orient2d_sign(p0.x, p0.y, p1.x, p1.y, p2.x, p2.y, false, nullptr) <= 0;
    const int tv[3] = {t0, t1, t2};
    for (int j = 0; j < 3; j++) {
      const int a = tv[(j + 1) % 3], b = tv[(j + 2) % 3];
      const int u = v.N[3 * t + j];
      const bool is_seg = hash_has(s.skeys, s.scap, edge_key64(a, b));
      if (u < 0) {
        bb += !is_seg;
        continue;
      }
      int found = -1;
      for (int k = 0; k < 3; k++)
        if (v.T[3 * u + (k + 1) % 3] == b && v.T[3 * u + (k + 2) % 3] == a) found = k;
      if (found < 0 || v.N[3 * u + found] != t) {
        bs++;
        continue;
      }
      if (is_seg || t >= u) continue;
      const double2 d = v.P[v.T[3 * u + found]];
      int sg = incircle_filtered(p0.x, p0.y, p1.x, p1.y, p2.x, p2.y, d.x, d.y, false, nullptr);
      if (sg == PRED_UNSURE) sg = incircle_dd_sign(p0.x, p0.y, p1.x, p1.y, p2.x, p2.y, d.x, d.y);
      if (sg == PRED_UNSURE || sg == 0) {
        // undecided (or a zero the filters cannot certify): exact stage
        int pos = atomicAdd(&s.defer[0], 1);
        s.defer[1 + pos] = 3 * t + j;
        continue;
      }
      bd += sg > 0;
    }
  }
  if (bo) atomicAdd(&s.cnt[V_ORIENT], bo);
  if (bb) atomicAdd(&s.cnt[V_BOUNDARY], bb);
  if (bs) atomicAdd(&s.cnt[V_SYM], bs);
  if (bd) atomicAdd(&s.cnt[V_DELAUNAY], bd);
}

// Exact incircle for the listed edges (the list's length is read here, so
// the grid is fixed at enqueue time).
__global__ void k_ck_exact(View v, Scratch s) {
  if (range_failed(s.cnt)) return;
  const int n = s.defer[0];
  const int lane = blockIdx.x * blockDim.x + threadIdx.x;
  double *w = s.ws + (size_t)lane * GCDT_INCIRCLE_WS;
  unsigned long long bd = 0;
  for (int e = lane; e < n; e += gridDim.x * blockDim.x) {
    const int te = s.defer[1 + e], t = te / 3, j = te % 3;
    const int t0 = v.T[3 * t], t1 = v.T[3 * t + 1], t2 = v.T[3 * t + 2];
    const int a = v.T[3 * t + (j + 1) % 3], b = v.T[3 * t + (j + 2) % 3];
    const int u = v.N[3 * t + j];
    int d = -1;
    for (int k = 0; k < 3; k++)
      if (v.T[3 * u + (k + 1) % 3] == b && v.T[3 * u + (k + 2) % 3] == a) d = v.T[3 * u + k];
    const double2 p0 = v.P[t0], p1 = v.P[t1], p2 = v.P[t2], pd = v.P[d];
    bd += incircle_exact(p0.x, p0.y, p1.x, p1.y, p2.x, p2.y, pd.x, pd.y, w) > 0.0;
  }
  if (bd) atomicAdd(&s.cnt[V_DELAUNAY], bd);
}

// ---- segment coverage: CSR of the subsegment graph, one walk per input
// segment (the CPU checker's rule, check_oracle.c) ----
__global__ void k_ck_deg(View v, Scratch s) {
  if (range_failed(s.cnt)) return;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < v.ns; q += gridDim.x * blockDim.x) {
    int2 e = v.S[q];
    atomicAdd(&s.deg[e.x], 1);
    atomicAdd(&s.deg[e.y], 1);
  }
}
// Exclusive scan of deg[0..n) in place, in chunks of SCAN_CHUNK: chunk sums,
// a one-block scan of the sums, then each chunk's own scan plus its offset.
__device__ __forceinline__ int block_exclusive_scan(int x, int *tmp, int *total) {
  // blockDim.x == SCAN_CHUNK; Hillis-Steele over LDS
  const int t = threadIdx.x;
  tmp[t] = x;
  __syncthreads();
  for (int off = 1; off < SCAN_CHUNK; off <<= 1) {
    int y = t >= off ? tmp[t - off] : 0;
    __syncthreads();
    tmp[t] += y;
    __syncthreads();
  }
  int incl = tmp[t];
  if (total) *total = tmp[SCAN_CHUNK - 1];
  __syncthreads();
  return incl - x;
}
__global__ __launch_bounds__(SCAN_CHUNK) void k_ck_scan_sums(Scratch s, int n) {
  if (range_failed(s.cnt)) return;
  __shared__ int tmp[SCAN_CHUNK];
  const long long i = (long long)blockIdx.x * SCAN_CHUNK + threadIdx.x;
  int tot = 0;
  block_exclusive_scan(i < n ? s.deg[i] : 0, tmp, &tot);
  if (threadIdx.x == 0) s.bsum[blockIdx.x] = tot;
}
__global__ __launch_bounds__(SCAN_CHUNK) void k_ck_scan_top(Scratch s, int nb) {
  if (range_failed(s.cnt)) return;
  __shared__ int tmp[SCAN_CHUNK];
  int carry = 0;
  for (int base = 0; base < nb; base += SCAN_CHUNK) {
    const int i = base + threadIdx.x;
    int tot = 0;
    int ex = block_exclusive_scan(i < nb ? s.bsum[i] : 0, tmp, &tot);
    if (i < nb) s.bsum[i] = ex + carry;
    carry += tot;
  }
}
__global__ __launch_bounds__(SCAN_CHUNK) void k_ck_scan_apply(Scratch s, int n) {
  if (range_failed(s.cnt)) return;
  __shared__ int tmp[SCAN_CHUNK];
  const long long i = (long long)blockIdx.x * SCAN_CHUNK + threadIdx.x;
  int ex = block_exclusive_scan(i < n ? s.deg[i] : 0, tmp, nullptr);
  if (i < n) s.deg[i] = ex + s.bsum[blockIdx.x];
}
__global__ void k_ck_adj(View v, Scratch s) {
  if (range_failed(s.cnt)) return;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < v.ns; q += gridDim.x * blockDim.x) {
    int2 e = v.S[q];
    s.adj[s.deg[e.x] + atomicAdd(&s.fill[e.x], 1)] = q;
    s.adj[s.deg[e.y] + atomicAdd(&s.fill[e.y], 1)] = q;
  }
}
// Walk input segment q from its origin to its destination along subsegments
// whose vertices lie on it (distance <= tol * length), always to the next
// vertex along the segment; ties go to the lowest subsegment index, as the
// CPU checker's ascending adjacency order does.  A vertex passed on the way
// records the highest segment index that passed it.
__global__ void k_ck_walk(View v, Scratch s, double tol) {
  if (range_failed(s.cnt)) return;
  unsigned long long bad = 0, used = 0;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < v.ns_in; q += gridDim.x * blockDim.x) {
    const int A = v.S_in[q].x, B = v.S_in[q].y;
    const double2 pa = v.P[A], pb = v.P[B];
    const double dx = pb.x - pa.x, dy = pb.y - pa.y;
    const double len2 = dx * dx + dy * dy, len = sqrt(len2);
    int cur = A;
    double tcur = 0.0;
    bool ok = false;
    for (long long steps = 0; steps <= v.np; steps++) {
      if (cur == B) {
        ok = true;
        break;
      }
      int best = -1, bseg = INT_MAX;
      double tb = 2.0;
      for (int e = s.deg[cur]; e < s.deg[cur + 1]; e++) {
        const int sq = s.adj[e];
        const int2 se = v.S[sq];
        const int w = se.x == cur ? se.y : se.x;
        const double2 pw = v.P[w];
        const double tw = ((pw.x - pa.x) * dx + (pw.y - pa.y) * dy) / len2;
        const double dist = fabs((pw.x - pa.x) * dy - (pw.y - pa.y) * dx) / len;
        if (dist <= tol * len && tw > tcur && (tw < tb || (tw == tb && sq < bseg))) {
          tb = tw;
          best = w;
          bseg = sq;
        }
      }
      if (best < 0) break;
      if (best != B) atomicMax(&s.onseg[best], q);
      cur = best;
      tcur = tb;
      used++;
    }
    bad += !ok;
  }
  if (bad) atomicAdd(&s.cnt[V_COVER], bad);
  if (used) atomicAdd(&s.cnt[C_USED], used);
}

// Quality: cos^2 of the angle opposite the shortest edge against cos^2(theta)
// (with the CPU checker's 1e-9 relative slack, folded into ga), the
// Miller-Pav-Walkington exemption (triangle.cpp:7331-7378): the shortest
// edge joins two vertices inside different input segments that share an
// endpoint, equidistant from it within 0.1%, and is not a subsegment.
__global__ void k_ck_quality(View v, Scratch s, double ga) {
  if (range_failed(s.cnt)) return;
  unsigned long long bad = 0, ex = 0;
  double minang = 180.0;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < v.nt; t += gridDim.x * blockDim.x) {
    const int tv[3] = {v.T[3 * t], v.T[3 * t + 1], v.T[3 * t + 2]};
    double x[3], y[3], l[3];
    for (int k = 0; k < 3; k++) {
      const double2 p = v.P[tv[k]];
      x[k] = p.x;
      y[k] = p.y;
    }
    for (int k = 0; k < 3; k++) {
      const int a = (k + 1) % 3, b = (k + 2) % 3;
      l[k] = (x[a] - x[b]) * (x[a] - x[b]) + (y[a] - y[b]) * (y[a] - y[b]);
    }
    int e = 0;
    if (l[1] < l[e]) e = 1;
    if (l[2] < l[e]) e = 2;
    const int a = (e + 1) % 3, b = (e + 2) % 3;
    const double dot = (x[a] - x[e]) * (x[b] - x[e]) + (y[a] - y[e]) * (y[b] - y[e]);
    const double cos2 = dot * dot / (l[a] * l[b]);
    const double ang = acos(sqrt(cos2 > 1.0 ? 1.0 : cos2)) * (180.0 / 3.14159265358979323846);
    if (cos2 > ga) {
      const int va = tv[a], vb = tv[b];
      bool exempt = false;
      if (va >= v.np_in && vb >= v.np_in && s.onseg[va] >= 0 && s.onseg[vb] >= 0 &&
          !hash_has(s.skeys, s.scap, edge_key64(va, vb))) {
        const int2 s1 = v.S_in[s.onseg[va]], s2 = v.S_in[s.onseg[vb]];
        int common = -1;
        if (s1.x == s2.x || s1.x == s2.y) common = s1.x;
        else if (s1.y == s2.x || s1.y == s2.y) common = s1.y;
        if (common >= 0) {
          const double2 J = v.P[common], PA = v.P[va], PB = v.P[vb];
          const double d1s = (PA.x - J.x) * (PA.x - J.x) + (PA.y - J.y) * (PA.y - J.y);
          const double d2s = (PB.x - J.x) * (PB.x - J.x) + (PB.y - J.y) * (PB.y - J.y);
          exempt = d1s < 1.0011 * d2s && d1s > 0.9989 * d2s;
        }
      }
      if (exempt) {
        ex++;
      } else {
        bad++;
        minang = fmin(minang, ang);
      }
    } else {
      minang = fmin(minang, ang);
    }
  }
  if (bad) atomicAdd(&s.cnt[V_QUALITY], bad);
  if (ex) atomicAdd(&s.cnt[C_EXEMPT], ex);
  // non-negative doubles order like their bit patterns
  atomicMin(&s.cnt[C_MINANG], (unsigned long long)__double_as_longlong(minang));
}

// Results: out[0..9] the categories, out[10] signs left to the exact stage,
// out[11] MPW-exempt triangles, out[12] the smallest angle (degrees, as the
// bits of a double) over the triangles that are not exempt.
__global__ void k_ck_finish(View v, Scratch s, long long *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned long long c[C_WORDS];
  for (int i = 0; i < C_WORDS; i++) c[i] = s.cnt[i];
  if (c[V_RANGE] == 0 && v.S_in) {
    const unsigned long long used = c[C_USED], ns = (unsigned long long)v.ns;
    c[V_COVER] += used > ns ? used - ns : ns - used;
  }
  if (v.np < v.np_in) c[V_MOVED] += (unsigned long long)(v.np_in - v.np);
  for (int i = 0; i < V_N; i++) out[i] = (long long)c[i];
  out[10] = c[V_RANGE] == 0 ? (long long)s.defer[0] : 0;
  out[11] = (long long)c[C_EXEMPT];
  out[12] = (long long)c[C_MINANG];
}

unsigned long long pow2_above(unsigned long long x) {
  unsigned long long p = 1024;
  while (p < x) p <<= 1;
  return p;
}
int grid_for(long long n, int block = 256, int max_grid = 65536) {
  long long g = (n + block - 1) / block;
  return (int)std::max<long long>(1, std::min<long long>(g, max_grid));
}

// Carves the scratch of a check out of base (nullptr: only sizes it) and
// returns the bytes it needs; every piece is 256-byte aligned.
size_t ck_layout(int np, int nt, int ns, int lanes, char *base, Scratch *out) {
  Scratch s{};
  s.scap = pow2_above(2ull * (unsigned long long)ns + 16);
  s.ecap = pow2_above(6ull * (unsigned long long)nt + 16);
  s.lanes = lanes;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *p = base ? base + off : nullptr;
    off += (bytes + 255) & ~(size_t)255;
    return p;
  };
  const size_t nb = ((size_t)np + 1 + SCAN_CHUNK - 1) / SCAN_CHUNK;
  s.cnt = (unsigned long long *)take(sizeof(unsigned long long) * C_WORDS);
  s.skeys = (unsigned long long *)take(sizeof(unsigned long long) * s.scap);
  s.ekeys = (unsigned long long *)take(sizeof(unsigned long long) * s.ecap);
  s.ecnt = (int *)take(sizeof(int) * s.ecap);
  s.defer = (int *)take(sizeof(int) * (1 + 3 * (size_t)nt));
  s.deg = (int *)take(sizeof(int) * ((size_t)np + 1));
  s.bsum = (int *)take(sizeof(int) * nb);
  s.adj = (int *)take(sizeof(int) * std::max<size_t>(1, 2 * (size_t)ns));
  s.fill = (int *)take(sizeof(int) * std::max<size_t>(1, (size_t)np));
  s.onseg = (int *)take(sizeof(int) * std::max<size_t>(1, (size_t)np));
  s.ws = (double *)take(sizeof(double) * (size_t)lanes * GCDT_INCIRCLE_WS);
  if (out) *out = s;
  return off;
}

// Enqueues the whole check on `stream`; results go to the device array out
// (13 long longs, see k_ck_finish).  No host synchronisation, no allocation,
// no memset node (fills are kernels).
void ck_enqueue(const View &v, const Scratch &s, double theta_deg, double seg_tol,
                long long *out, hipStream_t stream) {
  const bool segs_in = v.S_in != nullptr;
  k_ck_fill64<<<1, 64, 0, stream>>>(s.cnt, C_WORDS, 0ull);
  k_ck_fill64<<<1, 64, 0, stream>>>(s.cnt + C_MINANG, 1,
                                    (unsigned long long)__builtin_bit_cast(long long, 180.0));
  k_ck_fill64<<<grid_for((long long)s.scap), 256, 0, stream>>>(s.skeys, s.scap, EMPTY_KEY);
  k_ck_fill64<<<grid_for((long long)s.ecap), 256, 0, stream>>>(s.ekeys, s.ecap, EMPTY_KEY);
  k_ck_fill32<<<grid_for((long long)s.ecap), 256, 0, stream>>>(s.ecnt, s.ecap, 0);
  k_ck_fill32<<<1, 64, 0, stream>>>(s.defer, 1, 0);
  k_ck_range<<<grid_for(3LL * v.nt + 2LL * v.ns + (segs_in ? 2LL * v.ns_in : 0)), 256, 0,
               stream>>>(v, s.cnt);
  if (v.P_in && v.np_in > 0)
    k_ck_moved<<<grid_for(std::min(v.np, v.np_in)), 256, 0, stream>>>(v, s.cnt);
  k_ck_seg_insert<<<grid_for(v.ns), 256, 0, stream>>>(v, s);
  k_ck_edge_insert<<<grid_for(3LL * v.nt), 256, 0, stream>>>(v, s);
  k_ck_edge_count<<<grid_for((long long)s.ecap), 256, 0, stream>>>(s);
  k_ck_subseg<<<grid_for(v.ns), 256, 0, stream>>>(v, s);
  k_ck_tri<<<grid_for(v.nt), 256, 0, stream>>>(v, s);
  k_ck_exact<<<std::max(1, s.lanes / 64), 64, 0, stream>>>(v, s);
  if (segs_in) {
    const int n = v.np + 1;
    const int nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    k_ck_fill32<<<grid_for(n), 256, 0, stream>>>(s.deg, (unsigned long long)n, 0);
    k_ck_fill32<<<grid_for(std::max(1, v.np)), 256, 0, stream>>>(s.fill,
                                                                 (unsigned long long)v.np, 0);
    k_ck_fill32<<<grid_for(std::max(1, v.np)), 256, 0, stream>>>(s.onseg,
                                                                 (unsigned long long)v.np, -1);
    k_ck_deg<<<grid_for(v.ns), 256, 0, stream>>>(v, s);
    k_ck_scan_sums<<<nb, SCAN_CHUNK, 0, stream>>>(s, n);
    k_ck_scan_top<<<1, SCAN_CHUNK, 0, stream>>>(s, nb);
    k_ck_scan_apply<<<nb, SCAN_CHUNK, 0, stream>>>(s, n);
    k_ck_adj<<<grid_for(v.ns), 256, 0, stream>>>(v, s);
    k_ck_walk<<<grid_for(v.ns_in, 64), 64, 0, stream>>>(v, s, seg_tol);
    const long double c = cosl((long double)theta_deg * 3.14159265358979323846264338327950288L /
                               180.0L);
    const double ga = (double)(c * c * (1.0L + 1e-9L));
    k_ck_quality<<<grid_for(v.nt), 256, 0, stream>>>(v, s, ga);
  }
  k_ck_finish<<<1, 64, 0, stream>>>(v, s, out);
  CK(hipGetLastError());
}

View make_view(const Refiner::DeviceMesh &dm) {
  return View{dm.P, dm.np, dm.T, dm.N, dm.nt, dm.S, dm.ns, dm.P_in, dm.np_in, dm.S_in, dm.ns_in};
}

void unpack(const long long *h, long long *viol, long long *exact_signs, CheckExtra *extra) {
  for (int i = 0; i < V_N; i++) viol[i] = h[i];
  if (exact_signs) *exact_signs = h[10];
  if (extra) {
    extra->n_exempt = h[11];
    long long bits = h[12];
    std::memcpy(&extra->min_angle_deg, &bits, sizeof(double));
  }
}

constexpr int CK_EXACT_LANES = 4096;  // lanes of the exact-stage grid

}  // namespace

// ---- scratch of the non-allocating entry point ----
// gcdt_check_mesh_device_noalloc carves every buffer of a check out of this
// array (64 MiB per device, held until the process ends).  In host code its
// name is only a stand-in: the device address is taken once per device with
// hipGetSymbolAddress (noalloc_scratch) and used for every kernel argument.
// Calls that overlap on one device share the array and race.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
constexpr int NOALLOC_EXACT_LANES = 512;  // exact-stage lanes (8 KiB of arena each)
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[MAX_DEVICES] = {nullptr};

// Device address of `scratch` on the current device (looked up on the first
// call per device).
char *noalloc_scratch() {
  int dev = 0;
  CK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) throw std::runtime_error("device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") +
                               hipGetErrorString(e));
    g_scratch_ptr[dev] = (char *)p;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

// Runs every device check on a mesh already in device memory, with scratch
// from hipMalloc (the allocating entry points).
void check_device_mesh(const Refiner::DeviceMesh &dm, hipStream_t stream, double theta_deg,
                       long long *viol, long long *exact_signs, CheckExtra *extra) {
  const View v = make_view(dm);
  const size_t bytes = ck_layout(dm.np, dm.nt, dm.ns, CK_EXACT_LANES, nullptr, nullptr);
  char *base = nullptr;
  long long *dout = nullptr;
  long long h[13] = {0};
  try {
    CK(hipMalloc((void **)&base, bytes));
    CK(hipMalloc((void **)&dout, sizeof(h)));
    Scratch s;
    ck_layout(dm.np, dm.nt, dm.ns, CK_EXACT_LANES, base, &s);
    ck_enqueue(v, s, theta_deg, 1e-9, dout, stream);
    CK(hipMemcpyAsync(h, dout, sizeof(h), hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
  } catch (...) {
    (void)hipFree(base);
    (void)hipFree(dout);
    throw;
  }
  (void)hipFree(base);
  (void)hipFree(dout);
  unpack(h, viol, exact_signs, extra);
}

size_t check_noalloc_bytes(int np, int nt, int ns) {
  return ck_layout(np, nt, ns, NOALLOC_EXACT_LANES, nullptr, nullptr);
}
size_t check_noalloc_limit() { return NOALLOC_SCRATCH_BYTES; }

// The same checks in the static scratch array, results to the device array
// out[13]; no allocation and no host synchronisation.  Throws before
// enqueuing anything when the mesh needs more scratch than the array holds.
void check_device_mesh_noalloc(const Refiner::DeviceMesh &dm, hipStream_t stream,
                               double theta_deg, long long *out) {
  const size_t bytes = check_noalloc_bytes(dm.np, dm.nt, dm.ns);
  if (bytes > NOALLOC_SCRATCH_BYTES)
    throw std::runtime_error("mesh needs " + std::to_string(bytes) +
                             " bytes of check scratch; the static array holds " +
                             std::to_string(NOALLOC_SCRATCH_BYTES) +
                             " (use gcdt_check_mesh for larger meshes)");
  char *base = noalloc_scratch();
  Scratch s;
  ck_layout(dm.np, dm.nt, dm.ns, NOALLOC_EXACT_LANES, base, &s);
  ck_enqueue(make_view(dm), s, theta_deg, 1e-9, out, stream);
}

}  // namespace gcdt
