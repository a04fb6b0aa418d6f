// Device-side invariant checks of a refined constrained Delaunay
// triangulation, with exact predicates: the reference's end-of-run host
// checks (checkNeighbors main.cpp:79-113, checkIncircle main.cpp:153-242,
// the index ranges of checkResult main.cpp:336-362) plus orientation, the
// boundary / subsegment edge rules and edge multiplicity.  Counts per
// category (gcdt_check_mesh in include/gpu_cdt_refine.h):
//   0 index range          1 neighbour symmetry     2 orientation (<= 0)
//   3 boundary edge that is not a subsegment        4 Delaunay: exact
//     incircle > 0 across a non-subsegment edge     5 subsegment that is not
//     a mesh edge          7 input point moved      9 edge in >= 3 triangles
// (6 segment coverage and 8 quality need walks along the input segments and
// stay with the host checker.)
//
// Layout: one thread per triangle / subsegment / edge; subsegments and edges
// go into open-addressing hash tables of 64-bit keys (lo << 32 | hi vertex).
// Signs are decided by the FP64 filter and the double-double stage in the
// triangle kernel; the few that stay undecided are listed and re-evaluated
// with exact expansions by a small grid holding a per-lane arena.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "gcdt_internal.h"
#include "gpu_cdt_refine.h"
#include "predicates.h"

namespace gcdt {

namespace {

enum { V_RANGE = 0, V_SYM = 1, V_ORIENT = 2, V_BOUNDARY = 3, V_DELAUNAY = 4, V_SUBSEG = 5,
       V_MOVED = 7, V_MANIFOLD = 9, V_N = 10 };
constexpr unsigned long long EMPTY_KEY = ~0ull;
constexpr int CK_EXACT_LANES = 4096;  // lanes of the exact-stage grid

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

__global__ void k_ck_fill(unsigned long long *keys, unsigned long long n) {
  for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * blockDim.x)
    keys[i] = EMPTY_KEY;
}
__global__ void k_ck_zero(int *a, unsigned long long n) {
  for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * blockDim.x)
    a[i] = 0;
}

__global__ void k_ck_range(View v, unsigned long long *cnt) {
  unsigned long long bad = 0;
  const long long nT = 3LL * v.nt, nS = 2LL * v.ns;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nT + nS;
       i += (long long)gridDim.x * blockDim.x) {
    if (i < nT) {
      int a = v.T[i], u = v.N[i];
      bad += (a < 0 || a >= v.np) + (u < -1 || u >= v.nt);
    } else {
      long long k = i - nT;
      int2 s = v.S[k >> 1];
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

__global__ void k_ck_seg_insert(View v, unsigned long long *skeys, unsigned long long scap) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < v.ns; s += gridDim.x * blockDim.x) {
    int2 e = v.S[s];
    hash_insert(skeys, scap, edge_key64(e.x, e.y));
  }
}

__global__ void k_ck_edge_insert(View v, unsigned long long *ekeys, int *ecnt,
                                 unsigned long long ecap) {
  const long long n = 3LL * v.nt;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long t = i / 3;
    const int j = (int)(i % 3);
    int a = v.T[3 * t + (j + 1) % 3], b = v.T[3 * t + (j + 2) % 3];
    atomicAdd(&ecnt[hash_insert(ekeys, ecap, edge_key64(a, b))], 1);
  }
}

__global__ void k_ck_edge_count(const unsigned long long *ekeys, const int *ecnt,
                                unsigned long long ecap, unsigned long long *cnt) {
  unsigned long long bad = 0;
  for (unsigned long long h = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; h < ecap;
       h += (unsigned long long)gridDim.x * blockDim.x)
    bad += ekeys[h] != EMPTY_KEY && ecnt[h] >= 3;
  if (bad) atomicAdd(&cnt[V_MANIFOLD], bad);
}

__global__ void k_ck_subseg(View v, const unsigned long long *ekeys, unsigned long long ecap,
                            unsigned long long *cnt) {
  unsigned long long bad = 0;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < v.ns; s += gridDim.x * blockDim.x) {
    int2 e = v.S[s];
    bad += !hash_has(ekeys, ecap, edge_key64(e.x, e.y));
  }
  if (bad) atomicAdd(&cnt[V_SUBSEG], bad);
}

// Per triangle: orientation, then per edge the boundary rule, neighbour
// symmetry and (once per edge, from the lower index) the Delaunay test.
__global__ void k_ck_tri(View v, const unsigned long long *skeys, unsigned long long scap,
                         int *defer, unsigned long long *cnt) {
  unsigned long long bo = 0, bb = 0, bs = 0, bd = 0;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < v.nt; t += gridDim.x * blockDim.x) {
    const int t0 = v.T[3 * t], t1 = v.T[3 * t + 1], t2 = v.T[3 * t + 2];
    const double2 p0 = v.P[t0], p1 = v.P[t1], p2 = v.P[t2];
    bo += orient2d_sign(p0.x, p0.y, p1.x, p1.y, p2.x, p2.y, false, nullptr) <= 0;
    const int tv[3] = {t0, t1, t2};
    for (int j = 0; j < 3; j++) {
      const int a = tv[(j + 1) % 3], b = tv[(j + 2) % 3];
      const int u = v.N[3 * t + j];
      const bool is_seg = hash_has(skeys, scap, edge_key64(a, b));
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
        int pos = atomicAdd(&defer[0], 1);
        defer[1 + pos] = 3 * t + j;
        continue;
      }
      bd += sg > 0;
    }
  }
  if (bo) atomicAdd(&cnt[V_ORIENT], bo);
  if (bb) atomicAdd(&cnt[V_BOUNDARY], bb);
  if (bs) atomicAdd(&cnt[V_SYM], bs);
  if (bd) atomicAdd(&cnt[V_DELAUNAY], bd);
}

// Exact incircle for the listed edges; ws: GCDT_INCIRCLE_WS doubles per lane.
__global__ void k_ck_exact(View v, const int *defer, double *ws, unsigned long long *cnt) {
  const int n = defer[0];
  const int lane = blockIdx.x * blockDim.x + threadIdx.x;
  double *w = ws + (size_t)lane * GCDT_INCIRCLE_WS;
  unsigned long long bd = 0;
  for (int e = lane; e < n; e += gridDim.x * blockDim.x) {
    const int te = defer[1 + e], t = te / 3, j = te % 3;
    const int t0 = v.T[3 * t], t1 = v.T[3 * t + 1], t2 = v.T[3 * t + 2];
    const int a = v.T[3 * t + (j + 1) % 3], b = v.T[3 * t + (j + 2) % 3];
    const int u = v.N[3 * t + j];
    int d = -1;
    for (int k = 0; k < 3; k++)
      if (v.T[3 * u + (k + 1) % 3] == b && v.T[3 * u + (k + 2) % 3] == a) d = v.T[3 * u + k];
    const double2 p0 = v.P[t0], p1 = v.P[t1], p2 = v.P[t2], pd = v.P[d];
    bd += incircle_exact(p0.x, p0.y, p1.x, p1.y, p2.x, p2.y, pd.x, pd.y, w) > 0.0;
  }
  if (bd) atomicAdd(&cnt[V_DELAUNAY], bd);
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

template <class T>
void ck_alloc(T **p, size_t n) {
  CK(hipMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1)));
}

}  // namespace

// Runs every device check on a mesh already in device memory.
void check_device_mesh(const Refiner::DeviceMesh &dm, hipStream_t stream, long long *viol,
                       long long *exact_signs) {
  View v{dm.P, dm.np, dm.T, dm.N, dm.nt, dm.S, dm.ns, dm.P_in, dm.np_in};
  unsigned long long *cnt = nullptr, *skeys = nullptr, *ekeys = nullptr;
  int *ecnt = nullptr, *defer = nullptr;
  double *ws = nullptr;
  unsigned long long h[V_N] = {0};
  if (exact_signs) *exact_signs = 0;
  try {
    ck_alloc(&cnt, V_N);
    CK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * V_N, stream));
    k_ck_range<<<grid_for(3LL * v.nt + 2LL * v.ns), 256, 0, stream>>>(v, cnt);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, stream));
    CK(hipStreamSynchronize(stream));
    if (h[V_RANGE] == 0) {  // every index in range: the other checks may gather
      if (v.P_in && v.np_in > 0)
        k_ck_moved<<<grid_for(std::min(v.np, v.np_in)), 256, 0, stream>>>(v, cnt);
      const unsigned long long scap = pow2_above(2ull * (unsigned long long)v.ns + 16);
      const unsigned long long ecap = pow2_above(6ull * (unsigned long long)v.nt + 16);
      ck_alloc(&skeys, scap);
      ck_alloc(&ekeys, ecap);
      ck_alloc(&ecnt, ecap);
      ck_alloc(&defer, 1 + 3 * (size_t)v.nt);
      k_ck_fill<<<grid_for((long long)scap), 256, 0, stream>>>(skeys, scap);
      k_ck_fill<<<grid_for((long long)ecap), 256, 0, stream>>>(ekeys, ecap);
      k_ck_zero<<<grid_for((long long)ecap), 256, 0, stream>>>(ecnt, ecap);
      k_ck_zero<<<1, 64, 0, stream>>>(defer, 1);
      k_ck_seg_insert<<<grid_for(v.ns), 256, 0, stream>>>(v, skeys, scap);
      k_ck_edge_insert<<<grid_for(3LL * v.nt), 256, 0, stream>>>(v, ekeys, ecnt, ecap);
      k_ck_edge_count<<<grid_for((long long)ecap), 256, 0, stream>>>(ekeys, ecnt, ecap, cnt);
      k_ck_subseg<<<grid_for(v.ns), 256, 0, stream>>>(v, ekeys, ecap, cnt);
      k_ck_tri<<<grid_for(v.nt), 256, 0, stream>>>(v, skeys, scap, defer, cnt);
      CK(hipGetLastError());
      int nd = 0;
      CK(hipMemcpyAsync(&nd, defer, sizeof(int), hipMemcpyDeviceToHost, stream));
      CK(hipStreamSynchronize(stream));
      if (exact_signs) *exact_signs = nd;
      if (nd > 0) {
        ck_alloc(&ws, (size_t)CK_EXACT_LANES * GCDT_INCIRCLE_WS);
        k_ck_exact<<<CK_EXACT_LANES / 64, 64, 0, stream>>>(v, defer, ws, cnt);
        CK(hipGetLastError());
      }
      CK(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, stream));
      CK(hipStreamSynchronize(stream));
    }
  } catch (...) {
    (void)hipFree(cnt); (void)hipFree(skeys); (void)hipFree(ekeys);
    (void)hipFree(ecnt); (void)hipFree(defer); (void)hipFree(ws);
    throw;
  }
  (void)hipFree(cnt); (void)hipFree(skeys); (void)hipFree(ekeys);
  (void)hipFree(ecnt); (void)hipFree(defer); (void)hipFree(ws);
  if (v.np < v.np_in) h[V_MOVED] += (unsigned long long)(v.np_in - v.np);
  for (int i = 0; i < V_N; i++) viol[i] = (long long)h[i];
}

}  // namespace gcdt
