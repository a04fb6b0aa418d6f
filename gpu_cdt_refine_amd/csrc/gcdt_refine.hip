// MI355X (gfx950) CDT quality refinement: kernels and the host round driver.
//
// Round structure (reference GPU_Refine_Quality, refine.cu:6500-7139):
//   1. segment phase: while subsegments are encroached, one conflict pass
//      splits an independent set of them (refine.cu:3180-3778 splitEncsegs);
//   2. bad phase: test new triangles (refine.cu:376-570), then one conflict
//      pass inserts an independent set of Steiner points
//      (refine.cu:5107-5547 splitTriangles).
// A conflict pass is five launches over a dense candidate list:
//   cavity (walk + Bowyer-Watson cavity + ring)  ->  claim (64-bit epoch-
//   tagged atomicMax over cavity and ring)  ->  win (owner check)  ->
//   wave64 ballot/prefix rank of the winners  ->  apply (retriangulate the
//   cavity as a star, in place).
// Winners own their cavity and every triangle across its boundary, so an
// apply never races with another apply and no flip or link-repair phase
// (refine.cu:2328-2490) is needed.
#include <hip/hip_runtime.h>
#include <sys/mman.h>

#include <algorithm>
#include <climits>
#include <cstddef>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "gcdt_device.h"
#include "gcdt_internal.h"

namespace gcdt {

// ---------------------------------------------------------------------------
// Wave64 ballot / block-prefix stream compaction (replaces the reference's
// thrust::copy_if / exclusive_scan calls, refine.cu:149-315, 499-570).
// Output is in ascending index order, so the result is deterministic.
// ---------------------------------------------------------------------------
constexpr int CB = 256;       // threads per block
constexpr int CI = 16;        // items per thread
constexpr int CT = CB * CI;   // items per block tile
constexpr int NBUCKET = 8192;  // priority buckets (edge_bucket range)
constexpr int SUM_TILES = 4096;  // compactions of more tiles use a scan launch

struct PredSenc {
  const int *senc;
  __device__ bool operator()(int i) const { return senc[i] == 1; }
};
// PredSenc over the subsegments that exist after a pass whose new-subsegment
// count is still on the device: [0, ns0 + *nadd) of a range sized for every
// candidate splitting
struct PredSencN {
  const int *senc;
  int ns0;
  const unsigned long long *nadd;
  __device__ bool operator()(int i) const { return i < ns0 + (int)*nadd && senc[i] == 1; }
};
struct PredBad {
  const uint8_t *tflag;
  __device__ bool operator()(int i) const {
    return (tflag[i] & (TF_ALIVE | TF_BAD | TF_GIVEUP)) == (TF_ALIVE | TF_BAD);
  }
};
struct PredWin {
  const int *cwin;
  __device__ bool operator()(int i) const { return cwin[i] != 0; }
};
struct PredAlive {
  const uint8_t *tflag;
  __device__ bool operator()(int i) const { return (tflag[i] & TF_ALIVE) != 0; }
};
struct PredPrioEq {
  const int *prio;
  int b;
  __device__ bool operator()(int i) const { return prio[i] == b; }
};
struct PredPrioSel {  // (prio, index) <= (b, icut)
  const int *prio;
  int b, icut;
  __device__ bool operator()(int i) const {
    int p = prio[i];
    return p < b || (p == b && i <= icut);
  }
};

template <class Pred>
__global__ __launch_bounds__(CB) void k_count(Pred pred, int n, int *blockcnt) {
  __shared__ int wsum[CB / 64];
  int base = blockIdx.x * CT;
  int c = 0;
  for (int k = 0; k < CI; k++) {
    int i = base + k * CB + threadIdx.x;
    if (i < n && pred(i)) c++;
  }
  // wave reduction
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < CB / 64; w++) s += wsum[w];
    blockcnt[blockIdx.x] = s;
  }
}

// Scatter of one tile [base, base + CT): the selected indices in ascending
// order from position run on; returns the position after the tile.
template <class Pred>
__device__ __forceinline__ int scatter_tile(const Pred &pred, int n, int base, int run,
                                            int *out_list, int *out_rank, const int *src) {
  __shared__ int wcnt[CB / 64];
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int k = 0; k < CI; k++) {
    int i = base + k * CB + threadIdx.x;
    bool f = (i < n) && pred(i);
    unsigned long long b = __ballot(f);
    if (lane == 0) wcnt[wave] = __popcll(b);
    __syncthreads();
    int pre = 0, tot = 0;
    for (int w = 0; w < CB / 64; w++) {
      int c = wcnt[w];
      if (w < wave) pre += c;
      tot += c;
    }
    if (f) {
      int pos = run + pre + __popcll(b & lt);
      if (out_list) out_list[pos] = src ? src[i] : i;
      if (out_rank) out_rank[i] = pos;
    }
    run += tot;
    __syncthreads();
  }
  return run;
}

// Exclusive scan of the per-tile counts (one block: wave64 shuffle scans,
// then a scan of the 16 wave totals), for compactions of many tiles; the
// totals go where k_scatter_sum writes them.
__global__ __launch_bounds__(1024) void k_scan(const int *blockcnt, int nb,
                                               int *blockoff, int *total, int *host_total,
                                               unsigned long long *total64) {
  __shared__ int wsum[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int carry = 0;
  for (int base = 0; base < nb; base += 1024) {
    int i = base + threadIdx.x;
    int v = i < nb ? blockcnt[i] : 0;
    int x = v;  // inclusive scan within the wave
    for (int off = 1; off < 64; off <<= 1) {
      int t = __shfl_up(x, off, 64);
      if (lane >= off) x += t;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (wave == 0) {
      int w = lane < 16 ? wsum[lane] : 0;
      for (int off = 1; off < 16; off <<= 1) {
        int t = __shfl_up(w, off, 64);
        if (lane >= off) w += t;
      }
      if (lane < 16) wsum[lane] = w;
    }
    __syncthreads();
    if (i < nb) blockoff[i] = carry + (wave ? wsum[wave - 1] : 0) + x - v;
    carry += wsum[15];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *total = carry;
    if (total64) *total64 = (unsigned long long)carry;
    if (host_total) {
      __hip_atomic_store(host_total, carry, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __threadfence_system();
    }
  }
}

template <class Pred>
__global__ __launch_bounds__(CB) void k_scatter(Pred pred, int n, const int *blockoff,
                                                int *out_list, int *out_rank,
                                                const int *src) {
  scatter_tile(pred, n, blockIdx.x * CT, blockoff[blockIdx.x], out_list, out_rank, src);
}

// Scatter whose block offset is summed from the counts of the blocks before
// it, so no scan launch sits between count and scatter; the last block
// writes the totals: the device copy, and the pinned host-coherent copy
// (read by the host after a stream synchronisation, no copy) or the pass
// counter total64.
template <class Pred>
__global__ __launch_bounds__(CB) void k_scatter_sum(Pred pred, int n, const int *blockcnt,
                                                    int *out_list, int *out_rank,
                                                    const int *src, int *total,
                                                    int *host_total,
                                                    unsigned long long *total64) {
  __shared__ int wred[CB / 64];
  int s = 0;
  for (int b = threadIdx.x; b < (int)blockIdx.x; b += CB) s += blockcnt[b];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) wred[threadIdx.x >> 6] = s;
  __syncthreads();
  int run = 0;
  for (int w = 0; w < CB / 64; w++) run += wred[w];
  run = scatter_tile(pred, n, blockIdx.x * CT, run, out_list, out_rank, src);
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    *total = run;
    if (total64) *total64 = (unsigned long long)run;
    if (host_total) {
      __hip_atomic_store(host_total, run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __threadfence_system();
    }
  }
}

// The whole compaction in one launch when the input fits one tile (n <= CT):
// the scatter from position 0, then the totals as k_scatter_sum writes them.
template <class Pred>
__global__ __launch_bounds__(CB) void k_compact1(Pred pred, int n, int *out_list,
                                                 int *out_rank, const int *src, int *total,
                                                 int *host_total,
                                                 unsigned long long *total64) {
  const int run = scatter_tile(pred, n, 0, 0, out_list, out_rank, src);
  if (threadIdx.x == 0) {
    *total = run;
    if (total64) *total64 = (unsigned long long)run;
    if (host_total) {
      __hip_atomic_store(host_total, run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __threadfence_system();
    }
  }
}

// ---------------------------------------------------------------------------
// Initialisation kernels
// ---------------------------------------------------------------------------
__global__ void k_init_tris(const int *tris, const int *nbrs, const int *tinst,
                            const double2 *P, int nt, TriRec *T, uint8_t *tflag,
                            unsigned long long *claim, unsigned long long *cnt) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  int v0 = tris[3 * t], v1 = tris[3 * t + 1], v2 = tris[3 * t + 2];
  int code[3];
  int vv[3] = {v0, v1, v2};
  for (int j = 0; j < 3; j++) {
    int u = nbrs[3 * t + j];
    code[j] = -1;
    if (u >= 0) {
      int a = vv[(j + 1) % 3], b = vv[(j + 2) % 3];
      for (int k = 0; k < 3; k++)
        if (tris[3 * u + (k + 1) % 3] == b && tris[3 * u + (k + 2) % 3] == a)
          code[j] = (u << 2) | k;
      if (code[j] < 0) atomicAdd(&cnt[CNT_ERRORS], 1ull);
    }
  }
  TriRec r;
  r.V = make_int4(v0, v1, v2, tinst ? tinst[t] : 0);
  r.N = make_int4(code[0], code[1], code[2], 0);
  r.S = make_int4(-1, -1, -1, 0);
  r.X[0] = P[v0];
  r.X[1] = P[v1];
  r.X[2] = P[v2];
  r.pad[0] = r.pad[1] = make_int4(0, 0, 0, 0);
  T[t] = r;
  tflag[t] = TF_ALIVE | TF_CHECK;
  claim[2 * t] = 0ull;
  claim[2 * t + 1] = 0ull;
}

__device__ __forceinline__ unsigned long long edge_key(int a, int b) {
  unsigned long long lo = (unsigned)(a < b ? a : b), hi = (unsigned)(a < b ? b : a);
  return (lo << 32) | hi;
}
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

__global__ void k_init_segs(const int *segs, int ns, int2 *SEG, int2 *SV,
                            int *SP, int *ST, int *senc,
                            unsigned long long *hkey, int *hval,
                            unsigned long long hmask) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  int a = segs[2 * s], b = segs[2 * s + 1];
  SEG[s] = make_int2(a, b);
  SV[s] = make_int2(a, b);
  SP[s] = s;
  ST[s] = -1;
  senc[s] = 0;
  unsigned long long key = edge_key(a, b);
  unsigned long long h = mix64(key) & hmask;
  for (;;) {
    unsigned long long prev = atomicCAS(&hkey[h], ~0ull, key);
    if (prev == ~0ull || prev == key) {
      atomicMin(&hval[h], s);
      break;
    }
    h = (h + 1) & hmask;
  }
}

__global__ void k_attach_segs(MeshD m, int nt, const unsigned long long *hkey,
                              const int *hval, unsigned long long hmask) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  int4 tv = m.T[t].V;
  for (int j = 0; j < 3; j++) {
    int a = comp(tv, (j + 1) % 3), b = comp(tv, (j + 2) % 3);
    unsigned long long key = edge_key(a, b);
    unsigned long long h = mix64(key) & hmask;
    while (hkey[h] != ~0ull) {
      if (hkey[h] == key) {
        int s = hval[h];
        *seg_slot(m.T, t, j) = s;
        int code = (t << 2) | j;
        if (a == m.SEG[s].x) atomicExch(&m.ST[s], code);
        else atomicCAS(&m.ST[s], -1, code);
        break;
      }
      h = (h + 1) & hmask;
    }
  }
}

__global__ void k_init_enc(MeshD m, ParamsD pr, int nt) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const TriRec &r = m.T[t];
  int4 ts = r.S;
  for (int j = 0; j < 3; j++) {
    int s = comp(ts, j);
    if (s < 0) continue;
    double2 p = r.X[j];
    int2 sv = m.SV[s];
    if (encroaches(m, params_at(pr, m, t), sv.x, sv.y, p.x, p.y)) m.senc[s] = 1;
  }
}

__global__ void k_init_points(int np, uint8_t *pflag, int *ppar) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  pflag[i] = PF_INPUT;
  ppar[i] = -1;
}

__global__ void k_check_segs(const int *ST, int ns, unsigned long long *cnt) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < ns && ST[s] < 0) atomicAdd(&cnt[CNT_ERRORS], 1ull);
}

// ---------------------------------------------------------------------------
// Bad-triangle marking (refine.cu:376-497 kernelMarkBadTriangles)
// ---------------------------------------------------------------------------
// nt < 0: the live count and the go flag come from ctl (graph-driven loop).
// Quality test of triangle t if its check flag is set; returns its flags.
__device__ __forceinline__ uint8_t mark_bad_one(const MeshD &m, const ParamsD &pr, int t) {
  uint8_t f = m.tflag[t];
  if (!(f & TF_ALIVE) || !(f & TF_CHECK)) return f;
  f &= (uint8_t)~TF_CHECK;
  const TriRec &r = m.T[t];
  if (tri_is_bad(m, params_at(pr, m, t), r)) {
    f |= TF_BAD;
    // the candidate selection's priority, read later without the record
    m.tbucket[t] = (uint16_t)shortest_edge_bucket(r);
  } else {
    f &= (uint8_t)~TF_BAD;
  }
  m.tflag[t] = f;
  return f;
}

__global__ __launch_bounds__(256) void k_mark_bad(MeshD m, ParamsD pr, int nt,
                                                  const Ctl *ctl) {
  if (ctl) nt = ctl->i[C_GO] ? ctl->i[C_NT] : 0;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nt;
       t += gridDim.x * blockDim.x)
    mark_bad_one(m, pr, t);
}

// k_mark_bad over the compaction's tiles, also counting each tile's bad
// triangles (PredBad) for the scatter that follows: the count pass of the
// bad-triangle compaction rides on the marking pass.
__global__ __launch_bounds__(CB) void k_mark_bad_count(MeshD m, ParamsD pr, int nt,
                                                       int *blockcnt) {
  __shared__ int wsum[CB / 64];
  const int base = blockIdx.x * CT;
  int c = 0;
  for (int k = 0; k < CI; k++) {
    int t = base + k * CB + threadIdx.x;
    if (t < nt) {
      const uint8_t f = mark_bad_one(m, pr, t);
      c += (f & (TF_ALIVE | TF_BAD | TF_GIVEUP)) == (TF_ALIVE | TF_BAD);
    }
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int sum = 0;
    for (int w = 0; w < CB / 64; w++) sum += wsum[w];
    blockcnt[blockIdx.x] = sum;
  }
}

// ---------------------------------------------------------------------------
// Conflict-pass kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ void flush_stats(unsigned long long *cnt,
                                            const PredStats &st) {
  if (st.fp32_hits) atomicAdd(&cnt[CNT_FP32], st.fp32_hits);
  if (st.fp64_hits) atomicAdd(&cnt[CNT_FP64], st.fp64_hits);
  if (st.exact_calls) atomicAdd(&cnt[CNT_EXACT], st.exact_calls);
  if (st.dd_hits) atomicAdd(&cnt[CNT_DD], st.dd_hits);
}

__device__ __forceinline__ void mark_enc(const MeshD &m, int s,
                                         unsigned long long *cnt) {
  if (atomicCAS(&m.senc[s], 0, 1) == 0) atomicAdd(&cnt[CNT_MARKS], 1ull);
}
// Marks subsegment s encroached; true if this call marked it.
__device__ __forceinline__ bool mark_enc_local(const MeshD &m, int s) {
  return atomicCAS(&m.senc[s], 0, 1) == 0;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  return x;
}

// Per-lane event counts of the cavity kernel, summed over the wave and
// added to the pass counters once per wave: a counter shared by every
// candidate would otherwise serialise on one L2 address.
enum { CC_MARKS = 0, CC_GIVEUPS, CC_OVERFLOW, CC_REJ_ENC, CC_REJ_BLOCKED, CC_VALID,
       CC_FP32, CC_FP64, CC_DD, CC_EXACT, CC_EXAM, CC_N };
struct CavCounts {
  unsigned c[CC_N];
  __device__ void add(unsigned long long *cnt, int k, int dst) {
    unsigned long long v = wave_sum(c[k]);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(&cnt[dst], v);
  }
  __device__ void flush(unsigned long long *cnt, bool is_seg) {
    add(cnt, CC_EXAM, is_seg ? CNT_EXAM_SEG : CNT_EXAM_BAD);
    add(cnt, CC_MARKS, CNT_MARKS);
    add(cnt, CC_GIVEUPS, CNT_GIVEUPS);
    add(cnt, CC_OVERFLOW, CNT_OVERFLOW);
    add(cnt, CC_REJ_ENC, CNT_REJ_ENC);
    add(cnt, CC_REJ_BLOCKED, CNT_REJ_BLOCKED);
    add(cnt, CC_VALID, CNT_VALID);
    add(cnt, CC_FP32, CNT_FP32);
    add(cnt, CC_FP64, CNT_FP64);
    add(cnt, CC_DD, CNT_DD);
    add(cnt, CC_EXACT, CNT_EXACT);
  }
};

constexpr int BIG_LDS_BYTES = 160 * 1024;
// Largest grid of the exact-predicate cavity kernel (wave64 blocks); its
// arena holds GCDT_INCIRCLE_WS doubles per lane: 512 x 64 x 8 KiB = 256 MiB.
constexpr int EXACT_GRID = 512;

// ---------------------------------------------------------------------------
// Candidate evaluation: a wave64 evaluates 64/LG candidates at once, one per
// group of LG lanes.
//
// The lanes of a group compute the Steiner / split point and the walk (same
// loads, broadcast); the Bowyer-Watson cavity is then grown level by level,
// the group's lanes testing the three edges of every frontier triangle in
// parallel.  Membership is an LDS hash set (atomicCAS insert), the cavity
// list an LDS array appended with atomicAdd and finally sorted, so the
// result is the cavity *set* in ascending triangle order -- independent of
// timing and of LG.  The boundary / disk / orientation checks, the ring and
// the encroachment test run lane-parallel over the cavity's edges.  Loops
// that contain a barrier run a wave-uniform number of times (groups with
// nothing to do idle through them), so every barrier is reached by the whole
// block.  Several candidates per wave keep several dependent-load chains in
// flight per wave: the kernel is bound by memory latency, not bandwidth.
// The boundary edges are also linked into the star's cyclic order here (the
// successor of edge u->v is the boundary edge leaving v) and written to the
// candidate's workspace, so the apply kernel only scatters records.
// LDS per group: list (2K+2) | cavity hash (HC) | ring hash (HR) | 8 scalars |
// boundary edges 6 x (K+2).
// ---------------------------------------------------------------------------
__host__ __device__ inline int pow2_at_least(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}
// At most K + LG cavity inserts happen (see the overflow rule below).
__host__ __device__ inline int cav_hash_cap(int K, int LG) { return pow2_at_least(2 * (K + LG)); }
// Ring inserts come from boundary edges only: at most 3K of them.
__host__ __device__ inline int ring_hash_cap(int K) { return pow2_at_least(3 * K + 8); }
// list | cavity hash | ring hash | 8 scalars | boundary edges (origin,
// destination, outer otri, subsegment, successor, star order) x (K+2)
__host__ __device__ inline int group_lds_ints(int K, int LG) {
  return (2 * K + 2) + cav_hash_cap(K, LG) + ring_hash_cap(K) + 8 + 6 * (K + 2);
}

__device__ __forceinline__ unsigned hmix(int x) {
  unsigned h = (unsigned)x * 0x9E3779B1u;
  return h ^ (h >> 15);
}
// true if x was not yet present (and is now)
__device__ __forceinline__ bool hset_insert(int *h, int H, int x) {
  unsigned k = hmix(x) & (unsigned)(H - 1);
  for (;;) {
    int prev = atomicCAS(&h[k], -1, x);
    if (prev == -1) return true;
    if (prev == x) return false;
    k = (k + 1) & (unsigned)(H - 1);
  }
}
__device__ __forceinline__ bool hset_has(const int *h, int H, int x) {
  unsigned k = hmix(x) & (unsigned)(H - 1);
  for (;;) {
    int v = h[k];
    if (v == x) return true;
    if (v == -1) return false;
    k = (k + 1) & (unsigned)(H - 1);
  }
}

enum { SC_N = 0, SC_R, SC_OVF, SC_GIVEUP, SC_NBND, SC_ENC, SC_DSEG, SC_SLOW };
constexpr int ST_NONE = -1;  // lane group without a candidate
constexpr int ST_DEFER = 7;  // not evaluated this pass (vertex deletion)
// vertex deletion (below): largest star, subsegment candidates per pass
// that nominate vertices, workspace ints per nomination (star, link, outer
// otri, subsegment: DEL_MAX each; new triangles: 3 (DEL_MAX - 2))
constexpr int DEL_MAX = 32;
constexpr int DEL_CANDS = 1 << 16;
constexpr int DEL_WS = 4 * DEL_MAX + 3 * (DEL_MAX - 2);
constexpr int ST_SLOW = 6;   // deferred: an undecided sign, or a cavity beyond
                             // the fast kernel's LDS capacity

// FAST: every candidate of the pass, with the FP32 / FP64 filters only and
// LDS sized for cavities of at most Kf = min(K, ps.Kf) triangles; a candidate
// that meets an undecided sign or a larger cavity is abandoned without side
// effects and listed in ps.slow.  !FAST: the listed candidates again, with
// the full cascade (double-double, then exact expansions in the per-lane
// arena) and LDS for K triangles.  Neither the rare stages' registers nor the
// largest cavities' LDS thus limit the occupancy of the hot kernel.
template <bool IS_SEG, int LG, bool FAST>
__device__ __forceinline__ void cavity_body(const MeshD &m, const ParamsD &pr,
                                            const PassD &ps, int Kmax) {
  static_assert(LG >= 8 && 64 % LG == 0, "lane groups of 8..64 lanes");
  constexpr int NG = 64 / LG;
  extern __shared__ int sm[];
  // K: the cavity capacity this kernel handles (LDS layout); ps.K: the
  // pass's capacity (workspace layout, overflow rule)
  const int K = FAST ? min(ps.K, ps.Kf) : ps.K;
  const int HC = cav_hash_cap(K, LG), HR = ring_hash_cap(K);
  const int lane = threadIdx.x, grp = lane / LG, gl = lane % LG;
  int *L = sm + grp * group_lds_ints(K, LG);
  int *hc = L + (2 * K + 2);
  int *hr = hc + HC;
  int *sc = hr + HR;
  int *eu = sc + 8, *ev = eu + (K + 2), *eo = ev + (K + 2), *es = eo + (K + 2);
  int *esucc = es + (K + 2), *eord = esucc + (K + 2);
  const bool fp32 = pr.use_fp32 != 0;
  CavCounts cc;
  for (int k = 0; k < CC_N; k++) cc.c[k] = 0u;
  // the exact stages' arena: only the exact kernel (whose grid the arena
  // covers) uses it
  double *pws = FAST ? nullptr : ps.pws + ((size_t)blockIdx.x * 64 + lane) * GCDT_INCIRCLE_WS;
  const int nc = pass_nc(ps);
  const int nseg = pass_nseg(ps);
  const int *ids = pass_ids(ps);
  // work items: this kernel's kind of the pass's candidates (subsegments
  // [0, nseg), bad triangles [nseg, nc)), or the deferred ones of any kind
  const int lo = IS_SEG ? 0 : nseg;
  const int nwork = FAST ? (IS_SEG ? nseg : nc - nseg) : (nc > 0 ? (int)ps.cnt[CNT_SLOW] : 0);
  bool unsure = false;
  auto in_circle = [&](double2 a, double2 b, double2 c, double x, double y, PredStats *q) {
    if (!FAST) return incircle_sign(a.x, a.y, b.x, b.y, c.x, c.y, x, y, pws, fp32, q);
    int sg = incircle_filtered(a.x, a.y, b.x, b.y, c.x, c.y, x, y, fp32, q);
    if (sg == PRED_UNSURE) { unsure = true; sg = 0; }
    return sg;
  };
  auto orient = [&](double ax, double ay, double2 b, double2 c, PredStats *q) {
    if (!FAST) return orient2d_sign(ax, ay, b.x, b.y, c.x, c.y, fp32, q);
    int sg = orient2d_filtered(ax, ay, b.x, b.y, c.x, c.y, fp32, q);
    if (sg == PRED_UNSURE) { unsure = true; sg = 0; }
    return sg;
  };
  for (int base = blockIdx.x * NG; base < nwork; base += gridDim.x * NG) {
    bool act = base + grp < nwork;
    int i = nc;
    if (act) {
      i = FAST ? lo + base + grp : ps.slow[base + grp];
      if (!FAST && (i < nseg) != IS_SEG) {  // deferred candidate of the other kind
        act = false;
        i = nc;
      }
    }
    PredStats cst = {0, 0, 0, 0};  // this candidate's predicate counts
    // ---- point + seeds (uniform across the group) ----
    int status = act ? ST_VALID : ST_NONE, prio = 0, seed0 = -1, seed1 = -1, split_s = -1,
        t0 = -1;
    double cx = 0.0, cy = 0.0;
    double near2 = 0.0;  // a cavity vertex closer than this coincides with the point
    ParamsD lp = pr;
    if (act) {
      // skipped this pass: a bad triangle whose slot the deletion step
      // rewrote
      if (!IS_SEG && (m.tflag[ids[i]] & (TF_ALIVE | TF_BAD)) != (TF_ALIVE | TF_BAD)) {
        status = ST_DEFER;
      } else if (IS_SEG) {
        split_s = ids[i];
        int t1, j1, t2, j2;
        seg_sides(m, split_s, t1, j1, t2, j2);
        if (gl == 0) cc.c[CC_EXAM] += (t1 >= 0) + (t2 >= 0);
        lp = params_at(pr, m, t1 >= 0 ? t1 : t2);
        if (!split_point(m, split_s, cx, cy)) {
          status = ST_GIVEUP;
        } else {
          seed0 = t1 >= 0 ? t1 : t2;
          seed1 = t1 >= 0 ? t2 : -1;
        }
      } else {
        t0 = ids[i];
        lp = params_at(pr, m, t0);
        const TriRec &r0 = m.T[t0];
        const double sh2 = shortest_edge_sq(r0);
        near2 = 1e-12 * sh2;
        if (pr.priority == 1 || pr.priority == 2) prio = edge_bucket(sh2);
        const int sp = steiner_point<FAST>(lp, r0, cx, cy);
        if (sp == STEINER_DEFER) {
          status = ST_SLOW;
        } else if (!sp) {
          status = ST_GIVEUP;
        } else {
          int tl = -1, jl = -1, sg = -1;
          // the walk's signs are counted once per group (every lane walks)
          PredStats wst = {0, 0, 0, 0};
          int wsteps = 0;
          int w = walk<FAST>(m, t0, cx, cy, fp32, &wst, tl, jl, sg, wsteps);
          if (gl == 0) {
            cc.c[CC_EXAM] += (unsigned)wsteps;  // the first step reads the bad triangle
            cst.fp32_hits += wst.fp32_hits;
            cst.fp64_hits += wst.fp64_hits;
            cst.dd_hits += wst.dd_hits;
            cst.exact_calls += wst.exact_calls;
          }
          // Blocked by a subsegment: mark it, or give the triangle up when
          // the subsegment was given up (it is never split, so neither is
          // the triangle).  senc only turns 2 in segment launches, so the
          // group's lanes read the same value.
          if (w == 5) {
            status = ST_SLOW;
          } else if (w == 3) {
            if (m.senc[sg] == 2) {
              status = ST_GIVEUP;
            } else {
              status = ST_REJ_BLOCKED;
              if (gl == 0 && mark_enc_local(m, sg)) cc.c[CC_MARKS]++;
            }
          } else if (w == 2 || w == 4) {
            status = ST_GIVEUP;
          } else {
            seed0 = tl;
            if (w == 1) {
              int s = comp(m.T[tl].S, jl);
              if (s >= 0 && m.senc[s] == 2) {
                status = ST_GIVEUP;
              } else if (s >= 0) {
                status = ST_REJ_BLOCKED;
                if (gl == 0 && mark_enc_local(m, s)) cc.c[CC_MARKS]++;
              } else {
                int nb = comp(m.T[tl].N, jl);
                if (nb < 0) status = ST_GIVEUP;
                else seed1 = nb >> 2;
              }
            }
          }
        }
      }
    }
    // ---- breadth-first growth of the cavity ----
    const bool grow = status == ST_VALID;
    for (int k = gl; k < HC; k += LG) hc[k] = -1;
    for (int k = gl; k < HR; k += LG) hr[k] = -1;
    if (gl < 8) sc[gl] = 0;
    if (gl == 0) sc[SC_DSEG] = 0x7f7fffff;  // +FLT_MAX bits: no subsegment yet
    __syncthreads();
    if (grow && gl == 0) {
      L[0] = seed0;
      hset_insert(hc, HC, seed0);
      int c0 = 1;
      if (seed1 >= 0) {
        L[1] = seed1;
        hset_insert(hc, HC, seed1);
        c0 = 2;
      }
      sc[SC_N] = c0;
    }
    __syncthreads();
    int fb = 0;
    for (;;) {
      int fe = sc[SC_N];
      int ovf = sc[SC_OVF] | sc[SC_SLOW];
      bool more = grow && !ovf && fb < fe;
      __syncthreads();
      if (__ballot(more) == 0ull) break;  // wave-uniform exit
      if (more) {
        int items = (fe - fb) * 3;
        for (int it = gl; it < items; it += LG) {
          int t = L[fb + it / 3], j = it % 3;
          if (comp(m.T[t].S, j) >= 0) continue;  // subsegments are never crossed
          int nb = comp(m.T[t].N, j);
          if (nb < 0) continue;
          int u = nb >> 2;
          if (hset_has(hc, HC, u)) continue;
          const TriRec &ru = m.T[u];
          double2 a = ru.X[0], b = ru.X[1], c = ru.X[2];
          cc.c[CC_EXAM]++;
          int ins = in_circle(a, b, c, cx, cy, &cst);
          if (FAST && unsure) {
            sc[SC_SLOW] = 1;
            unsure = false;
            continue;
          }
          if (ins > 0) {
            // u is in the cavity.  Once K triangles are in, a triangle that is
            // still absent proves the cavity larger than K; inserts stay
            // below K + LG, so the hash set never fills.
            if (hset_has(hc, HC, u)) continue;
            if (sc[SC_N] >= K) {
              sc[SC_OVF] = 1;
            } else if (hset_insert(hc, HC, u)) {
              int pos = atomicAdd(&sc[SC_N], 1);
              if (pos < K) L[pos] = u;
              else sc[SC_OVF] = 1;
            }
          }
        }
      }
      __syncthreads();
      if (more) fb = fe;
    }
    int n = 0, r = 0;
    bool cav_ok = false;
    if (grow) {
      n = sc[SC_N];
      if (FAST && sc[SC_SLOW]) status = ST_SLOW;
      else if ((sc[SC_OVF] || n > K) && K < ps.K) status = ST_SLOW;  // larger cavity
      else if (sc[SC_OVF] || n > K) status = (K >= Kmax) ? ST_GIVEUP : ST_OVERFLOW;
      else cav_ok = true;
    }
    // ---- sort the cavity (rank sort into the free half of L) ----
    int *tmp = L + K + 2;
    if (cav_ok)
      for (int e = gl; e < n; e += LG) {
        int v = L[e], rank = 0;
        for (int x = 0; x < n; x++) rank += L[x] < v;
        tmp[rank] = v;
      }
    __syncthreads();
    if (cav_ok)
      for (int e = gl; e < n; e += LG) L[e] = tmp[e];
    __syncthreads();
    // ---- disk and star checks over every cavity edge ----
    if (cav_ok)
      for (int it = gl; it < 3 * n; it += LG) {
        int t = L[it / 3], j = it % 3;
        const TriRec &rt = m.T[t];
        int s = comp(rt.S, j);
        int nb = comp(rt.N, j);
        bool inside = nb >= 0 && hset_has(hc, HC, nb >> 2);
        if (s >= 0 && s != split_s && inside) sc[SC_GIVEUP] = 1;
        bool bnd = (s >= 0 && s != split_s) || nb < 0 || !inside;
        if (bnd) {
          int4 tv = rt.V;
          if (!IS_SEG) {
            // Steiner point on an existing vertex up to rounding: give up
            // (in exact arithmetic the empty circumcircle keeps every vertex
            // at least half the shortest edge away)
            double2 q = rt.X[(j + 1) % 3];
            double dx = cx - q.x, dy = cy - q.y;
            if (dx * dx + dy * dy < near2) sc[SC_GIVEUP] = 1;
          }
          int bpos = atomicAdd(&sc[SC_NBND], 1);
          if (bpos < K + 2) {
            eu[bpos] = comp(tv, (j + 1) % 3);
            ev[bpos] = comp(tv, (j + 2) % 3);
            eo[bpos] = nb;
            es[bpos] = s;
          }
          if (!(s >= 0 && s == split_s)) {
            double2 u = rt.X[(j + 1) % 3], v = rt.X[(j + 2) % 3];
            int o = orient(cx, cy, u, v, &cst);
            if (FAST && unsure) {
              sc[SC_SLOW] = 1;
              unsure = false;
            } else if (o <= 0) {
              sc[SC_GIVEUP] = 1;
            }
          }
          if (nb >= 0 && !inside && hset_insert(hr, HR, nb >> 2)) {
            int pos = atomicAdd(&sc[SC_R], 1);
            if (pos < K + 2) L[n + pos] = nb >> 2;
          }
        }
        if (!IS_SEG && s >= 0) {
          // The subsegment is this edge.  Ruppert's test (dot < 0) is exactly
          // symmetric in the endpoints, so the record's coordinates serve;
          // Chew's lens product is not, so it takes the subsegment's order.
          double2 A, B;
          if (lp.ruppert) {
            A = rt.X[(j + 1) % 3];
            B = rt.X[(j + 2) % 3];
          } else {
            const int2 sv = m.SV[s];
            A = m.P[sv.x];
            B = m.P[sv.y];
          }
          if (encroaches_xy(lp, A, B, cx, cy)) sc[SC_ENC] = 1;
          if (ps.feat) {  // distance of the Steiner point to this subsegment
            float ex = (float)(B.x - A.x), ey = (float)(B.y - A.y);
            float px = (float)(cx - A.x), py = (float)(cy - A.y);
            float len2 = ex * ex + ey * ey;
            float tt = len2 > 0.f ? fminf(fmaxf((px * ex + py * ey) / len2, 0.f), 1.f) : 0.f;
            float ddx = px - tt * ex, ddy = py - tt * ey;
            atomicMin(&sc[SC_DSEG], __float_as_int(sqrtf(ddx * ddx + ddy * ddy)));
          }
        }
      }
    __syncthreads();
    if (cav_ok) {
      r = sc[SC_R];
      if (FAST && sc[SC_SLOW]) {
        status = ST_SLOW;
      } else if (sc[SC_GIVEUP] || sc[SC_NBND] != n + 2 || r > n + 2) {
        status = ST_GIVEUP;
      } else if (!IS_SEG && sc[SC_ENC]) {
        // Ruppert/Chew rule: a Steiner point that encroaches a subsegment on
        // its cavity boundary is not inserted; every such subsegment is
        // marked.  When every subsegment it encroaches was given up, the
        // triangle is given up too.
        bool live = false;
        for (int it = gl; it < 3 * n; it += LG) {
          int t = L[it / 3], j = it % 3;
          const TriRec &rt = m.T[t];
          int s = comp(rt.S, j);
          if (s < 0) continue;
          bool enc;
          if (lp.ruppert) {
            enc = encroaches_xy(lp, rt.X[(j + 1) % 3], rt.X[(j + 2) % 3], cx, cy);
          } else {
            const int2 sv = m.SV[s];
            enc = encroaches(m, lp, sv.x, sv.y, cx, cy);
          }
          if (enc) {
            live |= m.senc[s] != 2;
            if (mark_enc_local(m, s)) cc.c[CC_MARKS]++;
          }
        }
        // every lane of this group is here (the branch is group-uniform)
        const unsigned long long gmask = LG == 64 ? ~0ull : ((1ull << (LG % 64)) - 1ull);
        const bool any_live = ((__ballot(live) >> (grp * LG)) & gmask) != 0ull;
        status = any_live ? ST_REJ_ENC : ST_GIVEUP;
      }
    }
    // ---- star order of the boundary (valid candidates) ----
    // The n+2 boundary edges form one cycle.  Closed star: it starts at the
    // edge with the smallest origin vertex; a hull subsegment split (open
    // star) starts after the split edge and leaves it out.
    const bool star = status == ST_VALID;
    const int nbe = n + 2;
    if (star)
      for (int e = gl; e < nbe; e += LG) {
        int v = ev[e], f = 0;
        while (f < nbe && eu[f] != v) f++;
        esucc[e] = f < nbe ? f : 0;
      }
    __syncthreads();
    int nstar = 0;
    if (star) {
      int start = 0;
      bool open = false;
      for (int e = 0; e < nbe; e++) {
        if (IS_SEG && es[e] == split_s) {  // only a hull split edge is on the boundary
          open = true;
          start = esucc[e];
          break;
        }
        if (eu[e] < eu[start]) start = e;
      }
      nstar = open ? nbe - 1 : nbe;
      if (gl == 0) {
        int e = start;
        for (int q = 0; q < nstar; q++) {
          eord[q] = e;
          e = esucc[e];
        }
      }
    }
    __syncthreads();
    if (star) {
      int *G = ps.ws + (size_t)i * ps.R + (2 * ps.K + 2);
      for (int q = gl; q < nstar; q += LG) {
        int e = eord[q];
        G[q] = eu[e];
        G[(ps.K + 2) + q] = ev[e];
        G[2 * (ps.K + 2) + q] = eo[e];
        G[3 * (ps.K + 2) + q] = es[e];
      }
    }
    // ---- outputs ----
    if (status == ST_VALID) {
      int *G = ps.ws + (size_t)i * ps.R;
      for (int q = gl; q < n + r; q += LG) G[q] = L[q];
    }
    if (status == ST_SLOW) {  // defer: no output, no side effect, no counts
      if (gl == 0) ps.slow[atomicAdd(&ps.cnt[CNT_SLOW], 1ull)] = i;
    } else {
      cc.c[CC_FP32] += (unsigned)cst.fp32_hits;
      cc.c[CC_FP64] += (unsigned)cst.fp64_hits;
      cc.c[CC_DD] += (unsigned)cst.dd_hits;
      cc.c[CC_EXACT] += (unsigned)cst.exact_calls;
    }
    if (!IS_SEG && ps.feat && act && status != ST_SLOW) {
      // per-candidate features of the learned priority
      float l0 = 0.f, l1 = 0.f, l2 = 0.f, area2 = 0.f;
      if (status == ST_VALID) {
        const TriRec &r0 = m.T[t0];
        double2 A = r0.X[0], O = r0.X[1], D = r0.X[2];
        l0 = (float)((O.x - D.x) * (O.x - D.x) + (O.y - D.y) * (O.y - D.y));
        l1 = (float)((D.x - A.x) * (D.x - A.x) + (D.y - A.y) * (D.y - A.y));
        l2 = (float)((A.x - O.x) * (A.x - O.x) + (A.y - O.y) * (A.y - O.y));
        area2 = fabsf((float)((O.x - A.x) * (D.y - A.y) - (O.y - A.y) * (D.x - A.x)));
      }
      float sh2 = fminf(l0, fminf(l1, l2));
      float R = sqrtf(l0 * l1 * l2) / fmaxf(2.f * area2, 1e-30f);
      float sh = sqrtf(sh2);
      // smallest angle from the law of sines opposite the shortest edge
      float sina = fminf(sh / fmaxf(2.f * R, 1e-30f), 1.f);
      float dseg = __int_as_float(sc[SC_DSEG]);
      for (int f = gl; f < FEAT; f += LG) {
        float fv = 0.f;
        if (status == ST_VALID) {
          if (f == 0) fv = asinf(sina) * (180.f / 3.14159265f) / 60.f;
          else if (f == 1) fv = log2f(fmaxf(R / fmaxf(sh, 1e-30f), 1e-6f)) / 4.f;
          else if (f == 2) fv = (float)n / 16.f;
          else if (f == 3) fv = (float)r / 16.f;
          else if (f == 4) fv = dseg > 1e30f ? 1.f : fminf(dseg / fmaxf(sh, 1e-30f), 4.f) / 4.f;
          else if (f == 5) fv = (float)pass_index_of(ps) / 256.f;
          else if (f == 7) fv = 1.f;
        }
        ps.feat[(size_t)i * FEAT + f] = fv;
      }
    }
    if (act && gl == 0 && status != ST_SLOW) {
      ps.cstatus[i] = status;
      ps.cpt[i] = make_double2(cx, cy);
      ps.cn[i] = status == ST_VALID ? n : 0;
      ps.cr[i] = status == ST_VALID ? r : 0;
      ps.cprio[i] = prio;
      if (status == ST_VALID) cc.c[CC_VALID]++;
      else if (status == ST_OVERFLOW) cc.c[CC_OVERFLOW]++;
      else if (status == ST_REJ_ENC) cc.c[CC_REJ_ENC]++;
      else if (status == ST_REJ_BLOCKED) cc.c[CC_REJ_BLOCKED]++;
      else if (status == ST_GIVEUP) {
        cc.c[CC_GIVEUPS]++;
        if (IS_SEG) m.senc[split_s] = 2;
        else m.tflag[t0] |= TF_GIVEUP;
      }
    }
    __syncthreads();  // LDS is reused by the next candidates
  }
  cc.flush(ps.cnt, IS_SEG);
}

// The filtered kernel is held to 5 waves per SIMD (<= 96 VGPRs, no spills);
// the exact kernel runs on few candidates and keeps its full register file.
template <bool IS_SEG, int LG>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void k_cavity_fast(
    MeshD m, ParamsD pr, PassD ps, int Kmax) {
  cavity_body<IS_SEG, LG, true>(m, pr, ps, Kmax);
}
template <bool IS_SEG, int LG>
__global__ __launch_bounds__(64) void k_cavity_slow(MeshD m, ParamsD pr, PassD ps, int Kmax) {
  cavity_body<IS_SEG, LG, false>(m, pr, ps, Kmax);
}

// Claim key: higher wins.  Epoch in the top 11 bits makes older claims lose
// without a reset pass; then (2^21-1 - priority) and the inverted candidate
// index, so the smallest (priority, index) wins -- the reference's
// atomicMin(trimarker, index) rule (refine.cu:4159-4172) with a priority.
__device__ __forceinline__ unsigned long long claim_key(unsigned long long epoch,
                                                        int prio, int i) {
  return (epoch << 53) | ((unsigned long long)(0x1fffff - prio) << 32) |
         (unsigned long long)(0xffffffffu - (unsigned)i);
}

// Per-candidate list kernels (claims, win test, block test, apply): a group of
// G lanes serves one candidate, one entry of its cavity + ring list (or one
// triangle of its star) per lane, so the candidate's loads and atomics are in
// flight together instead of one dependent memory round trip per entry.
// Group decisions are ballots over the group's lanes.  Groups are aligned
// within the wave and every lane of a group follows the same candidate, so a
// group reaches each ballot / shuffle with all of its lanes.
constexpr int LIST_BLOCK = 256;
static_assert(LIST_BLOCK == CB, "k_rank_apply ranks the winners with the compaction's tile");
constexpr int CLAIM_LANES = 16;
constexpr int APPLY_LANES = 4;

template <int G>
__device__ __forceinline__ unsigned long long group_bits(unsigned long long b) {
  static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "power-of-two groups");
  const int base = (threadIdx.x & 63) & ~(G - 1);
  return G == 64 ? b : (b >> base) & ((1ull << (G == 64 ? 0 : G)) - 1ull);
}

// Claim iteration ps.iter.  Iteration 0: every valid candidate claims its
// cavity and ring.  Later iterations: a candidate still in play (cact) whose
// region touches a winner's stamped region drops out (it is blocked, the
// oracle's rule); the others claim again among themselves.  Stamps carry an
// epoch above every claim of the pass, so concurrent claims never hide them.
template <int G>
__global__ __launch_bounds__(LIST_BLOCK) void k_claim(MeshD m, PassD ps, int *cact) {
  const int nc = pass_nc(ps);
  const unsigned long long ep = pass_epoch(ps);
  const unsigned long long sticky = pass_sticky(ps);
  const int gl = threadIdx.x % G, per = blockDim.x / G;
  const int nseg = pass_nseg(ps);
  unsigned ws_seg = 0, ws_bad = 0;  // workspace ints the cavity kernels wrote
  for (int i = blockIdx.x * per + threadIdx.x / G; i < nc; i += gridDim.x * per) {
    const int *L = ps.ws + (size_t)i * ps.R;
    int prio;
    if (ps.iter == 0) {
      const bool on = ps.cstatus[i] == ST_VALID;
      if (gl == 0) {
        cact[i] = on ? 1 : 0;
        ps.cwin[i] = 0;
      }
      if (!on) continue;
      prio = ps.cprio[i];
      if (ps.score && i >= nseg) {  // the learned priority, kept for the later kernels
        prio = learned_prio(prio, ps.score[i], ps.lshift);
        if (gl == 0) ps.cprio[i] = prio;
      }
      if (gl == 0) {  // cavity + ring list, then the star boundary (4 ints per edge)
        const unsigned w = (unsigned)(ps.cn[i] + ps.cr[i] + 4 * (ps.cn[i] + 2));
        if (i < nseg) ws_seg += w;
        else ws_bad += w;
      }
    } else {
      if (!cact[i]) continue;
      const int n = ps.cn[i], k = n + ps.cr[i];
      bool hit = false;
      for (int q = gl; q < k; q += G) {
        const unsigned long long *c = &m.claim[2 * (size_t)L[q]];
        // shared rings: a ring entry is blocked only by a winner's cavity
        // (stamped in word 0), a cavity entry also by a winner's ring (word 1)
        if (!ps.shared_ring) hit |= (c[0] >> 53) == sticky;
        else if (q >= n) hit |= (c[0] >> 53) == sticky;
        else hit |= ((c[0] >> 53) == sticky) | ((c[1] >> 53) == sticky);
      }
      if (group_bits<G>(__ballot(hit)) != 0ull) {
        if (gl == 0) cact[i] = 0;
        continue;
      }
      prio = ps.cprio[i];
    }
    const unsigned long long key = claim_key(ep, prio, i);
    const int n = ps.cn[i], k = n + ps.cr[i];
    // shared rings: cavity entries claim word 0, ring entries word 1;
    // otherwise every entry claims word 0
    for (int q = gl; q < k; q += G)
      atomicMax(&m.claim[2 * (size_t)L[q] + ((ps.shared_ring && q >= n) ? 1 : 0)], key);
  }
  if (ps.iter == 0) {
    const unsigned long long a = wave_sum(ws_seg), b = wave_sum(ws_bad);
    if ((threadIdx.x & 63) == 0) {
      if (a) atomicAdd(&ps.cnt[CNT_WS_SEG], a);
      if (b) atomicAdd(&ps.cnt[CNT_WS_BAD], b);
    }
  }
}

// A candidate wins when every triangle of its cavity and ring holds its key;
// a winner then stamps its region.  A stamp only replaces the winner's own
// key, which no other candidate's test could match either.
template <int G>
__global__ __launch_bounds__(LIST_BLOCK) void k_win(MeshD m, PassD ps, const int *cact) {
  const int nc = pass_nc(ps);
  const unsigned long long ep = pass_epoch(ps);
  const unsigned long long sticky = pass_sticky(ps);
  const int gl = threadIdx.x % G, per = blockDim.x / G;
  for (int i = blockIdx.x * per + threadIdx.x / G; i < nc; i += gridDim.x * per) {
    if (!cact[i]) continue;
    const unsigned long long key = claim_key(ep, ps.cprio[i], i);
    const int *L = ps.ws + (size_t)i * ps.R;
    const int n = ps.cn[i], k = n + ps.cr[i];
    bool lose = false;
    for (int q = gl; q < k; q += G) {
      const unsigned long long *c = &m.claim[2 * (size_t)L[q]];
      if (!ps.shared_ring) {
        lose |= c[0] != key;  // every entry must hold the key
      } else if (q < n) {
        // cavity: the best cavity claim, and no better ring claim
        lose |= (c[0] != key) | (c[1] > key);
      } else {
        lose |= c[0] > key;  // ring: no better candidate's cavity
      }
    }
    if (group_bits<G>(__ballot(lose)) != 0ull) continue;
    if (gl == 0) ps.cwin[i] = 1;
    if (ps.iter + 1 < ps.iters) {
      const unsigned long long stamp = claim_key(sticky, ps.cprio[i], i);
      for (int q = gl; q < k; q += G)
        m.claim[2 * (size_t)L[q] + ((ps.shared_ring && q >= n) ? 1 : 0)] = stamp;
    }
  }
}

// Retriangulate a winner's cavity as the star of its new vertex; same star
// order and slot rule as the oracle.  Also writes the new point, and for a
// segment split the two subsegment halves.  Lane q of the group writes star
// triangle q and repairs the outer neighbour and subsegment across its
// boundary edge.  The group's records are read (uniform loads) before any of
// its lanes writes, and no lane reads a record another lane writes: the new
// point's coordinates come from the candidate, not from P.
template <int G>
__device__ __forceinline__ void apply_one(const MeshD &m, const ParamsD &pr, const PassD &ps,
                                          int i, int gl, int nseg, const int *ids, int pbase,
                                          int tbase, int sbase, unsigned long long &nflip) {
  {
    if (!ps.cwin[i]) return;
    // winners split subsegments first (candidates [0, nseg)), so a segment
    // winner's rank is also its rank among the new subsegments
    const bool is_seg = i < nseg;
    {
      const unsigned long long wb = __ballot(gl == 0 && is_seg);
      if (wb != 0ull && (int)(threadIdx.x & 63) == __builtin_ctzll(wb))
        atomicAdd(&ps.cnt[CNT_SEGWIN], (unsigned long long)__popcll(wb));
    }
    const int rank = ps.crank[i];
    const int p = pbase + rank;
    const int tb = tbase + 2 * rank;
    const int split_s = is_seg ? ids[i] : -1;
    const double2 c = ps.cpt[i];
    const int n = ps.cn[i];
    const int *L = ps.ws + (size_t)i * ps.R;
    // the boundary edges in star order (origin Bu, destination Bv, outer
    // otri Bo, subsegment Bs) were written by the cavity kernel
    const int *Bu = L + (2 * ps.K + 2);
    const int *Bv = Bu + (ps.K + 2);
    const int *Bo = Bv + (ps.K + 2);
    const int *Bs = Bo + (ps.K + 2);
    const int inst = m.batched ? m.T[L[0]].V.w : 0;  // one instance: no gather
    const ParamsD lp = params_at(pr, m, L[0]);
    bool open = false;
    int a = -1, b = -1, par = -1;
    double2 Pa = make_double2(0.0, 0.0), Pb = Pa;
    if (is_seg) {
      int t1, j1, t2, j2;
      seg_sides(m, split_s, t1, j1, t2, j2);
      open = t1 < 0 || t2 < 0;  // hull subsegment: the star is an open fan
      const int2 sv = m.SV[split_s];
      a = sv.x;
      b = sv.y;
      par = m.SP[split_s];
      Pa = m.P[a];
      Pb = m.P[b];
    }
    const int nstar = open ? n + 1 : n + 2;
    // flips of the same insertion by Lawson's algorithm: the first triangle
    // (both triangles of a two-sided subsegment) split, then one flip per
    // further cavity triangle
    if (gl == 0) nflip += (unsigned long long)(n - ((is_seg && !open) ? 2 : 1));
    const int sA = split_s, sB = is_seg ? sbase + rank : -1;
    // first star triangle (in star order) at each end of the split, packed
    // as (q << 32 | otri) so that a minimum finds it
    long long kA = LLONG_MAX, kB = LLONG_MAX;
    bool encA = false, encB = false;
    for (int q = gl; q < nstar; q += G) {
      const int st = q < n ? L[q] : tb + (q - n);
      const int nx = (q + 1 < nstar) ? q + 1 : (open ? -1 : 0);
      const int pv = (q > 0) ? q - 1 : (open ? -1 : nstar - 1);
      const int snx = nx < 0 ? -1 : (nx < n ? L[nx] : tb + (nx - n));
      const int spv = pv < 0 ? -1 : (pv < n ? L[pv] : tb + (pv - n));
      const int u = Bu[q], v = Bv[q], o = Bo[q], s = Bs[q];
      const double2 Pu = m.P[u], Pv = m.P[v];
      TriRec r;
      r.V = make_int4(p, u, v, inst);
      r.N = make_int4(o, snx >= 0 ? (snx << 2) | 2 : -1, spv >= 0 ? (spv << 2) | 1 : -1, 0);
      r.S = make_int4(s, (v == a) ? sA : (v == b) ? sB : -1, (u == a) ? sA : (u == b) ? sB : -1,
                      0);
      r.X[0] = c;
      r.X[1] = Pu;
      r.X[2] = Pv;
      r.pad[0] = r.pad[1] = make_int4(0, 0, 0, 0);
      m.T[st] = r;
      m.tflag[st] = TF_ALIVE | TF_CHECK;
      if (o >= 0) *nbr_slot(m.T, o >> 2, o & 3) = (st << 2) | 0;
      if (s >= 0) {
        m.ST[s] = (st << 2) | 0;
        if (is_seg && m.senc[s] == 0) {
          // s is the boundary edge (u, v); see the cavity kernel on endpoint order
          bool enc;
          if (lp.ruppert) {
            enc = encroaches_xy(lp, Pu, Pv, c.x, c.y);
          } else {
            const int2 sv = m.SV[s];
            enc = encroaches(m, lp, sv.x, sv.y, c.x, c.y);
          }
          if (enc) m.senc[s] = 1;
        }
      }
      if (is_seg) {
        const long long hq = (long long)q << 32;
        if (u == a) {
          kA = min(kA, hq | ((st << 2) | 2));
          encA |= encroaches_xy(lp, Pa, c, Pv.x, Pv.y);
        }
        if (v == a) {
          kA = min(kA, hq | ((st << 2) | 1));
          encA |= encroaches_xy(lp, Pa, c, Pu.x, Pu.y);
        }
        if (u == b) {
          kB = min(kB, hq | ((st << 2) | 2));
          encB |= encroaches_xy(lp, c, Pb, Pv.x, Pv.y);
        }
        if (v == b) {
          kB = min(kB, hq | ((st << 2) | 1));
          encB |= encroaches_xy(lp, c, Pb, Pu.x, Pu.y);
        }
      }
    }
    if (is_seg) {
      for (int off = G / 2; off > 0; off >>= 1) {
        kA = min(kA, __shfl_xor(kA, off, G));
        kB = min(kB, __shfl_xor(kB, off, G));
      }
      const bool eA = group_bits<G>(__ballot(encA)) != 0ull;
      const bool eB = group_bits<G>(__ballot(encB)) != 0ull;
      if (gl == 0) {
        m.SV[sA] = make_int2(a, p);
        m.SV[sB] = make_int2(p, b);
        m.SP[sB] = par;
        m.ST[sA] = kA == LLONG_MAX ? -1 : (int)(kA & 0xffffffffll);
        m.ST[sB] = kB == LLONG_MAX ? -1 : (int)(kB & 0xffffffffll);
        m.senc[sA] = eA ? 1 : 0;
        m.senc[sB] = eB ? 1 : 0;
      }
    }
    if (gl == 0) {
      m.P[p] = c;
      m.pflag[p] = is_seg ? PF_SEGV : PF_FREE;
      m.ppar[p] = is_seg ? par : -1;
      if (open) {  // the second new slot stays empty
        const int st = tb + 1;
        TriRec r;
        r.V = make_int4(-1, -1, -1, 0);
        r.N = make_int4(-1, -1, -1, 0);
        r.S = make_int4(-1, -1, -1, 0);
        r.X[0] = r.X[1] = r.X[2] = make_double2(0.0, 0.0);
        r.pad[0] = r.pad[1] = make_int4(0, 0, 0, 0);
        m.T[st] = r;
        m.tflag[st] = 0;
      }
    }
  }
}

template <int G>
__global__ __launch_bounds__(LIST_BLOCK) void k_apply(MeshD m, ParamsD pr, PassD ps, int pbase,
                                                      int tbase, int sbase) {
  const int nc = pass_nc(ps);
  const int nseg = pass_nseg(ps);
  const int *ids = pass_ids(ps);
  if (ps.ctl) {
    pbase = ps.ctl->i[C_NP];
    tbase = ps.ctl->i[C_NT];
    sbase = ps.ctl->i[C_NS];
  }
  const int gl = threadIdx.x % G, per = blockDim.x / G;
  unsigned long long nflip = 0;
  for (int i = blockIdx.x * per + threadIdx.x / G; i < nc; i += gridDim.x * per)
    apply_one<G>(m, pr, ps, i, gl, nseg, ids, pbase, tbase, sbase, nflip);
  nflip = wave_sum(nflip);
  if ((threadIdx.x & 63) == 0 && nflip) atomicAdd(&ps.cnt[CNT_FLIPS], nflip);
}

// Small host-loop passes (nc <= SMALL_PASS): one block ranks the winners
// (the compaction's one-tile scatter) and then applies them -- one launch
// instead of two; the ranks reach the block's other threads through the
// barrier.
constexpr int SMALL_PASS = 256;
template <int G>
__global__ __launch_bounds__(LIST_BLOCK) void k_rank_apply(MeshD m, ParamsD pr, PassD ps,
                                                           int pbase, int tbase, int sbase) {
  const int nc = ps.nc;
  const int run = scatter_tile(PredWin{ps.cwin}, nc, 0, 0, nullptr, ps.crank, nullptr);
  if (threadIdx.x == 0) ps.cnt[CNT_NWIN] = (unsigned long long)run;
  __syncthreads();
  const int gl = threadIdx.x % G, per = blockDim.x / G;
  unsigned long long nflip = 0;
  for (int i = threadIdx.x / G; i < nc; i += per)
    apply_one<G>(m, pr, ps, i, gl, ps.nseg, ps.ids, pbase, tbase, sbase, nflip);
  nflip = wave_sum(nflip);
  if ((threadIdx.x & 63) == 0 && nflip) atomicAdd(&ps.cnt[CNT_FLIPS], nflip);
}

// ---------------------------------------------------------------------------
// Vertex deletion (Chew's rule as Triangle applies it before splitting an
// encroached subsegment, triangle.cpp:13319-13357; the reference marks the
// same vertices and removes them by flips, refine.cu:604-677, 778-1070).
// Before the cavities of a pass with subsegment candidates, each of the
// first DEL_CANDS candidates nominates the apexes on both sides of its
// subsegment (candidate d = 2 i + side) when the apex is a free vertex
// strictly inside the subsegment's diametral circle.  A nominee whose star
// is closed, has at most DEL_MAX triangles and no subsegment at the vertex,
// and whose link polygon the ear rule retriangulates, is a deletion
// candidate; it claims its star and the outer triangles across its link, the
// smallest d wins each, and a candidate holding all of its triangles deletes
// its vertex.  The pass's cavities are built afterwards: a subsegment splits
// in the same pass as its apexes are deleted (Triangle deletes, then splits),
// and bad-triangle candidates whose slot was rewritten skip the pass.  One
// apex per side and pass is deleted, where Triangle repeats on the new apex.
// The ear rule: repeatedly cut the first ear of the remaining link polygon
// that turns left and whose circumcircle holds no other remaining vertex
// strictly inside (exact signs); the cut triangles are the constrained
// Delaunay triangles inside the star once the vertex is gone.  One lane per
// candidate: stars are small and deletions rare.
// ---------------------------------------------------------------------------

struct DelD {
  const int *ids;  // subsegment candidates of the pass
  int nseg;        // host loop: their number (graph loop: from ctl)
  Ctl *ctl;
  unsigned long long epoch;  // host loop: the claim epoch
  int policy;                // 1: Chew-mode instances only, 2: every instance
  int *ws, *dv, *dk;
  double *pws;               // exact-predicate arena, one slot per lane
  unsigned long long *cnt;
  int use_fp32;
};

__device__ __forceinline__ int del_nseg(const DelD &d) {
  const int n = d.ctl ? (d.ctl->i[C_GO] ? d.ctl->i[C_NC] : 0) : d.nseg;
  return n < DEL_CANDS ? n : DEL_CANDS;
}
__device__ __forceinline__ unsigned long long del_epoch(const DelD &d) {
  return d.ctl ? (unsigned long long)d.ctl->i[C_EPOCH] + 1ull : d.epoch;
}

// Apex of triangle t (corner j) on subsegment s: a deletion candidate?
// Writes the star, link, outer otri, subsegments and the cut triangles to W
// and the star size to *kout; returns the vertex or -1.  The link's indices,
// coordinates (taken from the star's records) and the ear bookkeeping live
// in this lane's column of the block's LDS arrays (lane-interleaved: entry i
// of lane l at i * 64 + l).
constexpr int DEL_LANES = 64;
__device__ int del_candidate(const MeshD &m, int s, int t, int j, int *W, int *kout,
                             double *pws, bool fp32, int *lp, double2 *lc,
                             unsigned char *lrem) {
  const int ln = threadIdx.x;
  int *star = W, *p = W + DEL_MAX, *o = W + 2 * DEL_MAX, *sg = W + 3 * DEL_MAX;
  int *nt = W + 4 * DEL_MAX;
  const TriRec &r0 = m.T[t];
  const int v = comp(r0.V, j);
  if ((m.pflag[v] & (PF_FREE | PF_DEAD)) != PF_FREE) return -1;
  const int2 sv = m.SV[s];
  const double2 O = m.P[sv.x], D = m.P[sv.y], V = r0.X[j];
  if (!((O.x - V.x) * (D.x - V.x) + (O.y - V.y) * (D.y - V.y) < 0.0)) return -1;
  int cur = t, jc = j, k = 0;
  for (;;) {
    if (k == DEL_MAX) return -1;
    const TriRec &r = m.T[cur];
    const int4 rv = r.V, rn = r.N, rs = r.S;
    const int x = comp(rv, (jc + 1) % 3), y = comp(rv, (jc + 2) % 3);
    if (k == 0) {
      lp[ln] = x;
      lc[ln] = r.X[(jc + 1) % 3];
    } else if (lp[k * DEL_LANES + ln] != x) {
      return -1;
    }
    star[k] = cur;
    o[k] = comp(rn, jc);
    sg[k] = comp(rs, jc);
    if (comp(rs, (jc + 1) % 3) >= 0) return -1;  // edge (v, y) is a subsegment
    const int nb = comp(rn, (jc + 1) % 3);
    if (nb < 0) return -1;
    k++;
    const int nx = nb >> 2;
    if (nx == t) {
      if (y != lp[ln]) return -1;
      break;
    }
    if (k < DEL_MAX) {
      lp[k * DEL_LANES + ln] = y;
      lc[k * DEL_LANES + ln] = r.X[(jc + 2) % 3];
    }
    const int4 vn = m.T[nx].V;
    const int jn = vn.x == v ? 0 : (vn.y == v ? 1 : (vn.z == v ? 2 : -1));
    if (jn < 0) return -1;
    cur = nx;
    jc = jn;
  }
  if (k < 3) return -1;
  for (int q = 0; q < k; q++) p[q] = lp[q * DEL_LANES + ln];
  auto REM = [&](int q) -> int { return lrem[q * DEL_LANES + ln]; };
  auto PT = [&](int q) -> double2 { return lc[REM(q) * DEL_LANES + ln]; };
  auto ID = [&](int q) -> int { return lp[REM(q) * DEL_LANES + ln]; };
  int r = k, ne = 0;
  for (int q = 0; q < k; q++) lrem[q * DEL_LANES + ln] = (unsigned char)q;
  while (r > 3) {
    int found = -1;
    for (int q = 0; q < r && found < 0; q++) {
      const int qa = (q + r - 1) % r, qc = (q + 1) % r;
      const double2 A = PT(qa), B = PT(q), C = PT(qc);
      if (orient2d_sign(A.x, A.y, B.x, B.y, C.x, C.y, fp32, nullptr) <= 0) continue;
      bool ok = true;
      for (int w = 0; w < r && ok; w++) {
        if (w == qa || w == q || w == qc) continue;
        const double2 E = PT(w);
        if (incircle_sign(A.x, A.y, B.x, B.y, C.x, C.y, E.x, E.y, pws, fp32, nullptr) > 0) ok = false;
      }
      if (ok) found = q;
    }
    if (found < 0) return -1;
    nt[3 * ne] = ID((found + r - 1) % r);
    nt[3 * ne + 1] = ID(found);
    nt[3 * ne + 2] = ID((found + 1) % r);
    ne++;
    for (int q = found; q + 1 < r; q++) lrem[q * DEL_LANES + ln] = lrem[(q + 1) * DEL_LANES + ln];
    r--;
  }
  {
    const double2 A = PT(0), B = PT(1), C = PT(2);
    if (orient2d_sign(A.x, A.y, B.x, B.y, C.x, C.y, fp32, nullptr) <= 0) return -1;
  }
  nt[3 * ne] = ID(0);
  nt[3 * ne + 1] = ID(1);
  nt[3 * ne + 2] = ID(2);
  *kout = k;
  return v;
}

// Nominations, stars and retriangulations, one lane per nomination (both
// sides of a subsegment in parallel); a candidate then claims its star and
// the outer triangles across its link (claims commute, so no other lane's
// star needs to be known).
__global__ __launch_bounds__(DEL_LANES) void k_del_scan(MeshD m, ParamsD pr, DelD dd) {
  __shared__ int s_lp[DEL_MAX * DEL_LANES];
  __shared__ double2 s_lc[DEL_MAX * DEL_LANES];
  __shared__ unsigned char s_rem[DEL_MAX * DEL_LANES];
  const int nd = 2 * del_nseg(dd);
  const unsigned long long ep = del_epoch(dd);
  const int lane = blockIdx.x * blockDim.x + threadIdx.x;
  double *pws = dd.pws + (size_t)lane * GCDT_INCIRCLE_WS;
  for (int d = lane; d < nd; d += gridDim.x * blockDim.x) {
    const int s = dd.ids[d >> 1], side = d & 1;
    int t1, j1, t2, j2;
    seg_sides(m, s, t1, j1, t2, j2);
    const int t = side ? t2 : t1, j = side ? j2 : j1;
    int v = -1;
    int *W = dd.ws + (size_t)d * DEL_WS;
    if (t >= 0) {
      const ParamsD lp = params_at(pr, m, t);
      if (dd.policy == 2 || (dd.policy == 1 && !lp.ruppert))
        v = del_candidate(m, s, t, j, W, dd.dk + d, pws, dd.use_fp32 != 0, s_lp, s_lc, s_rem);
    }
    dd.dv[d] = v;
    if (v >= 0) {
      const int k = dd.dk[d];
      const unsigned long long key = claim_key(ep, 0, d);
      for (int q = 0; q < k; q++) {
        atomicMax(&m.claim[2 * (size_t)W[q]], key);
        const int o = W[2 * DEL_MAX + q];
        if (o >= 0) atomicMax(&m.claim[2 * (size_t)(o >> 2)], key);
      }
    }
  }
}

// A candidate holding every triangle of its star and ring deletes its
// vertex right away: winners' regions are disjoint and the other lanes' win
// tests read only claim words, which no apply writes.  The cut triangles
// take the star's slots in ascending order (the two largest slots die);
// links inside the star are matched edge by edge, link edges take the outer
// triangle (whose back-link is repaired) and the subsegment (whose adjacent
// otri is repaired).
__global__ __launch_bounds__(DEL_LANES) void k_del_apply(MeshD m, DelD dd) {
  // this lane's link, outer otri, subsegments, sorted slots and cut
  // triangles, lane-interleaved in LDS for the edge matching below
  __shared__ int a_p[DEL_MAX * DEL_LANES], a_o[DEL_MAX * DEL_LANES], a_s[DEL_MAX * DEL_LANES],
      a_sl[DEL_MAX * DEL_LANES], a_nt[3 * (DEL_MAX - 2) * DEL_LANES];
  const int ln = threadIdx.x;
  const int nd = 2 * del_nseg(dd);
  const unsigned long long ep = del_epoch(dd);
  unsigned long long ndel = 0, nflip = 0;
  for (int d = blockIdx.x * blockDim.x + threadIdx.x; d < nd; d += gridDim.x * blockDim.x) {
    if (dd.dv[d] < 0) continue;
    const int *W = dd.ws + (size_t)d * DEL_WS;
    const int k = dd.dk[d];
    auto P_ = [&](int i) -> int & { return a_p[i * DEL_LANES + ln]; };
    auto O_ = [&](int i) -> int & { return a_o[i * DEL_LANES + ln]; };
    auto S_ = [&](int i) -> int & { return a_s[i * DEL_LANES + ln]; };
    auto SL = [&](int i) -> int & { return a_sl[i * DEL_LANES + ln]; };
    auto NT = [&](int i) -> int & { return a_nt[i * DEL_LANES + ln]; };
    // the workspace first (independent loads), then every claim word of the
    // star and ring (independent again): two round trips, not two per entry
    // (unrolled by 8: eight loads in flight before their first use)
#pragma unroll 8
    for (int q = 0; q < k; q++) {
      SL(q) = W[q];
      O_(q) = W[2 * DEL_MAX + q];
    }
    {
      const unsigned long long key = claim_key(ep, 0, d);
      bool w = true;
#pragma unroll 8
      for (int q = 0; q < k; q++) {
        const int o = O_(q);
        const bool a = m.claim[2 * (size_t)SL(q)] == key;
        const bool b = o < 0 || m.claim[2 * (size_t)(o >> 2)] == key;
        w = w & a & b;
      }
      if (!w) continue;
    }
#pragma unroll 8
    for (int q = 0; q < k; q++) {
      P_(q) = W[DEL_MAX + q];
      S_(q) = W[3 * DEL_MAX + q];
    }
#pragma unroll 8
    for (int q = 0; q < 3 * (k - 2); q++) NT(q) = W[4 * DEL_MAX + q];
    const int inst = m.batched ? m.T[SL(0)].V.w : 0;
    for (int q = 1; q < k; q++) {
      const int x = SL(q);
      int w = q - 1;
      while (w >= 0 && SL(w) > x) {
        SL(w + 1) = SL(w);
        w--;
      }
      SL(w + 1) = x;
    }
    for (int q = 0; q < k - 2; q++) {
      const int st = SL(q);
      const int V0 = NT(3 * q), V1 = NT(3 * q + 1), V2 = NT(3 * q + 2);
      int nn[3], ss[3];
      for (int e = 0; e < 3; e++) {
        const int u = e == 0 ? V1 : (e == 1 ? V2 : V0);
        const int w = e == 0 ? V2 : (e == 1 ? V0 : V1);
        int nb = -2, sgv = -1;
        for (int i = 0; i < k && nb == -2; i++)
          if (P_(i) == u && P_((i + 1) % k) == w) {
            nb = O_(i);
            sgv = S_(i);
            if (nb >= 0) *nbr_slot(m.T, nb >> 2, nb & 3) = (st << 2) | e;
            if (sgv >= 0) m.ST[sgv] = (st << 2) | e;
          }
        for (int q2 = 0; q2 < k - 2 && nb == -2; q2++) {
          if (q2 == q) continue;
          for (int e2 = 0; e2 < 3; e2++)
            if (NT(3 * q2 + (e2 + 1) % 3) == w && NT(3 * q2 + (e2 + 2) % 3) == u) {
              nb = (SL(q2) << 2) | e2;
              break;
            }
        }
        nn[e] = nb;
        ss[e] = sgv;
      }
      TriRec r;
      r.V = make_int4(V0, V1, V2, inst);
      r.N = make_int4(nn[0], nn[1], nn[2], 0);
      r.S = make_int4(ss[0], ss[1], ss[2], 0);
      r.X[0] = m.P[V0];
      r.X[1] = m.P[V1];
      r.X[2] = m.P[V2];
      r.pad[0] = r.pad[1] = make_int4(0, 0, 0, 0);
      m.T[st] = r;
      m.tflag[st] = TF_ALIVE | TF_CHECK;
    }
    for (int q = k - 2; q < k; q++) {
      TriRec r;
      r.V = make_int4(-1, -1, -1, 0);
      r.N = make_int4(-1, -1, -1, 0);
      r.S = make_int4(-1, -1, -1, 0);
      r.X[0] = r.X[1] = r.X[2] = make_double2(0.0, 0.0);
      r.pad[0] = r.pad[1] = make_int4(0, 0, 0, 0);
      m.T[SL(q)] = r;
      m.tflag[SL(q)] = 0;
    }
    m.pflag[dd.dv[d]] |= PF_DEAD;
    ndel++;
    nflip += (unsigned long long)(k - 3);
  }
  ndel = wave_sum(ndel);
  nflip = wave_sum(nflip);
  if ((threadIdx.x & 63) == 0) {
    if (ndel) atomicAdd(&dd.cnt[CNT_DEL], ndel);
    if (nflip) atomicAdd(&dd.cnt[CNT_FLIPS], nflip);
  }
}

// Export with deleted vertices: live points keep their order (input points
// are never deleted, so they keep their indices) and every reference is
// renumbered.
struct PredPAlive {
  const uint8_t *pflag;
  __device__ bool operator()(int i) const { return (pflag[i] & PF_DEAD) == 0; }
};
__global__ void k_pcompact(const double2 *P, const uint8_t *pflag, const int *ppar, int np,
                           const int *pnew, double2 *P2, uint8_t *pflag2, int *ppar2) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
    if (pflag[i] & PF_DEAD) continue;
    const int q = pnew[i];
    P2[q] = P[i];
    pflag2[q] = pflag[i];
    ppar2[q] = ppar[i];
  }
}
__global__ void k_premap(MeshD m, int nt, int ns, const int *pnew) {
  const int n = nt > ns ? nt : ns;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (i < nt && (m.tflag[i] & TF_ALIVE)) {
      int4 v = m.T[i].V;
      v.x = pnew[v.x];
      v.y = pnew[v.y];
      v.z = pnew[v.z];
      m.T[i].V = v;
    }
    if (i < ns) {
      const int2 sv = m.SV[i];
      m.SV[i] = make_int2(pnew[sv.x], pnew[sv.y]);
    }
  }
}

// ---------------------------------------------------------------------------
// Export (refine.cu:5908-6484 compaction + format conversion)
// ---------------------------------------------------------------------------
__global__ void k_export(MeshD m, int nt, const int *newid, int *otri,
                         int *onbr, uint8_t *otstat, int *oinst) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  uint8_t f = m.tflag[t];
  if (!(f & TF_ALIVE)) return;
  int q = newid[t];
  int4 tv = m.T[t].V, tn = m.T[t].N;
  otri[3 * q] = tv.x;
  otri[3 * q + 1] = tv.y;
  otri[3 * q + 2] = tv.z;
  onbr[3 * q] = tn.x < 0 ? -1 : newid[tn.x >> 2];
  onbr[3 * q + 1] = tn.y < 0 ? -1 : newid[tn.y >> 2];
  onbr[3 * q + 2] = tn.z < 0 ? -1 : newid[tn.z >> 2];
  // TStatus bits: 0 non-null, 1 bad, 2 check
  otstat[q] = (uint8_t)(1 | ((f & TF_BAD) ? 2 : 0) | ((f & TF_CHECK) ? 4 : 0));
  if (oinst) oinst[q] = tv.w;
}

// Given-up elements left at the end of a run: alive triangles still flagged
// TF_GIVEUP (bad, never split) and subsegments whose split was given up.
__global__ __launch_bounds__(256) void k_count_left(const uint8_t *tflag, int nt, const int *senc,
                                                    int ns, unsigned long long *cnt) {
  unsigned long long a = 0, b = 0;
  const int n = nt > ns ? nt : ns;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (i < nt && (tflag[i] & (TF_ALIVE | TF_GIVEUP)) == (TF_ALIVE | TF_GIVEUP)) a++;
    if (i < ns && senc[i] == 2) b++;
  }
  a = wave_sum(a);
  b = wave_sum(b);
  if ((threadIdx.x & 63) == 0) {
    if (a) atomicAdd(&cnt[CNT_LEFT_TRI], a);
    if (b) atomicAdd(&cnt[CNT_LEFT_SEG], b);
  }
}

__global__ void k_pstatus(const uint8_t *pflag, int np, uint8_t *out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  uint8_t f = pflag[i];
  // PStatus bits: 1 Steiner, 2 segment split, 3 new (refine.h:36-106)
  out[i] = (uint8_t)(((f & PF_INPUT) ? 0 : 2) | ((f & PF_SEGV) ? 4 : 0) | ((f & PF_NEW) ? 8 : 0));
}

__global__ void k_mark_new(uint8_t *pflag, int from, int np) {
  for (int i = from + blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x)
    pflag[i] |= PF_NEW;
}

// Device predicate evaluation for tests.
__global__ void k_predicates(const double *q, int n, int pred_mode, int *o,
                             int *ic, double *pws, unsigned long long *cnt) {
  PredStats st = {0, 0, 0, 0};
  int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  double *ws = pws + (size_t)gtid * GCDT_INCIRCLE_WS;
  for (int i = gtid; i < n; i += gridDim.x * blockDim.x) {
    const double *p = q + 8 * (size_t)i;
    o[i] = orient2d_sign(p[0], p[1], p[2], p[3], p[4], p[5], pred_mode == 1, &st);
    ic[i] = incircle_sign(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], ws,
                          pred_mode == 1, &st);
  }
  flush_stats(cnt, st);
}

// Constructions on the device, for tests: q = (p0, p1, p2) per query;
// ccw = Triangle's adaptive counterclockwise(p0, p1, p2) value, and the
// refinement's Steiner point of the triangle with apex p0, org p1, dest p2.
__global__ void k_constructions(const double *q, int n, ParamsD pr, double *ccw, double *stp,
                                int *ok) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double *p = q + 6 * (size_t)i;
    ccw[i] = orient2d_value(p[0], p[1], p[2], p[3], p[4], p[5]);
    TriRec r;
    r.X[0] = make_double2(p[0], p[1]);
    r.X[1] = make_double2(p[2], p[3]);
    r.X[2] = make_double2(p[4], p[5]);
    double cx = 0.0, cy = 0.0;
    ok[i] = steiner_point<false>(pr, r, cx, cy);
    stp[2 * (size_t)i] = cx;
    stp[2 * (size_t)i + 1] = cy;
  }
}

// Learned priority: cprio = (bucket >> shift) * 256 + (255 - quantised
// predicted conflict degree).  Inside a (coarsened) shortest-edge bucket the
// candidate whose cavity and ring overlap the most other candidates' regions
// wins: it sits inside a cluster of bad triangles and its insertion clears
// the most of it, so clusters resolve in fewer rounds.  The graph loop
// applies it with this kernel; the host loop inside its first claim
// iteration (PassD::score).
__global__ void k_learned_prio(PassD ps, const float *score, int shift) {
  const int nc = pass_nc(ps);
  for (int i = pass_nseg(ps) + blockIdx.x * blockDim.x + threadIdx.x; i < nc;
       i += gridDim.x * blockDim.x) {
    if (ps.cstatus[i] != ST_VALID) continue;
    ps.cprio[i] = learned_prio(ps.cprio[i], score[i], shift);
  }
}

// Conflict logs: how many candidates claim each triangle, then per candidate
// the number of other claims on its cavity and ring.
__global__ void k_claim_count(PassD ps, int *tcnt) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ps.nc; i += gridDim.x * blockDim.x) {
    if (ps.cstatus[i] != ST_VALID) continue;
    const int *L = ps.ws + (size_t)i * ps.R;
    int k = ps.cn[i] + ps.cr[i];
    for (int q = 0; q < k; q++) atomicAdd(&tcnt[L[q]], 1);
  }
}
__global__ void k_conflict_degree(PassD ps, const int *tcnt, int *deg) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ps.nc; i += gridDim.x * blockDim.x) {
    int d = -1;
    if (ps.cstatus[i] == ST_VALID) {
      const int *L = ps.ws + (size_t)i * ps.R;
      int k = ps.cn[i] + ps.cr[i];
      d = 0;
      for (int q = 0; q < k; q++) d += tcnt[L[q]] - 1;
    }
    deg[i] = d;
  }
}

// ===========================================================================
// Graph-driven round loop: fixed-grid ("persistent") kernels whose work
// counts, epochs and loop decisions live in the device control block Ctl, so
// a block of rounds is captured once as a hipGraph and replayed with a single
// host synchronisation per launch.  Every decision is the host loop's
// (Refiner::run), made on the device by single-lane gate / finish kernels.
// ===========================================================================
constexpr int GC = 1024;  // blocks of the fixed-grid compaction

__device__ __forceinline__ bool ctl_on(const Ctl *c, int gidx) {
  return c->i[C_GO] && (gidx < 0 || c->i[gidx]);
}
__device__ __forceinline__ int ctl_cands_per_pass(const Ctl *c, int K) {
  long long v = (long long)(c->acc[A_BUDGET] / (6ull * (unsigned long long)K + 16ull));
  if (v < 1) v = 1;
  if (v > (1LL << 30)) v = 1LL << 30;
  return (int)v;
}

// Tiles of CT items, split into contiguous per-block ranges: block b scans
// tiles [b*tpb, (b+1)*tpb), so the scatter keeps ascending index order.
template <class Pred>
__global__ __launch_bounds__(CB) void k_dcount(Pred pred, const Ctl *c, int nidx,
                                               int gidx, int *blockcnt) {
  __shared__ int wsum[CB / 64];
  int n = ctl_on(c, gidx) ? c->i[nidx] : 0;
  int ntiles = (n + CT - 1) / CT;
  int tpb = (ntiles + gridDim.x - 1) / gridDim.x;
  int tb0 = blockIdx.x * tpb, tb1 = min(tb0 + tpb, ntiles);
  int cnt = 0;
  for (int tile = tb0; tile < tb1; tile++)
    for (int k = 0; k < CI; k++) {
      int i = tile * CT + k * CB + threadIdx.x;
      if (i < n && pred(i)) cnt++;
    }
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int sum = 0;
    for (int w = 0; w < CB / 64; w++) sum += wsum[w];
    blockcnt[blockIdx.x] = sum;
  }
}

// Exclusive scan of the GC block counts; the finisher sees the total.
template <class Fin>
__global__ __launch_bounds__(1024) void k_dscan(const int *blockcnt, int *blockoff,
                                                Ctl *c, int gidx, Fin fin) {
  __shared__ int sh[GC];
  if (!ctl_on(c, gidx)) return;
  int v = blockcnt[threadIdx.x];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < GC; off <<= 1) {
    int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  blockoff[threadIdx.x] = sh[threadIdx.x] - v;
  int total = sh[GC - 1];
  __syncthreads();
  fin(c, total, (int)threadIdx.x);
}

template <class Pred>
__global__ __launch_bounds__(CB) void k_dscatter(Pred pred, const Ctl *c, int nidx,
                                                 int gidx, const int *blockoff,
                                                 int *out_list, int *out_rank,
                                                 const int *src) {
  __shared__ int wcnt[CB / 64];
  int n = ctl_on(c, gidx) ? c->i[nidx] : 0;
  int ntiles = (n + CT - 1) / CT;
  int tpb = (ntiles + gridDim.x - 1) / gridDim.x;
  int tb0 = blockIdx.x * tpb, tb1 = min(tb0 + tpb, ntiles);
  if (tb0 >= tb1) return;
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int run = blockoff[blockIdx.x];
  unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int tile = tb0; tile < tb1; tile++)
    for (int k = 0; k < CI; k++) {
      int i = tile * CT + k * CB + threadIdx.x;
      bool f = (i < n) && pred(i);
      unsigned long long b = __ballot(f);
      if (lane == 0) wcnt[wave] = __popcll(b);
      __syncthreads();
      int pre = 0, tot = 0;
      for (int w = 0; w < CB / 64; w++) {
        int cw = wcnt[w];
        if (w < wave) pre += cw;
        tot += cw;
      }
      if (f) {
        int pos = run + pre + __popcll(b & lt);
        if (out_list) out_list[pos] = src ? src[i] : i;
        if (out_rank) out_rank[i] = pos;
      }
      run += tot;
      __syncthreads();
    }
}

// Priority predicates whose cut comes from the control block.
struct PredPrioEqC {
  const int *prio;
  const Ctl *c;
  __device__ bool operator()(int i) const { return prio[i] == c->i[C_BCUT]; }
};
struct PredPrioSelC {
  const int *prio;
  const Ctl *c;
  __device__ bool operator()(int i) const {
    int p = prio[i], b = c->i[C_BCUT];
    return p < b || (p == b && i <= c->i[C_ICUT]);
  }
};

__device__ __forceinline__ bool ctl_fits(const Ctl *c, int nc, bool seg) {
  return (long long)c->i[C_NP] + nc <= c->i[C_CAPP] &&
         (long long)c->i[C_NT] + 2LL * nc <= c->i[C_CAPT] &&
         (!seg || (long long)c->i[C_NS] + nc <= c->i[C_CAPS]);
}

// Segment phase: encroached subsegments counted (refine.cu:3180-3778 loop).
struct FinSeg {
  __device__ void operator()(Ctl *c, int total, int tid) const {
    if (tid != 0) return;
    c->i[C_NE] = total;
    c->i[C_USESEL] = 0;
    if (total == 0) {
      c->i[C_SEG_DONE] = 1;
      c->i[C_GO] = 0;
      return;
    }
    int nc = min(total, ctl_cands_per_pass(c, c->i[C_K]));
    c->i[C_NC] = nc;
    if (!ctl_fits(c, nc, true)) {
      c->i[C_SUSPEND] = 1;
      c->i[C_GO] = 0;
    }
  }
};

// Bad phase: bad triangles counted; candidate cap and selection decided.
struct FinBad {
  unsigned *hist;
  __device__ void operator()(Ctl *c, int total, int tid) const {
    int nb = total;
    long long cap = ctl_cands_per_pass(c, c->i[C_K]);
    int prev = c->i[C_PREVWIN];
    if (c->i[C_CANDFACTOR] > 0 && prev >= 0) {
      long long f = (long long)c->i[C_CANDFACTOR] * prev;
      long long lo = c->i[C_CANDMIN];
      cap = min(cap, f > lo ? f : lo);
    }
    int prio = c->i[C_PRIORITY];
    bool sel = nb > cap && (prio == 1 || prio == 2);
    int nc = nb > cap ? (int)cap : nb;
    bool fits = ctl_fits(c, nc, false);
    if (nb > 0 && fits && sel)
      for (int b = tid; b < NBUCKET; b += 1024) hist[b] = 0u;
    if (tid != 0) return;
    c->i[C_NB] = nb;
    c->i[C_USESEL] = 0;
    c->i[C_SEL] = 0;
    if (nb == 0) {
      c->i[C_DONE] = 1;
      c->i[C_STATUS] = 0;
      c->i[C_GO] = 0;
      return;
    }
    c->i[C_NC] = nc;
    if (!fits) {
      c->i[C_SUSPEND] = 1;
      c->i[C_GO] = 0;
      return;
    }
    c->i[C_SEL] = sel ? 1 : 0;
  }
};

struct FinNone {
  __device__ void operator()(Ctl *, int, int) const {}
};
struct FinSel {
  __device__ void operator()(Ctl *c, int total, int tid) const {
    if (tid != 0) return;
    c->i[C_NC] = total;
    c->i[C_USESEL] = 1;
  }
};
struct FinWin {
  __device__ void operator()(Ctl *c, int total, int tid) const {
    if (tid == 0) c->i[C_NWIN] = total;
  }
};
__global__ void k_gate_seg(Ctl *c) {
  int go = !c->i[C_DONE] && !c->i[C_SUSPEND] && !c->i[C_SEG_DONE] &&
           c->i[C_K] == c->i[C_KCAP];
  if (go && !c->i[C_ROUND_OPEN]) {
    if (c->i[C_ROUNDS] >= c->i[C_MAXROUNDS]) {
      c->i[C_DONE] = 1;  // round cap: status stays 1
      go = 0;
    } else {
      c->i[C_ROUND_OPEN] = 1;
      c->i[C_ROUNDS]++;
      c->i[C_NP_ROUND] = c->i[C_NP];
    }
  }
  c->i[C_GO] = go;
  if (go)
    for (int k = 0; k < CNT_N; k++) c->cnt[k] = 0ull;
}

__global__ void k_gate_bad(Ctl *c) {
  int go = !c->i[C_DONE] && !c->i[C_SUSPEND] && c->i[C_SEG_DONE] &&
           c->i[C_K] == c->i[C_KCAP];
  c->i[C_GO] = go;
  c->i[C_SEL] = 0;
  if (go)
    for (int k = 0; k < CNT_N; k++) c->cnt[k] = 0ull;
}

__device__ void ctl_accumulate(Ctl *c, int nwin) {
  const unsigned long long *n = c->cnt;
  c->acc[A_REJ_OVF] += n[CNT_OVERFLOW];
  c->acc[A_REJ_ENC] += n[CNT_REJ_ENC];
  c->acc[A_REJ_BLK] += n[CNT_REJ_BLOCKED];
  c->acc[A_GIVEUPS] += n[CNT_GIVEUPS];
  c->acc[A_LOST] += n[CNT_VALID] - (unsigned long long)nwin;
  c->acc[A_FP32] += n[CNT_FP32];
  c->acc[A_FP64] += n[CNT_FP64];
  c->acc[A_EXACT] += n[CNT_EXACT];
  c->acc[A_DD] += n[CNT_DD];
  c->acc[A_BIG] += n[CNT_BIG];
  c->acc[A_EXAM_SEG] += n[CNT_EXAM_SEG];
  c->acc[A_EXAM_BAD] += n[CNT_EXAM_BAD];
  c->acc[A_WS_SEG] += n[CNT_WS_SEG];
  c->acc[A_WS_BAD] += n[CNT_WS_BAD];
  c->acc[A_DELETED] += n[CNT_DEL];
  c->acc[A_FLIPS] += n[CNT_FLIPS];
}

__global__ void k_end_seg(Ctl *c, int iters, int eskip) {
  if (!c->i[C_GO]) return;
  int nwin = c->i[C_NWIN];
  ctl_accumulate(c, nwin);
  c->acc[A_SEGSPLITS] += nwin;
  c->i[C_NP] += nwin;
  c->i[C_NT] += 2 * nwin;
  c->i[C_NS] += nwin;
  c->i[C_SEG_PASSES]++;
  // the deletion step's epoch, claim epochs, the winners' stamp
  c->i[C_EPOCH] += eskip + iters + 1;
  int K = c->i[C_K];
  if (c->cnt[CNT_OVERFLOW] > 0 && K < c->i[C_KMAX]) c->i[C_K] = min(2 * K, c->i[C_KMAX]);
}

__global__ void k_end_bad(Ctl *c, int iters) {
  if (!c->i[C_GO]) return;
  int nwin = c->i[C_NWIN];
  ctl_accumulate(c, nwin);
  c->acc[A_INSERTED] += nwin;
  c->acc[A_CANDEVALS] += c->i[C_NC];
  c->i[C_NP] += nwin;
  c->i[C_NT] += 2 * nwin;
  c->i[C_PREVWIN] = nwin;
  c->i[C_PASS_INDEX]++;
  c->i[C_BAD_PASSES]++;
  c->i[C_EPOCH] += iters + 1;  // claim epochs + the winners' stamp
  long long prog = nwin + (long long)c->cnt[CNT_MARKS] + (long long)c->cnt[CNT_GIVEUPS];
  int K = c->i[C_K];
  if (c->cnt[CNT_OVERFLOW] > 0 && K < c->i[C_KMAX]) {
    c->i[C_K] = min(2 * K, c->i[C_KMAX]);
    prog++;
  }
  if (prog == 0) {
    c->i[C_DONE] = 1;
    c->i[C_STATUS] = 2;
  }
  c->i[C_ROUND_OPEN] = 0;
  c->i[C_SEG_DONE] = 0;
}

// Priority histogram of the nb bad triangles (selection pass only).
__global__ __launch_bounds__(256) void k_dprio_hist(MeshD m, const int *ids,
                                                    const Ctl *c, int *prio,
                                                    unsigned *hist) {
  __shared__ unsigned h[NBUCKET];
  if (!ctl_on(c, C_SEL)) return;
  int n = c->i[C_NB];
  for (int b = threadIdx.x; b < NBUCKET; b += blockDim.x) h[b] = 0;
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    int p = m.tbucket[ids[i]];
    prio[i] = p;
    atomicAdd(&h[p], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < NBUCKET; b += blockDim.x)
    if (h[b]) atomicAdd(&hist[b], h[b]);
}

// Cut bucket: the bucket holding the cap-th smallest priority, and how many
// of its members are taken (one block of 1024 lanes, 8 buckets each).
__global__ __launch_bounds__(1024) void k_dcut(const unsigned *hist, Ctl *c) {
  __shared__ long long sh[1024];
  if (!ctl_on(c, C_SEL)) return;
  const int per = NBUCKET / 1024;
  long long cap = c->i[C_NC];
  long long mine = 0;
  for (int k = 0; k < per; k++) mine += hist[threadIdx.x * per + k];
  sh[threadIdx.x] = mine;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    long long t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  long long before = sh[threadIdx.x] - mine;
  if (before < cap && sh[threadIdx.x] >= cap) {
    long long acc = before;
    for (int k = 0; k < per; k++) {
      int b = threadIdx.x * per + k;
      if (acc + hist[b] >= cap) {
        c->i[C_BCUT] = b;
        c->i[C_TAKE] = (int)(cap - acc);
        break;
      }
      acc += hist[b];
    }
  }
}

__global__ void k_dicut(const int *sel_tmp, Ctl *c) {
  if (!ctl_on(c, C_SEL)) return;
  c->i[C_ICUT] = sel_tmp[c->i[C_TAKE] - 1];
}

// ===========================================================================
// Host side
// ===========================================================================
static void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
#define HC(x) hip_check((x), #x)

template <class T>
static void dalloc(T **p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  HC(hipMalloc((void **)p, n * sizeof(T)));
}
template <class T>
static void dfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}
template <class T>
static void dgrow(T **p, size_t oldn, size_t newn, hipStream_t s) {
  T *q;
  dalloc(&q, newn);
  if (*p && oldn) HC(hipMemcpyAsync(q, *p, oldn * sizeof(T), hipMemcpyDeviceToDevice, s));
  HC(hipStreamSynchronize(s));
  dfree(*p);
  *p = q;
}

static int nblk(long long n, int b = 256) { return (int)((n + b - 1) / b); }

// Largest mesh the 32-bit encodings hold: triangle ids travel as
// otri = tri << 2 | edge in an int, point and subsegment ids as ints.
constexpr long long MAX_TRIANGLES = (1LL << 29) - 1;
constexpr long long MAX_POINTS = (1LL << 31) - 1;

Refiner::Refiner(const gcdt_options &o) : opt(o) {
  HC(hipSetDevice(opt.device));
  HC(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  HC(hipHostMalloc((void **)&h_scalar, 64 * sizeof(int), hipHostMallocDefault));
  // compaction totals are written here by the scan kernel
  HC(hipHostMalloc((void **)&h_total, 64 * sizeof(int), hipHostMallocCoherent | hipHostMallocMapped));
  HC(hipHostGetDevicePointer((void **)&d_host_total, h_total, 0));
  // pass counters: written here by k_flush_counters
  HC(hipHostMalloc((void **)&h_cnt, CNT_N * sizeof(unsigned long long),
                   hipHostMallocCoherent | hipHostMallocMapped));
  HC(hipHostGetDevicePointer((void **)&d_host_cnt, h_cnt, 0));
  HC(hipEventCreate(&ev0));
  HC(hipEventCreate(&ev1));
  // cavity and apply kernels take up to BIG_LDS_BYTES of dynamic LDS
  const void *cav[] = {
      (const void *)k_cavity_fast<true, 8>,  (const void *)k_cavity_fast<false, 8>,
      (const void *)k_cavity_fast<true, 16>, (const void *)k_cavity_fast<false, 16>,
      (const void *)k_cavity_fast<true, 32>, (const void *)k_cavity_fast<false, 32>,
      (const void *)k_cavity_fast<true, 64>, (const void *)k_cavity_fast<false, 64>,
      (const void *)k_cavity_slow<true, 8>,  (const void *)k_cavity_slow<false, 8>,
      (const void *)k_cavity_slow<true, 16>, (const void *)k_cavity_slow<false, 16>,
      (const void *)k_cavity_slow<true, 32>, (const void *)k_cavity_slow<false, 32>,
      (const void *)k_cavity_slow<true, 64>, (const void *)k_cavity_slow<false, 64>};
  for (const void *f : cav)
    HC(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, BIG_LDS_BYTES));

  dalloc(&d_total, 4);
  dalloc(&d_cnt, CNT_N);
  dalloc(&d_ctl, 1);
  dalloc(&g_cnt, GC);
  dalloc(&g_off, GC);
  HC(hipHostMalloc((void **)&h_ctl, sizeof(Ctl), hipHostMallocDefault));
  dalloc(&d_val, 4);
  HC(hipHostMalloc((void **)&h_val4, 4 * sizeof(unsigned long long), hipHostMallocDefault));
  for (int k = 0; k < NSTAGE; k++)
    HC(hipEventCreateWithFlags(&ev_stage[k], hipEventDisableTiming));
  set_angle(opt.theta_deg, opt.offcenter, opt.mode);
  pr.use_fp32 = opt.pred_mode == 1;
  pr.priority = opt.priority;
  pr.ipar = nullptr;
}

// Quality threshold, off-center constant and encroachment rule of theta
// (refine.cu:6557-6567 with the constant kept in double, as Triangle does,
// triangle.cpp:3527-3533).
void Refiner::set_angle(double theta_deg, int offcenter, int mode) {
  opt.theta_deg = theta_deg;
  opt.offcenter = offcenter;
  opt.mode = mode;
  double ga = std::cos(theta_deg * 3.141592653589793238462643383279502884197169399375105820974944592308 / 180.0);
  pr.offconst = (ga == 1.0) ? 0.0 : 0.475 * std::sqrt((1.0 + ga) / (1.0 - ga));
  pr.goodangle = ga * ga;
  pr.lens = (2.0 * pr.goodangle - 1.0) * (2.0 * pr.goodangle - 1.0);
  pr.offcenter = offcenter;
  pr.ruppert = mode == 1;
}

Refiner::~Refiner() {
  (void)hipStreamSynchronize(stream);
  if (use_vm) {  // owned by the reserved ranges (released by their destructors)
    tmp_list = nullptr;
    sel_prio = sel_tmp = sel_list = nullptr;
  }
  free_mesh();
  for (auto &in : inputs) {
    dfree(in.P); dfree(in.T); dfree(in.N); dfree(in.S); dfree(in.tinst); dfree(in.ipar);
  }
  dfree(seg_in);
  dfree(c_pt); dfree(c_status); dfree(c_n); dfree(c_r);
  dfree(c_prio); dfree(c_win); dfree(c_rank); dfree(c_ws); dfree(c_pws);
  dfree(c_act); dfree(c_slow); dfree(tmp_list); dfree(mix_list);
  dfree(sel_prio); dfree(sel_tmp); dfree(sel_list); dfree(d_hist);
  dfree(c_feat); dfree(c_score); dfree(c_deg); dfree(t_cnt); dfree(mlp_params);
  dfree(cb_cnt); dfree(cb_off); dfree(d_total); dfree(d_cnt);
  dfree(o_T); dfree(o_N); dfree(o_tstat); dfree(o_pstat); dfree(newid);
  dfree(h_key); dfree(h_val);
  dfree(d_ctl); dfree(g_cnt); dfree(g_off);
  dfree(o_inst);
  dfree(del_ws); dfree(del_dv); dfree(del_dk);
  dfree(p_new); dfree(p_tmp); dfree(pf_tmp); dfree(pp_tmp);
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  if (h_ctl) (void)hipHostFree(h_ctl);
  if (h_scalar) (void)hipHostFree(h_scalar);
  if (h_total) (void)hipHostFree(h_total);
  if (h_cnt) (void)hipHostFree(h_cnt);
  dfree(d_val);
  if (h_val4) (void)hipHostFree(h_val4);
  delete copier;
  for (int k = 0; k < NSTAGE; k++) {
    if (h_stage[k]) (void)hipHostFree(h_stage[k]);
    (void)hipEventDestroy(ev_stage[k]);
  }
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  for (auto &kt : ktimes) {
    for (auto &e : kt.pending) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
  }
  if (stream) (void)hipStreamDestroy(stream);
}

void Refiner::free_mesh() {
  if (use_vm) return;  // the reserved ranges live as long as the refiner
  dfree(m.P); dfree(m.pflag); dfree(m.ppar);
  dfree(m.T); dfree(m.tflag); dfree(m.tbucket); dfree(m.claim);
  dfree(m.SV); dfree(m.SP); dfree(m.ST); dfree(m.senc);
  capP = capT = capS = 0;
}

// Reserves the ranges of every growable array at its largest size (the
// 32-bit encodings' limits); false, with nothing held, when the device or the
// runtime cannot map memory into reserved ranges.
bool Refiner::vm_setup() {
  const int dev = opt.device;
  if (!VmRange::supported(dev)) return false;
  const size_t mp = (size_t)MAX_POINTS, mt = (size_t)MAX_TRIANGLES;
  const bool ok =
      vm_P.reserve(dev, mp * sizeof(double2)) && vm_pflag.reserve(dev, mp) &&
      vm_ppar.reserve(dev, mp * sizeof(int)) && vm_T.reserve(dev, mt * sizeof(TriRec)) &&
      vm_tflag.reserve(dev, mt) && vm_tbucket.reserve(dev, mt * sizeof(uint16_t)) &&
      vm_claim.reserve(dev, 2 * mt * sizeof(unsigned long long)) &&
      vm_SV.reserve(dev, mp * sizeof(int2)) && vm_SP.reserve(dev, mp * sizeof(int)) &&
      vm_ST.reserve(dev, mp * sizeof(int)) && vm_senc.reserve(dev, mp * sizeof(int)) &&
      vm_tmp.reserve(dev, mp * sizeof(int)) && vm_sel_prio.reserve(dev, mt * sizeof(int)) &&
      vm_sel_tmp.reserve(dev, mt * sizeof(int)) && vm_sel_list.reserve(dev, mt * sizeof(int));
  if (!ok) {
    vm_release_all();
    return false;
  }
  m.P = (double2 *)vm_P.ptr();
  m.pflag = (uint8_t *)vm_pflag.ptr();
  m.ppar = (int *)vm_ppar.ptr();
  m.T = (TriRec *)vm_T.ptr();
  m.tflag = (uint8_t *)vm_tflag.ptr();
  m.tbucket = (uint16_t *)vm_tbucket.ptr();
  m.claim = (unsigned long long *)vm_claim.ptr();
  m.SV = (int2 *)vm_SV.ptr();
  m.SP = (int *)vm_SP.ptr();
  m.ST = (int *)vm_ST.ptr();
  m.senc = (int *)vm_senc.ptr();
  return true;
}

void Refiner::vm_release_all() {
  for (VmRange *r : {&vm_P, &vm_pflag, &vm_ppar, &vm_T, &vm_tflag, &vm_tbucket, &vm_claim, &vm_SV,
                     &vm_SP, &vm_ST, &vm_senc, &vm_tmp, &vm_sel_prio, &vm_sel_tmp, &vm_sel_list})
    r->release();
  m.P = nullptr; m.pflag = nullptr; m.ppar = nullptr; m.T = nullptr; m.tflag = nullptr;
  m.tbucket = nullptr; m.claim = nullptr; m.SV = nullptr; m.SP = nullptr; m.ST = nullptr;
  m.senc = nullptr;
  capP = capT = capS = 0;
}

// Maps the mesh arrays up to the given capacities (in place; new claim
// words start at zero).
void Refiner::vm_map_mesh(long long cp, long long ct, long long cs) {
  if (cp > capP) {
    vm_P.map_to((size_t)cp * sizeof(double2));
    vm_pflag.map_to((size_t)cp);
    vm_ppar.map_to((size_t)cp * sizeof(int));
    capP = cp;
  }
  if (ct > capT) {
    vm_T.map_to((size_t)ct * sizeof(TriRec));
    vm_tflag.map_to((size_t)ct);
    vm_tbucket.map_to((size_t)ct * sizeof(uint16_t));
    vm_claim.map_to((size_t)ct * 2 * sizeof(unsigned long long));
    HC(hipMemsetAsync(m.claim + 2 * (size_t)capT, 0,
                      sizeof(unsigned long long) * 2 * (size_t)(ct - capT), stream));
    capT = ct;
  }
  if (cs > capS) {
    vm_SV.map_to((size_t)cs * sizeof(int2));
    vm_SP.map_to((size_t)cs * sizeof(int));
    vm_ST.map_to((size_t)cs * sizeof(int));
    vm_senc.map_to((size_t)cs * sizeof(int));
    capS = cs;
  }
}

void Refiner::upload(const double *points, int np, const int *tris,
                     const int *nbrs, int nt, const int *segs, int ns) {
  // a single resident input whose buffers are large enough is overwritten
  // in place (a process that refines many CDTs allocates once)
  if (inputs.size() == 1 && !inputs[0].tinst && inputs[0].capP >= np && inputs[0].capT >= nt &&
      inputs[0].capS >= ns) {
    InputSet &in = inputs[0];
    copy_input(in, points, np, tris, nbrs, nt, segs, ns, nullptr, nullptr, 0);
    select_input(0);
    return;
  }
  for (auto &in : inputs) {
    dfree(in.P); dfree(in.T); dfree(in.N); dfree(in.S); dfree(in.tinst); dfree(in.ipar);
  }
  inputs.clear();
  add_input(points, np, tris, nbrs, nt, segs, ns, nullptr, nullptr, 0);
  select_input(0);
}

// Every index the init kernels dereference, counted on the device: vertex
// ids of triangles, neighbour ids, segment vertex ids, instance ids.
__global__ void k_validate_input(const int *T, const int *N, int nt, const int *S, int ns, int np,
                                 const int *tinst, int n_inst, unsigned long long *bad) {
  unsigned long long b0 = 0, b1 = 0, b2 = 0, b3 = 0;
  const long long n = max(3LL * nt, 2LL * ns);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    if (i < 3LL * nt) {
      const int a = T[i], u = N[i];
      b0 += a < 0 || a >= np;
      b1 += u < -1 || u >= nt;
      if (tinst && i < nt) b3 += tinst[i] < 0 || tinst[i] >= n_inst;
    }
    if (i < 2LL * ns) b2 += S[i] < 0 || S[i] >= np;
  }
  if (b0) atomicAdd(&bad[0], b0);
  if (b1) atomicAdd(&bad[1], b1);
  if (b2) atomicAdd(&bad[2], b2);
  if (b3) atomicAdd(&bad[3], b3);
}

// Pinned staging buffers (shared by upload and download) and the host copy
// threads.
void Refiner::ensure_stage() {
  if (!copier) copier = new CopyPool(CopyPool::default_workers());
  if (!h_stage[0]) {
    stage_bytes = size_t(32) << 20;
    for (int k = 0; k < NSTAGE; k++)
      HC(hipHostMalloc((void **)&h_stage[k], stage_bytes, hipHostMallocDefault));
  }
}

// Host-to-device through the staging buffers: the host threads fill one
// while the copy engine drains the others (a pageable copy is staged by
// the runtime at about half the pinned rate).  k counts the chunks issued
// so far; a buffer is refilled only after its previous transfer ended.
void Refiner::upload_staged(void *dst, const void *src, size_t bytes, size_t &k) {
  for (size_t off = 0; off < bytes; off += stage_bytes, k++) {
    const size_t b = std::min(stage_bytes, bytes - off);
    const int s = (int)(k % NSTAGE);
    if (k >= (size_t)NSTAGE) HC(hipEventSynchronize(ev_stage[s]));
    copier->copy(h_stage[s], (const char *)src + off, b);
    HC(hipMemcpyAsync((char *)dst + off, h_stage[s], b, hipMemcpyHostToDevice, stream));
    HC(hipEventRecord(ev_stage[s], stream));
  }
}

// Copies one input CDT into `in` (buffers large enough) and validates its
// indices on the device before any kernel gathers through them.
void Refiner::copy_input(InputSet &in, const double *points, int np, const int *tris,
                         const int *nbrs, int nt, const int *segs, int ns, const int *tri_inst,
                         const InstPar *ipar, int n_inst) {
  in.np = np; in.nt = nt; in.ns = ns;
  in.n_inst = tri_inst ? n_inst : 0;
  in.any_chew = false;
  if (tri_inst)
    for (int i = 0; i < n_inst; i++) in.any_chew |= ipar[i].ruppert == 0;
  ensure_stage();
  HC(hipStreamSynchronize(stream));  // the staging buffers are free
  size_t k = 0;
  upload_staged(in.P, points, sizeof(double2) * np, k);
  upload_staged(in.T, tris, sizeof(int) * 3 * (size_t)nt, k);
  upload_staged(in.N, nbrs, sizeof(int) * 3 * (size_t)nt, k);
  if (ns) upload_staged(in.S, segs, sizeof(int) * 2 * (size_t)ns, k);
  if (tri_inst) {
    upload_staged(in.tinst, tri_inst, sizeof(int) * (size_t)nt, k);
    upload_staged(in.ipar, ipar, sizeof(InstPar) * (size_t)n_inst, k);
  }
  HC(hipMemsetAsync(d_val, 0, 4 * sizeof(unsigned long long), stream));
  k_validate_input<<<std::min(nblk(std::max(3LL * nt, 2LL * ns)), 4096), 256, 0, stream>>>(
      in.T, in.N, nt, in.S, ns, np, tri_inst ? in.tinst : nullptr, n_inst, d_val);
  HC(hipGetLastError());
  HC(hipMemcpyAsync(h_val4, d_val, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
  HC(hipStreamSynchronize(stream));
  if (h_val4[0]) throw std::runtime_error("triangle vertex index out of range");
  if (h_val4[1]) throw std::runtime_error("neighbour index out of range");
  if (h_val4[2]) throw std::runtime_error("segment vertex index out of range");
  if (h_val4[3]) throw std::runtime_error("triangle instance index out of range");
}

int Refiner::add_input(const double *points, int np, const int *tris,
                       const int *nbrs, int nt, const int *segs, int ns,
                       const int *tri_inst, const InstPar *ipar, int n_inst) {
  if (np < 3 || nt < 1) throw std::runtime_error("input mesh is empty");
  if ((long long)nt * 4 >= (1LL << 31) / 2)
    throw std::runtime_error("too many triangles for 32-bit otri encoding");
  if (ns < 0 || (tri_inst && n_inst < 1)) throw std::runtime_error("invalid input counts");
  InputSet in;
  in.capP = np; in.capT = nt; in.capS = ns;
  dalloc(&in.P, (size_t)np);
  dalloc(&in.T, 3 * (size_t)nt);
  dalloc(&in.N, 3 * (size_t)nt);
  dalloc(&in.S, 2 * (size_t)(ns > 0 ? ns : 1));
  if (tri_inst) {
    dalloc(&in.tinst, (size_t)nt);
    dalloc(&in.ipar, (size_t)n_inst);
  }
  try {
    copy_input(in, points, np, tris, nbrs, nt, segs, ns, tri_inst, ipar, n_inst);
  } catch (...) {
    dfree(in.P); dfree(in.T); dfree(in.N); dfree(in.S); dfree(in.tinst); dfree(in.ipar);
    throw;
  }
  inputs.push_back(in);
  return (int)inputs.size() - 1;
}

void Refiner::select_input(int k) {
  if (k < 0 || k >= (int)inputs.size()) throw std::runtime_error("no such input");
  const InputSet &in = inputs[k];
  np0 = in.np; nt0 = in.nt; ns0 = in.ns;
  in_P = in.P; in_T = in.T; in_N = in.N; in_S = in.S;
  in_tinst = in.tinst;
  d_ipar = in.ipar;
  n_inst = in.n_inst;
  in_any_chew = in.any_chew;
}


void Refiner::ensure_capacity(long long needP, long long needT, long long needS) {
  if (needT > MAX_TRIANGLES || needP > MAX_POINTS || needS > MAX_POINTS)
    throw std::runtime_error("the refined mesh would exceed 2^29 triangles / 2^31 points "
                             "(32-bit encodings); the refinement does not converge on this input");
  if (needP > capP || needT > capT || needS > capS) st.growths++;
  if (use_vm) {  // in place: map more, nothing moves
    vm_map_mesh(needP > capP ? std::min(std::max<long long>(needP, capP + capP / 2), MAX_POINTS) : capP,
                needT > capT ? std::min(std::max<long long>(needT, capT + capT / 2), MAX_TRIANGLES) : capT,
                needS > capS ? std::min(std::max<long long>(needS, capS + capS / 2), MAX_POINTS) : capS);
    return;
  }
  if (needP > capP) {
    long long c = std::min(std::max<long long>(needP, capP + capP / 2), MAX_POINTS);
    dgrow(&m.P, (size_t)np, (size_t)c, stream);
    dgrow(&m.pflag, (size_t)np, (size_t)c, stream);
    dgrow(&m.ppar, (size_t)np, (size_t)c, stream);
    capP = c;
  }
  if (needT > capT) {
    long long c = std::min(std::max<long long>(needT, capT + capT / 2), MAX_TRIANGLES);
    dgrow(&m.T, (size_t)nt, (size_t)c, stream);
    dgrow(&m.tflag, (size_t)nt, (size_t)c, stream);
    dgrow(&m.tbucket, (size_t)nt, (size_t)c, stream);
    unsigned long long *nc;
    dalloc(&nc, 2 * (size_t)c);
    HC(hipMemsetAsync(nc, 0, sizeof(unsigned long long) * 2 * c, stream));
    if (nt) HC(hipMemcpyAsync(nc, m.claim, sizeof(unsigned long long) * 2 * nt, hipMemcpyDeviceToDevice, stream));
    HC(hipStreamSynchronize(stream));
    dfree(m.claim);
    m.claim = nc;
    capT = c;
  }
  if (needS > capS) {
    long long c = std::min(std::max<long long>(needS, capS + capS / 2), MAX_POINTS);
    dgrow(&m.SV, (size_t)ns, (size_t)c, stream);
    dgrow(&m.SP, (size_t)ns, (size_t)c, stream);
    dgrow(&m.ST, (size_t)ns, (size_t)c, stream);
    dgrow(&m.senc, (size_t)ns, (size_t)c, stream);
    capS = c;
  }
}

// Kernel-time bookkeeping (options.profile): one event pair per launch.
void Refiner::tbegin(int id) {
  if (!opt.profile || capturing) return;
  hipEvent_t a, b;
  HC(hipEventCreate(&a));
  HC(hipEventCreate(&b));
  HC(hipEventRecord(a, stream));
  ktimes[id].pending.push_back({a, b});
}
void Refiner::tend(int id, long long items) {
  if (!opt.profile || capturing) return;
  HC(hipEventRecord(ktimes[id].pending.back().second, stream));
  ktimes[id].launches++;
  ktimes[id].items += items;
}
void Refiner::collect_times() {
  for (auto &kt : ktimes) {
    for (auto &e : kt.pending) {
      float ms = 0.f;
      HC(hipEventSynchronize(e.second));
      HC(hipEventElapsedTime(&ms, e.first, e.second));
      kt.total_ms += ms;
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
    kt.pending.clear();
  }
}

const char *Refiner::kname(int id) {
  static const char *names[] = {"init",       "mark_bad", "compact",
                                "seg_cavity", "bad_cavity", "claim",
                                "win",        "apply",      "export",
                                "mlp",        "del_scan",   "del_apply"};
  return names[id];
}

// Stream compaction of [0, n) by pred, in ascending index order: one launch
// when n fits one tile, else per-tile counts then a scatter that sums the
// counts before its tile (counted: the counts are already in cb_cnt).  The
// summing costs O(tiles^2) reads in all, so beyond SUM_TILES tiles a scan
// launch computes the tile offsets instead.  With
// total64 the count stays on the device and -1 is returned; otherwise the
// call synchronises and returns it.
template <class Pred>
int Refiner::compact(Pred pred, int n, int *out_list, int *out_rank,
                     const int *src, unsigned long long *total64, bool counted) {
  if (n <= 0) {
    if (total64) HC(hipMemsetAsync(total64, 0, sizeof(unsigned long long), stream));
    return 0;
  }
  int nb = (n + CT - 1) / CT;
  ensure_cb(nb);
  tbegin(K_COMPACT);
  if (nb == 1) {
    k_compact1<Pred><<<1, CB, 0, stream>>>(pred, n, out_list, out_rank, src, d_total,
                                           total64 ? nullptr : d_host_total, total64);
  } else {
    if (!counted) k_count<Pred><<<nb, CB, 0, stream>>>(pred, n, cb_cnt);
    if (nb <= SUM_TILES) {
      k_scatter_sum<Pred><<<nb, CB, 0, stream>>>(pred, n, cb_cnt, out_list, out_rank, src,
                                                 d_total, total64 ? nullptr : d_host_total,
                                                 total64);
    } else {
      k_scan<<<1, 1024, 0, stream>>>(cb_cnt, nb, cb_off, d_total,
                                      total64 ? nullptr : d_host_total, total64);
      k_scatter<Pred><<<nb, CB, 0, stream>>>(pred, n, cb_off, out_list, out_rank, src);
    }
  }
  tend(K_COMPACT, n);
  HC(hipGetLastError());
  if (total64) return -1;  // the count is read later with the pass counters
  HC(hipStreamSynchronize(stream));
  return __atomic_load_n(h_total, __ATOMIC_ACQUIRE);
}

// Per-tile counts of the compaction (kept across calls: a counting pass may
// fill them before compact() runs).
void Refiner::ensure_cb(int nb) {
  if (nb <= cb_cap) return;
  dfree(cb_cnt);
  dfree(cb_off);
  cb_cap = std::max(nb, 2 * cb_cap);
  dalloc(&cb_cnt, (size_t)cb_cap);
  dalloc(&cb_off, (size_t)cb_cap);
}

void Refiner::read_counters() {
  HC(hipMemcpyAsync(h_cnt, d_cnt, CNT_N * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
  HC(hipStreamSynchronize(stream));
  cnt_zero = false;
}

// End of a pass: the counters to the host's mapped copy (read after the
// stream synchronisation) and back to zero for the next pass -- one launch
// in place of a copy and a fill.
__global__ void k_flush_counters(unsigned long long *cnt, unsigned long long *host) {
  const int k = threadIdx.x;
  if (k < CNT_N) {
    __hip_atomic_store(host + k, cnt[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    cnt[k] = 0ull;
  }
  __threadfence_system();
}

void Refiner::flush_counters() {
  k_flush_counters<<<1, 64, 0, stream>>>(d_cnt, d_host_cnt);
  HC(hipGetLastError());
  HC(hipStreamSynchronize(stream));
  cnt_zero = true;
}

// Candidate workspace of at least `ints` ints (grown by half again at least,
// capped by the budget; passes never need more than the budget).
void Refiner::ensure_ws(long long ints) {
  ints = std::min(ints, opt.cand_budget);
  if (c_ws && ws_ints >= ints) return;
  long long c = std::min(opt.cand_budget, std::max(ints, ws_ints + ws_ints / 2));
  dfree(c_ws);
  dalloc(&c_ws, (size_t)std::max(c, 1LL));
  ws_ints = c;
}

void Refiner::ensure_cand(int nc) {
  if (nc <= c_cap) return;
  int c = std::max(nc, c_cap * 2);
  dfree(c_pt); dfree(c_status); dfree(c_n); dfree(c_r);
  dfree(c_prio); dfree(c_win); dfree(c_rank); dfree(c_act); dfree(c_slow);
  dalloc(&c_act, (size_t)c);
  dalloc(&c_slow, (size_t)c);
  dalloc(&c_pt, (size_t)c);
  dalloc(&c_status, (size_t)c);
  dalloc(&c_n, (size_t)c);
  dalloc(&c_r, (size_t)c);
  dalloc(&c_prio, (size_t)c);
  dalloc(&c_win, (size_t)c);
  dalloc(&c_rank, (size_t)c);
  c_cap = c;
}

int Refiner::cands_per_pass(int K) const {
  long long c = opt.cand_budget / (6LL * K + 16);
  if (c < 1) c = 1;
  if (c > (1LL << 30)) c = 1LL << 30;
  return (int)c;
}

// Bad-pass candidate selection: when more triangles are bad than a pass
// takes, keep the `cap` smallest (priority, index) candidates, in index order
// -- the order Triangle's bad-triangle queue serves them
// (triangle.cpp:6932-6968).  Histogram of priority buckets -> cut bucket ->
// ordered compactions; no host-side sort.

__global__ __launch_bounds__(256) void k_prio_hist(MeshD m, const int *ids,
                                                   int n, int *prio,
                                                   unsigned *hist) {
  __shared__ unsigned h[NBUCKET];
  for (int b = threadIdx.x; b < NBUCKET; b += blockDim.x) h[b] = 0;
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    int p = m.tbucket[ids[i]];
    prio[i] = p;
    atomicAdd(&h[p], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < NBUCKET; b += blockDim.x)
    if (h[b]) atomicAdd(&hist[b], h[b]);
}

int Refiner::select_by_priority(int *ids, int nb, int cap, int **out_ids) {
  *out_ids = ids;
  if (nb <= cap) return nb;
  if (opt.priority != 1 && opt.priority != 2) return cap;  // index order
  if (nb > sel_cap) {
    dfree(sel_prio); dfree(sel_tmp); dfree(sel_list);
    sel_cap = std::max(nb, 2 * sel_cap);
    dalloc(&sel_prio, (size_t)sel_cap);
    dalloc(&sel_tmp, (size_t)sel_cap);
    dalloc(&sel_list, (size_t)sel_cap);
  }
  if (!d_hist) dalloc(&d_hist, NBUCKET);
  HC(hipMemsetAsync(d_hist, 0, sizeof(unsigned) * NBUCKET, stream));
  k_prio_hist<<<std::min(nblk(nb), 1024), 256, 0, stream>>>(m, ids, nb, sel_prio, d_hist);
  HC(hipGetLastError());
  std::vector<unsigned> h(NBUCKET);
  HC(hipMemcpyAsync(h.data(), d_hist, sizeof(unsigned) * NBUCKET, hipMemcpyDeviceToHost, stream));
  HC(hipStreamSynchronize(stream));
  long long acc = 0;
  int bcut = 0;
  for (; bcut < NBUCKET; bcut++) {
    if (acc + h[bcut] >= (unsigned long long)cap) break;
    acc += h[bcut];
  }
  int take = (int)(cap - acc);  // 1 <= take <= h[bcut]
  compact(PredPrioEq{sel_prio, bcut}, nb, sel_tmp, nullptr, nullptr);
  HC(hipMemcpyAsync(h_scalar + 1, sel_tmp + (take - 1), sizeof(int), hipMemcpyDeviceToHost, stream));
  HC(hipStreamSynchronize(stream));
  int icut = h_scalar[1];
  int got = compact(PredPrioSel{sel_prio, bcut, icut}, nb, sel_list, nullptr, ids);
  *out_ids = sel_list;
  return got;
}

// Pass descriptor common to both loops (counts and epochs filled later).
PassD Refiner::pass_base(int Kc) const {
  PassD ps{};
  ps.K = Kc;
  ps.Kf = opt.cav_lds_cap;
  ps.R = 6 * Kc + 16;
  ps.cpt = c_pt;
  ps.cstatus = c_status;
  ps.cn = c_n;
  ps.cr = c_r;
  ps.cprio = c_prio;
  ps.cwin = c_win;
  ps.crank = c_rank;
  ps.ws = c_ws;
  ps.pws = c_pws;
  ps.cnt = d_cnt;
  ps.feat = nullptr;
  ps.ctl = nullptr;
  ps.ids_sel = nullptr;
  ps.slow = c_slow;
  ps.iters = std::max(1, opt.claim_iters);
  ps.shared_ring = opt.shared_ring;
  return ps;
}

DelD Refiner::del_base() const {
  DelD dd{};
  dd.policy = opt.deletion;
  dd.ws = del_ws;
  dd.dv = del_dv;
  dd.dk = del_dk;
  dd.pws = c_pws;
  dd.cnt = d_cnt;
  dd.use_fp32 = pr.use_fp32;
  return dd;
}

// The deletion step of a pass (host loop: dd.nseg and dd.epoch set; graph
// loop: both come from the control block).  The scan kernel's grid is the
// exact-predicate arena's (one arena slot per lane).
void Refiner::enqueue_deletion(const DelD &dd) {
  // host loop: grids sized by the pass's subsegment candidates (most
  // segment passes hold few); graph loop: sized for the largest pass
  const long long n = dd.ctl ? DEL_CANDS : std::min(dd.nseg, DEL_CANDS);
  const int gs = (int)std::max(1LL, std::min<long long>(pws_blocks * 4, (2 * n + 63) / 64));
  const int g = (int)std::max(1LL, std::min<long long>(2048, (2 * n + 63) / 64));
  tbegin(K_DEL_SCAN);
  k_del_scan<<<gs, DEL_LANES, 0, stream>>>(m, pr, dd);
  tend(K_DEL_SCAN, dd.ctl ? 0 : n);
  tbegin(K_DEL_APPLY);
  k_del_apply<<<g, DEL_LANES, 0, stream>>>(m, dd);
  tend(K_DEL_APPLY, 0);
  HC(hipGetLastError());
}

// Drops deleted vertices: live points move to the front in index order and
// triangles and subsegments are renumbered.
void Refiner::compact_points() {
  if (np > p_tmp_cap) {
    dfree(p_new); dfree(p_tmp); dfree(pf_tmp); dfree(pp_tmp);
    p_tmp_cap = np;
    dalloc(&p_new, (size_t)p_tmp_cap);
    dalloc(&p_tmp, (size_t)p_tmp_cap);
    dalloc(&pf_tmp, (size_t)p_tmp_cap);
    dalloc(&pp_tmp, (size_t)p_tmp_cap);
  }
  const int np2 = compact(PredPAlive{m.pflag}, np, nullptr, p_new, nullptr);
  k_pcompact<<<std::min(nblk(np), 4096), 256, 0, stream>>>(m.P, m.pflag, m.ppar, np, p_new,
                                                             p_tmp, pf_tmp, pp_tmp);
  k_premap<<<std::min(nblk(std::max(nt, ns)), 4096), 256, 0, stream>>>(m, nt, ns, p_new);
  HC(hipGetLastError());
  HC(hipMemcpyAsync(m.P, p_tmp, sizeof(double2) * np2, hipMemcpyDeviceToDevice, stream));
  HC(hipMemcpyAsync(m.pflag, pf_tmp, np2, hipMemcpyDeviceToDevice, stream));
  HC(hipMemcpyAsync(m.ppar, pp_tmp, sizeof(int) * np2, hipMemcpyDeviceToDevice, stream));
  np = np2;
}

// Claim iterations: claim -> win, repeated; from the second iteration on a
// claim skips candidates blocked by an earlier iteration's winner stamps.  In
// the host loop ps.epoch advances here; in the graph loop the kernels derive
// it from the control block.
// Claim iterations of a pass: segment-only passes may take fewer
// (options.seg_claim_iters) than passes with bad triangles.
int Refiner::iters_for(bool seg_only) const {
  return std::max(1, seg_only && opt.seg_claim_iters > 0 ? opt.seg_claim_iters : opt.claim_iters);
}

void Refiner::enqueue_claims(PassD ps, int gc, int iters) {
  ps.iters = iters;
  for (int it = 0; it < iters; it++) {
    if (!ps.ctl) ps.epoch = ++epoch;
    ps.iter = it;
    tbegin(K_CLAIM);
    k_claim<CLAIM_LANES><<<gc, LIST_BLOCK, 0, stream>>>(m, ps, c_act);
    tend(K_CLAIM, ps.nc);
    tbegin(K_WIN);
    k_win<CLAIM_LANES><<<gc, LIST_BLOCK, 0, stream>>>(m, ps, c_act);
    tend(K_WIN, ps.nc);
    HC(hipGetLastError());
  }
  if (!ps.ctl) ++epoch;  // the winners' stamp epoch
}

// One conflict pass over ids[0..nc): candidates below nseg split
// subsegments, the rest are bad triangles.  Returns the number of winners,
// *nwin_seg of them subsegment splits.
//
// list_enc: also list the subsegments encroached after the pass in tmp_list
// (listed_enc), queued behind the apply kernel so that one synchronisation
// serves both; the next segment phase then starts from that list.
int Refiner::run_pass(const int *ids, int nc, int nseg, long long *progress, bool *overflowed,
                      int *nwin_seg, bool list_enc) {
  const int nbad = nc - nseg;
  // room for every candidate winning: the winners' count is only read back
  // with the pass counters, after the apply kernel
  ensure_capacity(np + (long long)nc, nt + 2LL * nc, ns + (long long)nseg);
  ensure_ws((long long)nc * (6LL * K + 16));
  PassD ps = pass_base(K);
  ps.ids = ids;
  ps.nc = nc;
  ps.nseg = nseg;
  const int iters = iters_for(nseg == nc);
  ps.iters = iters;
  if (epoch + iters + 3 >= (1ull << 11)) {  // epoch wrap: clear claims
    HC(hipMemsetAsync(m.claim, 0, sizeof(unsigned long long) * 2 * capT, stream));
    epoch = 0;
  }
  ps.iter = 0;
  bool want_feat = nbad > 0 && (opt.priority == 2 || opt.log_training);
  if (want_feat && nc > feat_cap) {
    dfree(c_feat); dfree(c_score); dfree(c_deg);
    feat_cap = std::max(nc, 2 * feat_cap);
    dalloc(&c_feat, (size_t)feat_cap * FEAT);
    dalloc(&c_score, (size_t)feat_cap);
    dalloc(&c_deg, (size_t)feat_cap);
  }
  ps.feat = want_feat ? c_feat : nullptr;
  ps.pass_index = pass_index;
  if (!cnt_zero) HC(hipMemsetAsync(d_cnt, 0, CNT_N * sizeof(unsigned long long), stream));
  cnt_zero = false;
  if (del_on && nseg > 0) {  // vertex deletion first, with its own claim epoch
    DelD dd = del_base();
    dd.ids = ids;
    dd.nseg = nseg;
    dd.epoch = ++epoch;
    enqueue_deletion(dd);
  }
  ps.epoch0 = epoch + 1;
  ps.epoch = epoch + 1;
  // 64/lg candidates per wave64 block (the exact kernel's grid is capped by
  // its arena in launch_cavity)
  int lg = cavity_lanes(std::min(K, opt.cav_lds_cap));
  auto cav_grid = [&](int n) {
    return std::min((n + 64 / lg - 1) / (64 / lg), opt.grid_cavity);
  };
  if (nseg > 0) {
    tbegin(K_SEGCAV);
    launch_cavity(true, cav_grid(nseg), ps);
    tend(K_SEGCAV, nseg);
  }
  if (nbad > 0) {
    tbegin(K_BADCAV);
    launch_cavity(false, cav_grid(nbad), ps);
    tend(K_BADCAV, nbad);
  }
  HC(hipGetLastError());
  int gc = std::min(nblk((long long)nc * CLAIM_LANES, LIST_BLOCK), opt.grid_claim);
  if (nbad > 0 && opt.priority == 2) {
    if (!mlp_params) throw std::runtime_error("priority 2 needs a model (set_priority_model)");
    tbegin(K_MLP);
    mlp_forward_device(c_feat + (size_t)nseg * FEAT, nbad, mlp_params, c_score + nseg, stream);
    tend(K_MLP, nbad);
    // the first claim iteration turns the scores into claim priorities
    ps.score = c_score;
    ps.lshift = opt.learned_shift;
    HC(hipGetLastError());
  }
  if (nbad > 0 && opt.log_training) {  // conflict logs for training (pure bad passes)
    if (nt > t_cnt_cap) {
      dfree(t_cnt);
      t_cnt_cap = std::max<long long>(nt, capT);
      dalloc(&t_cnt, (size_t)t_cnt_cap);
    }
    HC(hipMemsetAsync(t_cnt, 0, sizeof(int) * nt, stream));
    k_claim_count<<<gc, 256, 0, stream>>>(ps, t_cnt);
    k_conflict_degree<<<gc, 256, 0, stream>>>(ps, t_cnt, c_deg);
    std::vector<float> hf((size_t)nc * FEAT);
    std::vector<int> hd(nc);
    HC(hipMemcpyAsync(hf.data(), c_feat, sizeof(float) * hf.size(), hipMemcpyDeviceToHost, stream));
    HC(hipMemcpyAsync(hd.data(), c_deg, sizeof(int) * nc, hipMemcpyDeviceToHost, stream));
    HC(hipStreamSynchronize(stream));
    for (int i = nseg; i < nc; i++) {
      if (hd[i] < 0) continue;
      log_X.insert(log_X.end(), hf.begin() + (size_t)i * FEAT, hf.begin() + (size_t)(i + 1) * FEAT);
      log_t.push_back(std::log2(1.0f + (float)hd[i]));
    }
  }
  enqueue_claims(ps, gc, iters);
  if (nc <= SMALL_PASS) {  // ranks and apply in one launch
    tbegin(K_APPLY);
    k_rank_apply<APPLY_LANES><<<1, LIST_BLOCK, 0, stream>>>(m, pr, ps, np, nt, ns);
    tend(K_APPLY, 0);
  } else {
    compact(PredWin{c_win}, nc, nullptr, c_rank, nullptr, d_cnt + CNT_NWIN);
    tbegin(K_APPLY);
    k_apply<APPLY_LANES><<<std::min(nblk((long long)nc * APPLY_LANES, LIST_BLOCK),
                                    opt.grid_apply),
                           LIST_BLOCK, 0, stream>>>(m, pr, ps, np, nt, ns);
    tend(K_APPLY, 0);
  }
  HC(hipGetLastError());
  // every segment candidate may add one subsegment; the list buffer is not
  // grown here (the pass may still read it)
  const bool listed = list_enc && (long long)ns + nseg <= tmp_list_cap;
  if (listed)
    compact(PredSencN{m.senc, ns, d_cnt + CNT_SEGWIN}, ns + nseg, tmp_list, nullptr, nullptr,
            d_cnt + CNT_NENC);
  flush_counters();
  listed_enc = listed ? (int)h_cnt[CNT_NENC] : -1;
  const int nwin = (int)h_cnt[CNT_NWIN];
  const int nws = (int)h_cnt[CNT_SEGWIN];
  const long long ndel = (long long)h_cnt[CNT_DEL];
  st.deleted += ndel;
  st.flips += (long long)h_cnt[CNT_FLIPS];
  if (opt.profile && !capturing) ktimes[K_APPLY].items += nwin;  // winners applied
  np += nwin;
  nt += 2 * nwin;
  ns += nws;
  st.rej_overflow += h_cnt[CNT_OVERFLOW];
  st.rej_encroach += h_cnt[CNT_REJ_ENC];
  st.rej_blocked += h_cnt[CNT_REJ_BLOCKED];
  st.giveups += h_cnt[CNT_GIVEUPS];
  st.lost += (long long)h_cnt[CNT_VALID] - nwin;
  st.pred_fp32 += h_cnt[CNT_FP32];
  st.pred_fp64 += h_cnt[CNT_FP64];
  st.pred_exact += h_cnt[CNT_EXACT];
  st.pred_dd += h_cnt[CNT_DD];
  st.big_cavities += h_cnt[CNT_BIG];
  st.tri_examined_seg += h_cnt[CNT_EXAM_SEG];
  st.tri_examined_bad += h_cnt[CNT_EXAM_BAD];
  st.ws_ints_seg += h_cnt[CNT_WS_SEG];
  st.ws_ints_bad += h_cnt[CNT_WS_BAD];
  st.seg_splits += nws;
  st.inserted += nwin - nws;
  *progress = nwin + ndel + (long long)h_cnt[CNT_MARKS] + (long long)h_cnt[CNT_GIVEUPS];
  *overflowed = h_cnt[CNT_OVERFLOW] > 0;
  *nwin_seg = nws;
  return nwin;
}

void Refiner::run() {
  memset(&st, 0, sizeof(st));
  for (auto &kt : ktimes) { kt.launches = 0; kt.total_ms = 0; kt.items = 0; }
  HC(hipEventRecord(ev0, stream));
  // --- initialise working mesh from the uploaded input --------------------
  np = np0; nt = nt0; ns = ns0;
  np_round = np0;
  long long wantP = std::min(std::max<long long>(16, 6LL * np0), MAX_POINTS);
  long long wantT = std::min(std::max<long long>(16, 6LL * nt0), MAX_TRIANGLES);
  long long wantS = std::min(std::max<long long>(16, 8LL * ns0 + 16), MAX_POINTS);
  const bool want_inst = in_tinst != nullptr;
  if (!vm_checked) {
    // in-place growth when the device maps memory into reserved ranges; a
    // runtime that reserves but refuses the first mapping (some tools
    // intercepting the calls do) leaves the reallocating arrays in charge
    vm_checked = true;
    use_vm = false;
    if (opt.vmm && vm_setup()) {
      try {
        vm_map_mesh(wantP, wantT, wantS);
        use_vm = true;
      } catch (const std::exception &) {
        HC(hipStreamSynchronize(stream));
        vm_release_all();
      }
    }
  }
  st.grow_in_place = use_vm ? 1 : 0;
  if (use_vm) {
    vm_map_mesh(std::max(capP, wantP), std::max(capT, wantT), std::max(capS, wantS));
  } else if (capP < wantP || capT < wantT || capS < wantS) {
    free_mesh();
    capP = wantP; capT = wantT; capS = wantS;
    dalloc(&m.P, capP); dalloc(&m.pflag, capP); dalloc(&m.ppar, capP);
    dalloc(&m.T, capT);
    dalloc(&m.tflag, capT); dalloc(&m.tbucket, capT); dalloc(&m.claim, 2 * (size_t)capT);
    dalloc(&m.SV, capS); dalloc(&m.SP, capS); dalloc(&m.ST, capS); dalloc(&m.senc, capS);
  }
  if (ns0 > seg_in_cap || !seg_in) {
    dfree(seg_in);
    seg_in_cap = std::max(ns0, 1);
    dalloc(&seg_in, (size_t)seg_in_cap);
  }
  m.SEG = seg_in;
  pr.ipar = d_ipar;
  HC(hipMemsetAsync(d_cnt, 0, CNT_N * sizeof(unsigned long long), stream));
  tbegin(K_INIT);
  HC(hipMemcpyAsync(m.P, in_P, sizeof(double2) * np0, hipMemcpyDeviceToDevice, stream));
  m.batched = want_inst ? 1 : 0;
  k_init_points<<<nblk(np0), 256, 0, stream>>>(np0, m.pflag, m.ppar);
  k_init_tris<<<nblk(nt0), 256, 0, stream>>>(in_T, in_N, in_tinst, in_P, nt0, m.T, m.tflag,
                                              m.claim, d_cnt);
  // segment hash: capacity power of two >= 2*ns + 16
  unsigned long long hc = 1;
  while (hc < 2ull * (unsigned long long)ns0 + 16) hc <<= 1;
  if ((long long)hc > h_cap) {
    dfree(h_key); dfree(h_val);
    dalloc(&h_key, (size_t)hc);
    dalloc(&h_val, (size_t)hc);
    h_cap = (long long)hc;
  }
  HC(hipMemsetAsync(h_key, 0xff, sizeof(unsigned long long) * hc, stream));
  HC(hipMemsetAsync(h_val, 0x7f, sizeof(int) * hc, stream));
  if (ns0 > 0) {
    k_init_segs<<<nblk(ns0), 256, 0, stream>>>(in_S, ns0, seg_in, m.SV, m.SP, m.ST, m.senc, h_key, h_val, hc - 1);
    k_attach_segs<<<nblk(nt0), 256, 0, stream>>>(m, nt0, h_key, h_val, hc - 1);
    k_check_segs<<<nblk(ns0), 256, 0, stream>>>(m.ST, ns0, d_cnt);
    k_init_enc<<<nblk(nt0), 256, 0, stream>>>(m, pr, nt0);
  }
  tend(K_INIT, nt0);
  HC(hipGetLastError());
  read_counters();
  if (h_cnt[CNT_ERRORS])
    throw std::runtime_error("input is not a valid triangulation: " +
                             std::to_string(h_cnt[CNT_ERRORS]) +
                             " neighbour links or segments do not match triangle edges");
  // --- workspace ----------------------------------------------------------
  K = opt.cav_cap0;
  epoch = 0;
  pass_index = 0;
  log_X.clear();
  log_t.clear();
  HC(hipMemsetAsync(m.claim, 0, sizeof(unsigned long long) * 2 * capT, stream));
  // Candidate workspace: the graph loop sizes its passes on the device and
  // takes the whole budget; the host loop grows it to each pass's need.
  if (opt.graph) ensure_ws(opt.cand_budget);
  // Exact-stage arena: GCDT_INCIRCLE_WS doubles for each lane of the exact
  // kernel's grid (at most EXACT_GRID blocks of 64; the filtered kernel does
  // not use it).
  if (!c_pws) {
    pws_blocks = EXACT_GRID / 4;
    dalloc(&c_pws, (size_t)pws_blocks * 256 * GCDT_INCIRCLE_WS);
  }
  // vertex deletion: Triangle's rule deletes in Chew-mode instances only
  del_on = opt.deletion == 2 || (opt.deletion == 1 && (m.batched ? in_any_chew : !pr.ruppert));
  if (del_on && !del_ws) {
    dalloc(&del_ws, (size_t)2 * DEL_CANDS * DEL_WS);
    dalloc(&del_dv, (size_t)2 * DEL_CANDS);
    dalloc(&del_dk, (size_t)2 * DEL_CANDS);
  }
  st.status = 1;
  if (opt.graph) loop_graph();
  else if (opt.mixed) loop_mixed();
  else loop_host();
  st.final_cav_cap = K;
  // points inserted in the last round (the reference's PStatus "new" bit,
  // refine.h:36-106, cleared at every round's start)
  if (np > np_round) {
    k_mark_new<<<std::min(nblk(np - np_round), 4096), 256, 0, stream>>>(m.pflag, np_round, np);
    HC(hipGetLastError());
  }
  if (st.deleted > 0) compact_points();
  // given-up elements still in the mesh: a run that converged otherwise
  // reports them as status 3, not as converged
  HC(hipMemsetAsync(d_cnt, 0, CNT_N * sizeof(unsigned long long), stream));
  k_count_left<<<std::min(nblk(std::max(nt, ns)), 2048), 256, 0, stream>>>(m.tflag, nt, m.senc, ns,
                                                                           d_cnt);
  HC(hipGetLastError());
  read_counters();
  st.giveup_tris_left = (long long)h_cnt[CNT_LEFT_TRI];
  st.giveup_segs_left = (long long)h_cnt[CNT_LEFT_SEG];
  if (st.status == 0 && (st.giveup_tris_left > 0 || st.giveup_segs_left > 0)) st.status = 3;
  // --- export --------------------------------------------------------------
  if (nt > newid_cap) {
    dfree(newid);
    newid_cap = nt;
    dalloc(&newid, (size_t)newid_cap);
  }
  nt_out = compact(PredAlive{m.tflag}, nt, nullptr, newid, nullptr);
  if (nt_out > o_cap) {
    dfree(o_T); dfree(o_N); dfree(o_tstat); dfree(o_inst);
    o_cap = nt_out;
    dalloc(&o_T, 3 * (size_t)o_cap);
    dalloc(&o_N, 3 * (size_t)o_cap);
    dalloc(&o_tstat, (size_t)o_cap);
    if (m.batched) dalloc(&o_inst, (size_t)o_cap);
  }
  if (m.batched && !o_inst) dalloc(&o_inst, (size_t)std::max(o_cap, 1));
  tbegin(K_EXPORT);
  k_export<<<nblk(nt), 256, 0, stream>>>(m, nt, newid, o_T, o_N, o_tstat, m.batched ? o_inst : nullptr);
  tend(K_EXPORT, nt);
  HC(hipGetLastError());
  HC(hipEventRecord(ev1, stream));
  HC(hipStreamSynchronize(stream));
  float ms = 0.f;
  HC(hipEventElapsedTime(&ms, ev0, ev1));
  st.device_ms = ms;
  st.n_points = np;
  st.n_triangles = nt_out;
  st.n_subsegs = ns;
  collect_times();
}

// Host-driven rounds: the host reads every count and decides every step.
void Refiner::loop_host() {
  int prev_win = -1;
  listed_enc = -1;
  ensure_cand(std::min(cands_per_pass(K), std::max(4 * nt0, 1024)));
  for (int round = 0; round < opt.max_rounds; round++) {
    st.rounds = round + 1;
    np_round = np;
    // segment phase: until at most mix_below subsegments are encroached
    int ne;
    for (;;) {
      if (listed_enc >= 0) {  // listed behind the last pass
        ne = listed_enc;
        listed_enc = -1;
      } else {
        ensure_tmp(ns);
        ne = compact(PredSenc{m.senc}, ns, tmp_list, nullptr, nullptr);
      }
      if (ne <= opt.mix_below) break;
      int nc = std::min(ne, cands_per_pass(K));
      ensure_cand(nc);
      long long prog = 0;
      bool of = false;
      int nws = 0;
      run_pass(tmp_list, nc, nc, &prog, &of, &nws, true);
      st.seg_passes++;
      if (of && K < opt.cav_cap_max) K = std::min(2 * K, opt.cav_cap_max);
    }
    if (bad_pass(ne, &prev_win)) break;
  }
}

// Mixed rounds (options.mixed): one pass per round over the encroached
// subsegments followed by the selected bad triangles.
void Refiner::loop_mixed() {
  int prev_win = -1;
  ensure_cand(std::min(cands_per_pass(K), std::max(4 * nt0, 1024)));
  for (int round = 0; round < opt.max_rounds; round++) {
    st.rounds = round + 1;
    np_round = np;
    ensure_tmp(ns);
    int ne = compact(PredSenc{m.senc}, ns, tmp_list, nullptr, nullptr);
    if (bad_pass(ne, &prev_win)) break;
  }
}

// A round's bad pass: the ne encroached subsegments listed in tmp_list (none
// in the reference's round) first, then the selected bad triangles.
// Returns true when the refinement has ended (st.status set).
bool Refiner::bad_pass(int ne, int *prev_win) {
  const int cap = cands_per_pass(K);
  ne = std::min(ne, cap);
  ensure_mix(std::min<long long>((long long)ns + nt, (long long)cap + 1));
  if (ne > 0)
    HC(hipMemcpyAsync(mix_list, tmp_list, sizeof(int) * (size_t)ne, hipMemcpyDeviceToDevice, stream));
  ensure_tmp(nt);
  const int ntile = (nt + CT - 1) / CT;
  tbegin(K_MARKBAD);
  if (ntile > 1) {
    ensure_cb(ntile);
    k_mark_bad_count<<<ntile, CB, 0, stream>>>(m, pr, nt, cb_cnt);
  } else {
    k_mark_bad<<<std::min(nblk(nt), 65536), 256, 0, stream>>>(m, pr, nt, nullptr);
  }
  tend(K_MARKBAD, nt);
  int nb = compact(PredBad{m.tflag}, nt, tmp_list, nullptr, nullptr, nullptr, ntile > 1);
  if (ne == 0 && nb == 0) {
    st.status = 0;
    return true;
  }
  int capb = cap - ne;
  if (opt.cand_factor > 0 && *prev_win >= 0)
    capb = (int)std::min<long long>(capb, std::max<long long>(opt.cand_min, (long long)opt.cand_factor * *prev_win));
  int *bids = tmp_list;
  int nbc = nb > 0 && capb > 0 ? select_by_priority(tmp_list, nb, capb, &bids) : 0;
  const int *ids = bids;
  if (ne > 0) {
    if (nbc > 0)
      HC(hipMemcpyAsync(mix_list + ne, bids, sizeof(int) * (size_t)nbc, hipMemcpyDeviceToDevice, stream));
    ids = mix_list;
  }
  const int nc = ne + nbc;
  ensure_cand(nc);
  long long prog = 0;
  bool of = false;
  int nws = 0;
  int nwin = run_pass(ids, nc, ne, &prog, &of, &nws, !opt.mixed);
  if (ne > 0) st.seg_passes++;
  if (nbc > 0) {
    pass_index++;
    st.bad_passes++;
    st.cand_evals += nbc;
    *prev_win = nwin - nws;
  }
  if (of && K < opt.cav_cap_max) {
    K = std::min(2 * K, opt.cav_cap_max);
    prog++;
  }
  if (prog == 0) {
    st.status = 2;
    return true;
  }
  return false;
}

void Refiner::ensure_mix(long long n) {
  if (n <= mix_cap) return;
  dfree(mix_list);
  mix_cap = std::max<long long>(n, 2 * mix_cap);
  dalloc(&mix_list, (size_t)mix_cap);
}

// ---------------------------------------------------------------------------
// Graph-driven rounds
// ---------------------------------------------------------------------------
// Buffers whose size depends on K and on the mesh capacity.
void Refiner::size_pass_buffers() {
  ensure_cand(cands_per_pass(K));
  ensure_tmp((int)std::max(capT, capS));
  if (capT > sel_cap) {
    sel_cap = (int)capT;
    if (use_vm) {
      vm_sel_prio.map_to((size_t)sel_cap * sizeof(int));
      vm_sel_tmp.map_to((size_t)sel_cap * sizeof(int));
      vm_sel_list.map_to((size_t)sel_cap * sizeof(int));
      sel_prio = (int *)vm_sel_prio.ptr();
      sel_tmp = (int *)vm_sel_tmp.ptr();
      sel_list = (int *)vm_sel_list.ptr();
    } else {
      dfree(sel_prio); dfree(sel_tmp); dfree(sel_list);
      dalloc(&sel_prio, (size_t)sel_cap);
      dalloc(&sel_tmp, (size_t)sel_cap);
      dalloc(&sel_list, (size_t)sel_cap);
    }
  }
  if (!d_hist) dalloc(&d_hist, NBUCKET);
  if (opt.priority == 2 && feat_cap < c_cap) {
    dfree(c_feat); dfree(c_score); dfree(c_deg);
    feat_cap = c_cap;
    dalloc(&c_feat, (size_t)feat_cap * FEAT);
    dalloc(&c_score, (size_t)feat_cap);
    dalloc(&c_deg, (size_t)feat_cap);
  }
}

// Lanes per candidate in the cavity kernel: the configured group size,
// widened until the groups' LDS fits one block.
int Refiner::cavity_lanes(int Kc) const {
  int lg = opt.cavity_lanes;
  while (lg < 64 && (size_t)(64 / lg) * group_lds_ints(Kc, lg) * sizeof(int) > BIG_LDS_BYTES)
    lg *= 2;
  return lg;
}

template <bool IS_SEG, bool FAST>
static void launch_cavity_t(int lg, int grid, size_t lds, hipStream_t s, const MeshD &m,
                            const ParamsD &pr, const PassD &ps, int kmax) {
  if (FAST) {
    switch (lg) {
      case 8: k_cavity_fast<IS_SEG, 8><<<grid, 64, lds, s>>>(m, pr, ps, kmax); break;
      case 16: k_cavity_fast<IS_SEG, 16><<<grid, 64, lds, s>>>(m, pr, ps, kmax); break;
      case 32: k_cavity_fast<IS_SEG, 32><<<grid, 64, lds, s>>>(m, pr, ps, kmax); break;
      default: k_cavity_fast<IS_SEG, 64><<<grid, 64, lds, s>>>(m, pr, ps, kmax); break;
    }
  } else {
    switch (lg) {
      case 8: k_cavity_slow<IS_SEG, 8><<<grid, 64, lds, s>>>(m, pr, ps, kmax); break;
      case 16: k_cavity_slow<IS_SEG, 16><<<grid, 64, lds, s>>>(m, pr, ps, kmax); break;
      case 32: k_cavity_slow<IS_SEG, 32><<<grid, 64, lds, s>>>(m, pr, ps, kmax); break;
      default: k_cavity_slow<IS_SEG, 64><<<grid, 64, lds, s>>>(m, pr, ps, kmax); break;
    }
  }
}

// Filtered-predicate pass over all candidates, then the exact-predicate pass
// over the candidates it deferred (their count stays on the device).
void Refiner::launch_cavity(bool is_seg, int grid, const PassD &ps) {
  int kf = std::min(ps.K, ps.Kf);
  // the exact kernel takes a whole wave per candidate: its candidates are
  // the cavities beyond the fast kernel's LDS (33..K triangles, whose sort
  // and boundary cycle cost O(n^2 / lanes)) and the undecided signs
  int lg = cavity_lanes(kf), lgs = 64;
  size_t lds = sizeof(int) * (size_t)(64 / lg) * group_lds_ints(kf, lg);
  size_t lds_slow = sizeof(int) * (size_t)(64 / lgs) * group_lds_ints(ps.K, lgs);
  if (lds_slow > (size_t)BIG_LDS_BYTES)
    throw std::runtime_error("cavity capacity " + std::to_string(ps.K) + " needs more LDS than a block has");
  if (grid < 1) grid = 1;
  int slow_grid = std::min(grid, pws_blocks * 4);
  if (is_seg) {
    launch_cavity_t<true, true>(lg, grid, lds, stream, m, pr, ps, opt.cav_cap_max);
    launch_cavity_t<true, false>(lgs, slow_grid, lds_slow, stream, m, pr, ps, opt.cav_cap_max);
  } else {
    launch_cavity_t<false, true>(lg, grid, lds, stream, m, pr, ps, opt.cav_cap_max);
    launch_cavity_t<false, false>(lgs, slow_grid, lds_slow, stream, m, pr, ps, opt.cav_cap_max);
  }
  HC(hipGetLastError());
}

// Per-pass counters of the graph loop live in the control block.
unsigned long long *Refiner::ctl_cnt_ptr() const {
  return reinterpret_cast<unsigned long long *>(reinterpret_cast<char *>(d_ctl) +
                                                offsetof(Ctl, cnt));
}

void Refiner::enqueue_seg_pass() {
  k_gate_seg<<<1, 1, 0, stream>>>(d_ctl);
  k_dcount<PredSenc><<<GC, CB, 0, stream>>>(PredSenc{m.senc}, d_ctl, C_NS, -1, g_cnt);
  k_dscan<FinSeg><<<1, GC, 0, stream>>>(g_cnt, g_off, d_ctl, -1, FinSeg{});
  k_dscatter<PredSenc><<<GC, CB, 0, stream>>>(PredSenc{m.senc}, d_ctl, C_NS, -1, g_off,
                                              tmp_list, nullptr, nullptr);
  PassD ps = pass_base(K);
  ps.ids = tmp_list;
  ps.nseg = -1;
  ps.ctl = d_ctl;
  ps.cnt = ctl_cnt_ptr();
  if (del_on) {  // vertex deletion first, with its own claim epoch
    DelD dd = del_base();
    dd.ids = tmp_list;
    dd.ctl = d_ctl;
    dd.cnt = ctl_cnt_ptr();
    enqueue_deletion(dd);
    ps.eskip = 1;
  }
  launch_cavity(true, opt.grid_cavity, ps);
  const int seg_iters = iters_for(true);
  ps.iters = seg_iters;
  enqueue_claims(ps, opt.grid_claim, seg_iters);
  k_dcount<PredWin><<<GC, CB, 0, stream>>>(PredWin{c_win}, d_ctl, C_NC, -1, g_cnt);
  k_dscan<FinWin><<<1, GC, 0, stream>>>(g_cnt, g_off, d_ctl, -1, FinWin{});
  k_dscatter<PredWin><<<GC, CB, 0, stream>>>(PredWin{c_win}, d_ctl, C_NC, -1, g_off,
                                             nullptr, c_rank, nullptr);
  k_apply<APPLY_LANES><<<opt.grid_apply, LIST_BLOCK, 0, stream>>>(m, pr, ps, 0, 0, 0);
  k_end_seg<<<1, 1, 0, stream>>>(d_ctl, seg_iters, del_on ? 1 : 0);
  HC(hipGetLastError());
}

void Refiner::enqueue_bad_pass() {
  k_gate_bad<<<1, 1, 0, stream>>>(d_ctl);
  k_mark_bad<<<4096, 256, 0, stream>>>(m, pr, 0, d_ctl);
  k_dcount<PredBad><<<GC, CB, 0, stream>>>(PredBad{m.tflag}, d_ctl, C_NT, -1, g_cnt);
  k_dscan<FinBad><<<1, GC, 0, stream>>>(g_cnt, g_off, d_ctl, -1, FinBad{d_hist});
  k_dscatter<PredBad><<<GC, CB, 0, stream>>>(PredBad{m.tflag}, d_ctl, C_NT, -1, g_off,
                                             tmp_list, nullptr, nullptr);
  if (opt.priority == 1 || opt.priority == 2) {
    k_dprio_hist<<<512, 256, 0, stream>>>(m, tmp_list, d_ctl, sel_prio, d_hist);
    k_dcut<<<1, 1024, 0, stream>>>(d_hist, d_ctl);
    PredPrioEqC peq{sel_prio, d_ctl};
    k_dcount<PredPrioEqC><<<GC, CB, 0, stream>>>(peq, d_ctl, C_NB, C_SEL, g_cnt);
    k_dscan<FinNone><<<1, GC, 0, stream>>>(g_cnt, g_off, d_ctl, C_SEL, FinNone{});
    k_dscatter<PredPrioEqC><<<GC, CB, 0, stream>>>(peq, d_ctl, C_NB, C_SEL, g_off,
                                                   sel_tmp, nullptr, nullptr);
    k_dicut<<<1, 1, 0, stream>>>(sel_tmp, d_ctl);
    PredPrioSelC psel{sel_prio, d_ctl};
    k_dcount<PredPrioSelC><<<GC, CB, 0, stream>>>(psel, d_ctl, C_NB, C_SEL, g_cnt);
    k_dscan<FinSel><<<1, GC, 0, stream>>>(g_cnt, g_off, d_ctl, C_SEL, FinSel{});
    k_dscatter<PredPrioSelC><<<GC, CB, 0, stream>>>(psel, d_ctl, C_NB, C_SEL, g_off,
                                                    sel_list, nullptr, tmp_list);
  }
  PassD ps = pass_base(K);
  ps.ids = tmp_list;
  ps.ids_sel = sel_list;
  ps.ctl = d_ctl;
  ps.cnt = ctl_cnt_ptr();
  ps.feat = opt.priority == 2 ? c_feat : nullptr;
  launch_cavity(false, opt.grid_cavity, ps);
  if (opt.priority == 2) {
    const int *ci = reinterpret_cast<const int *>(reinterpret_cast<const char *>(d_ctl) +
                                                  offsetof(Ctl, i));
    mlp_forward_device_n(c_feat, ci + C_NC, ci + C_GO, mlp_params, c_score, 4096, stream);
    k_learned_prio<<<4096, 256, 0, stream>>>(ps, c_score, opt.learned_shift);
  }
  const int bad_iters = iters_for(false);
  ps.iters = bad_iters;
  enqueue_claims(ps, opt.grid_claim, bad_iters);
  k_dcount<PredWin><<<GC, CB, 0, stream>>>(PredWin{c_win}, d_ctl, C_NC, -1, g_cnt);
  k_dscan<FinWin><<<1, GC, 0, stream>>>(g_cnt, g_off, d_ctl, -1, FinWin{});
  k_dscatter<PredWin><<<GC, CB, 0, stream>>>(PredWin{c_win}, d_ctl, C_NC, -1, g_off,
                                             nullptr, c_rank, nullptr);
  k_apply<APPLY_LANES><<<opt.grid_apply, LIST_BLOCK, 0, stream>>>(m, pr, ps, 0, 0, 0);
  k_end_bad<<<1, 1, 0, stream>>>(d_ctl, bad_iters);
  HC(hipGetLastError());
}

void Refiner::capture_graph() {
  if (gexec) { HC(hipGraphExecDestroy(gexec)); gexec = nullptr; }
  if (graph) { HC(hipGraphDestroy(graph)); graph = nullptr; }
  HC(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  capturing = true;
  try {
    for (int r = 0; r < opt.graph_rounds; r++) {
      for (int q = 0; q < opt.graph_seg_passes; q++) enqueue_seg_pass();
      enqueue_bad_pass();
    }
  } catch (...) {
    capturing = false;
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  capturing = false;
  HC(hipStreamEndCapture(stream, &graph));
  HC(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
  g_K = K;
}

void Refiner::loop_graph() {
  if (opt.priority == 2 && !mlp_params)
    throw std::runtime_error("priority 2 needs a model (set_priority_model)");
  Ctl &h = *h_ctl;
  memset(&h, 0, sizeof(Ctl));
  h.i[C_STATUS] = 1;
  h.i[C_NP] = np;
  h.i[C_NP_ROUND] = np;
  h.i[C_NT] = nt;
  h.i[C_NS] = ns;
  h.i[C_K] = K;
  h.i[C_KCAP] = K;
  h.i[C_PREVWIN] = -1;
  h.i[C_CAPP] = (int)std::min<long long>(capP, INT32_MAX);
  h.i[C_CAPT] = (int)std::min<long long>(capT, INT32_MAX);
  h.i[C_CAPS] = (int)std::min<long long>(capS, INT32_MAX);
  h.i[C_MAXROUNDS] = opt.max_rounds;
  h.i[C_KMAX] = opt.cav_cap_max;
  h.i[C_CANDMIN] = opt.cand_min;
  h.i[C_CANDFACTOR] = opt.cand_factor;
  h.i[C_ITERS] = std::max(1, opt.claim_iters);
  h.i[C_PRIORITY] = opt.priority;
  h.acc[A_BUDGET] = (unsigned long long)opt.cand_budget;
  size_pass_buffers();
  HC(hipMemcpyAsync(d_ctl, &h, sizeof(Ctl), hipMemcpyHostToDevice, stream));
  capture_graph();
  const int iters = std::max(iters_for(false), iters_for(true));
  const long long passes = (long long)opt.graph_rounds * (opt.graph_seg_passes + 1);
  const long long max_launches = 4LL * opt.max_rounds + 100000;
  for (long long launch = 0;; launch++) {
    if (launch > max_launches) throw std::runtime_error("graph loop made no progress");
    if (h.i[C_EPOCH] + passes * (iters + 2) + 1 >= (1LL << 11)) {  // epoch wrap
      HC(hipMemsetAsync(m.claim, 0, sizeof(unsigned long long) * 2 * capT, stream));
      h.i[C_EPOCH] = 0;
      HC(hipMemcpyAsync(d_ctl, &h, sizeof(Ctl), hipMemcpyHostToDevice, stream));
    }
    HC(hipGraphLaunch(gexec, stream));
    HC(hipMemcpyAsync(&h, d_ctl, sizeof(Ctl), hipMemcpyDeviceToHost, stream));
    HC(hipStreamSynchronize(stream));
    if (h.i[C_DONE]) break;
    bool recapture = false;
    if (h.i[C_SUSPEND]) {  // the next pass needs more mesh capacity
      np = h.i[C_NP];
      nt = h.i[C_NT];
      ns = h.i[C_NS];
      long long nc = h.i[C_NC];
      ensure_capacity(np + nc, nt + 2 * nc, ns + nc);
      h.i[C_CAPP] = (int)std::min<long long>(capP, INT32_MAX);
      h.i[C_CAPT] = (int)std::min<long long>(capT, INT32_MAX);
      h.i[C_CAPS] = (int)std::min<long long>(capS, INT32_MAX);
      h.i[C_SUSPEND] = 0;
      if (use_vm) {  // grown in place: the captured graph's arguments still hold
        size_pass_buffers();
        HC(hipMemcpyAsync(d_ctl, &h, sizeof(Ctl), hipMemcpyHostToDevice, stream));
      } else {
        recapture = true;
      }
    }
    if (h.i[C_K] != g_K) {  // cavity capacity doubled
      K = h.i[C_K];
      h.i[C_KCAP] = K;
      recapture = true;
    }
    if (recapture) {
      size_pass_buffers();
      HC(hipMemcpyAsync(d_ctl, &h, sizeof(Ctl), hipMemcpyHostToDevice, stream));
      capture_graph();
    }
  }
  np = h.i[C_NP];
  nt = h.i[C_NT];
  ns = h.i[C_NS];
  K = h.i[C_K];
  pass_index = h.i[C_PASS_INDEX];
  st.status = h.i[C_STATUS];
  st.rounds = h.i[C_ROUNDS];
  st.seg_passes = h.i[C_SEG_PASSES];
  st.bad_passes = h.i[C_BAD_PASSES];
  st.rej_overflow = (long long)h.acc[A_REJ_OVF];
  st.rej_encroach = (long long)h.acc[A_REJ_ENC];
  st.rej_blocked = (long long)h.acc[A_REJ_BLK];
  st.giveups = (long long)h.acc[A_GIVEUPS];
  st.lost = (long long)h.acc[A_LOST];
  st.pred_fp32 = (long long)h.acc[A_FP32];
  st.pred_fp64 = (long long)h.acc[A_FP64];
  st.pred_exact = (long long)h.acc[A_EXACT];
  st.pred_dd = (long long)h.acc[A_DD];
  st.big_cavities = (long long)h.acc[A_BIG];
  st.tri_examined_seg = (long long)h.acc[A_EXAM_SEG];
  st.tri_examined_bad = (long long)h.acc[A_EXAM_BAD];
  st.ws_ints_seg = (long long)h.acc[A_WS_SEG];
  st.ws_ints_bad = (long long)h.acc[A_WS_BAD];
  st.seg_splits = (long long)h.acc[A_SEGSPLITS];
  st.inserted = (long long)h.acc[A_INSERTED];
  st.cand_evals = (long long)h.acc[A_CANDEVALS];
  st.deleted = (long long)h.acc[A_DELETED];
  st.flips = (long long)h.acc[A_FLIPS];
  np_round = h.i[C_NP_ROUND];
}

Refiner::DeviceMesh Refiner::device_mesh() const {
  DeviceMesh d;
  d.P = m.P;
  d.np = np;
  d.T = o_T;
  d.N = o_N;
  d.nt = nt_out;
  d.S = m.SV;
  d.ns = ns;
  d.P_in = inputs.empty() ? nullptr : inputs[0].P;
  d.np_in = inputs.empty() ? 0 : inputs[0].np;
  d.S_in = inputs.empty() ? nullptr : (const int2 *)inputs[0].S;
  d.ns_in = inputs.empty() ? 0 : inputs[0].ns;
  return d;
}

void Refiner::ensure_tmp(int n) {
  if (n <= tmp_list_cap) return;
  const int c = (int)std::min<long long>(std::max<long long>(n, 2LL * tmp_list_cap), MAX_POINTS);
  if (use_vm) {  // in place: the graph loop's captured lists keep their address
    vm_tmp.map_to((size_t)c * sizeof(int));
    tmp_list = (int *)vm_tmp.ptr();
  } else {
    dfree(tmp_list);
    dalloc(&tmp_list, (size_t)c);
  }
  tmp_list_cap = c;
}

void Refiner::download(MeshHost &out) {
  out.np = np;
  out.nt = nt_out;
  out.ns = ns;
  out.P.resize(2 * (size_t)np);
  out.T.resize(3 * (size_t)nt_out);
  out.N.resize(3 * (size_t)nt_out);
  out.S.resize(2 * (size_t)ns);
  out.pstat.resize((size_t)np);
  out.tstat.resize((size_t)nt_out);
  out.tinst.clear();
  if (m.batched && nt_out) out.tinst.resize((size_t)nt_out);
  HostDst d;
  d.P = out.P.data();
  d.T = out.T.data();
  d.N = out.N.data();
  d.S = out.S.data();
  d.pstat = out.pstat.data();
  d.tstat = out.tstat.data();
  d.tinst = out.tinst.empty() ? nullptr : out.tinst.data();
  download_into(d);
}

// Device-to-host in chunks through NSTAGE pinned staging buffers: the copy
// engine fills the next ones while the host threads copy a filled one into
// the caller's arrays (pageable, typically never touched: the page faults
// are spread over the threads too).
void Refiner::download_into(const HostDst &d) {
  struct Piece {
    void *dst;
    const char *src;
    size_t bytes;
  };
  std::vector<Piece> pieces;
  if (d.pstat) {
    if (np > o_pcap) {
      dfree(o_pstat);
      o_pcap = np;
      dalloc(&o_pstat, (size_t)o_pcap);
    }
    k_pstatus<<<nblk(np), 256, 0, stream>>>(m.pflag, np, o_pstat);
    HC(hipGetLastError());
  }
  const int skip = d.P_in_host ? std::min(np, d.np_in_host) : 0;
  if (d.P && np > skip)
    pieces.push_back({d.P + 2 * (size_t)skip, (const char *)(m.P + skip),
                      sizeof(double2) * (size_t)(np - skip)});
  if (d.T && nt_out) pieces.push_back({d.T, (const char *)o_T, sizeof(int) * 3 * (size_t)nt_out});
  if (d.N && nt_out) pieces.push_back({d.N, (const char *)o_N, sizeof(int) * 3 * (size_t)nt_out});
  if (d.S && ns) pieces.push_back({d.S, (const char *)m.SV, sizeof(int2) * (size_t)ns});
  if (d.pstat && np) pieces.push_back({d.pstat, (const char *)o_pstat, (size_t)np});
  if (d.tstat && nt_out) pieces.push_back({d.tstat, (const char *)o_tstat, (size_t)nt_out});
  if (d.tinst && m.batched && nt_out)
    pieces.push_back({d.tinst, (const char *)o_inst, sizeof(int) * (size_t)nt_out});
  ensure_stage();
  // the caller's arrays are untouched: ask for 2 MiB pages, so their first
  // touch (in the copies below) faults once per 2 MiB instead of per 4 KiB
  // (a hint; ignored where transparent huge pages are off)
  for (const Piece &p : pieces) {
    const uintptr_t huge = uintptr_t(2) << 20;
    const uintptr_t lo = ((uintptr_t)p.dst + huge - 1) & ~(huge - 1);
    const uintptr_t hi = ((uintptr_t)p.dst + p.bytes) & ~(huge - 1);
    if (hi > lo) (void)madvise((void *)lo, hi - lo, MADV_HUGEPAGE);
  }
  struct Chunk {
    char *dst;
    const char *src;
    size_t bytes;
  };
  std::vector<Chunk> ch;
  for (const Piece &p : pieces)
    for (size_t off = 0; off < p.bytes; off += stage_bytes)
      ch.push_back({(char *)p.dst + off, p.src + off, std::min(stage_bytes, p.bytes - off)});
  auto issue = [&](size_t k) {
    HC(hipMemcpyAsync(h_stage[k % NSTAGE], ch[k].src, ch[k].bytes, hipMemcpyDeviceToHost,
                      stream));
    HC(hipEventRecord(ev_stage[k % NSTAGE], stream));
  };
  for (size_t k = 0; k < std::min<size_t>(NSTAGE, ch.size()); k++) issue(k);
  // the input points need no transfer: copied while the first chunks travel
  if (d.P && skip) copier->copy(d.P, d.P_in_host, sizeof(double) * 2 * (size_t)skip);
  for (size_t k = 0; k < ch.size(); k++) {
    HC(hipEventSynchronize(ev_stage[k % NSTAGE]));
    copier->copy(ch[k].dst, h_stage[k % NSTAGE], ch[k].bytes);
    if (k + NSTAGE < ch.size()) issue(k + NSTAGE);
  }
  HC(hipStreamSynchronize(stream));
}

void Refiner::set_priority_model(const float *host_params, int nparam) {
  if (!mlp_params) dalloc(&mlp_params, (size_t)nparam);
  HC(hipMemcpy(mlp_params, host_params, sizeof(float) * nparam, hipMemcpyHostToDevice));
}

int Refiner::kernel_times(gcdt_kernel_time *rec, int max_rec) const {
  int k = 0;
  for (int id = 0; id < K_NKERN && k < max_rec; id++) {
    std::snprintf(rec[k].name, sizeof(rec[k].name), "%s", kname(id));
    rec[k].launches = ktimes[id].launches;
    rec[k].total_ms = ktimes[id].total_ms;
    rec[k].items = ktimes[id].items;
    k++;
  }
  return k;
}

ParamsD construction_params(double theta_deg, int offcenter) {
  ParamsD pr{};
  double ga = std::cos(theta_deg * 3.141592653589793238462643383279502884197169399375105820974944592308 / 180.0);
  pr.offconst = (ga == 1.0) ? 0.0 : 0.475 * std::sqrt((1.0 + ga) / (1.0 - ga));
  pr.goodangle = ga * ga;
  pr.lens = (2.0 * pr.goodangle - 1.0) * (2.0 * pr.goodangle - 1.0);
  pr.offcenter = offcenter;
  pr.ruppert = 1;
  return pr;
}

void device_constructions(const double *tri6, int n, double theta_deg, int offcenter,
                          double *ccw, double *stp, int *ok) {
  if (n <= 0) return;
  double *dq, *dccw, *dstp;
  int *dok;
  dalloc(&dq, 6 * (size_t)n);
  dalloc(&dccw, (size_t)n);
  dalloc(&dstp, 2 * (size_t)n);
  dalloc(&dok, (size_t)n);
  HC(hipMemcpy(dq, tri6, sizeof(double) * 6 * n, hipMemcpyHostToDevice));
  k_constructions<<<std::min(nblk(n), 256), 256>>>(dq, n, construction_params(theta_deg, offcenter),
                                                   dccw, dstp, dok);
  HC(hipGetLastError());
  HC(hipDeviceSynchronize());
  HC(hipMemcpy(ccw, dccw, sizeof(double) * n, hipMemcpyDeviceToHost));
  HC(hipMemcpy(stp, dstp, sizeof(double) * 2 * n, hipMemcpyDeviceToHost));
  HC(hipMemcpy(ok, dok, sizeof(int) * n, hipMemcpyDeviceToHost));
  dfree(dq); dfree(dccw); dfree(dstp); dfree(dok);
}

void host_constructions(const double *tri6, int n, double theta_deg, int offcenter, double *ccw,
                        double *stp, int *ok) {
  const ParamsD pr = construction_params(theta_deg, offcenter);
  for (int i = 0; i < n; i++) {
    const double *p = tri6 + 6 * (size_t)i;
    ccw[i] = orient2d_value(p[0], p[1], p[2], p[3], p[4], p[5]);
    TriRec r;
    r.X[0] = make_double2(p[0], p[1]);
    r.X[1] = make_double2(p[2], p[3]);
    r.X[2] = make_double2(p[4], p[5]);
    double cx = 0.0, cy = 0.0;
    ok[i] = steiner_point<false>(pr, r, cx, cy);
    stp[2 * (size_t)i] = cx;
    stp[2 * (size_t)i + 1] = cy;
  }
}

void device_predicates(const double *pts8, int n, int pred_mode, int *o,
                       int *ic, long long *stage_counts) {
  double *dq;
  int *dout, *din;
  double *pws;
  unsigned long long *cnt;
  int blocks = std::min(nblk(n), 64);
  dalloc(&dq, 8 * (size_t)n);
  dalloc(&dout, (size_t)n);
  dalloc(&din, (size_t)n);
  dalloc(&pws, (size_t)blocks * 256 * GCDT_INCIRCLE_WS);
  dalloc(&cnt, CNT_N);
  HC(hipMemset(cnt, 0, CNT_N * sizeof(unsigned long long)));
  HC(hipMemcpy(dq, pts8, sizeof(double) * 8 * n, hipMemcpyHostToDevice));
  k_predicates<<<blocks, 256>>>(dq, n, pred_mode, dout, din, pws, cnt);
  HC(hipGetLastError());
  HC(hipDeviceSynchronize());
  HC(hipMemcpy(o, dout, sizeof(int) * n, hipMemcpyDeviceToHost));
  HC(hipMemcpy(ic, din, sizeof(int) * n, hipMemcpyDeviceToHost));
  unsigned long long h[CNT_N];
  HC(hipMemcpy(h, cnt, sizeof(h), hipMemcpyDeviceToHost));
  if (stage_counts) {
    stage_counts[0] = (long long)h[CNT_FP32];
    stage_counts[1] = (long long)h[CNT_FP64];
    stage_counts[2] = (long long)h[CNT_EXACT];
    stage_counts[3] = (long long)h[CNT_DD];
  }
  dfree(dq); dfree(dout); dfree(din); dfree(pws); dfree(cnt);
}

}  // namespace gcdt
