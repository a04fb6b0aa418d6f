// Batched refinement: many independent PSLG instances, each with its own
// theta, insertion policy and mode, refined together.
//
// The reference refines one CDT per GPU_Refine_Quality call
// (refine.cu:6500-7139).  Here the instances are packed into waves; a wave
// is the union of consecutive instances' CDTs (indices offset, one device
// input), with the instance of every triangle in MeshD::tinst and the
// per-instance parameters in an InstPar table.  Cavities never leave an
// instance (instances share no triangle or edge), so one bulk-synchronous
// refinement of the union refines every instance; the candidate cap,
// priority selection and cavity-capacity growth are shared by the wave.
// All waves' inputs stay resident in HBM; one Refiner (one set of work
// buffers, one captured graph loop) runs them in turn.  A wave's output is
// fetched to the host before the next wave reuses the buffers (the last one
// when the results are asked for) and split per instance there.
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "gcdt_internal.h"
#include "gpu_cdt_refine.h"

using gcdt::InstPar;
using gcdt::MeshHost;
using gcdt::Refiner;

namespace {
thread_local std::string g_batch_err;
}  // namespace

struct gcdt_batch {
  Refiner *r = nullptr;
  int n_inst = 0;
  std::vector<int> np, nt, ns;        // input counts per instance
  std::vector<int> wave_first;        // first instance of every wave (+ end)
  std::vector<MeshHost> out;          // per wave, after download
  std::vector<int> onp, ont, ons;     // output counts per instance
  std::vector<long long> op, ot, os;  // output offsets per instance
  std::vector<double> P;              // concatenated per-instance outputs
  std::vector<int> T, N, S;
  std::vector<unsigned char> ps, ts;
  std::vector<double> run_ms;         // host wall time of every wave's run
  bool ran = false, split = false, last_downloaded = true;
  gcdt_stats st{};
};

// Per-instance quality parameters, computed as the single-mesh path does
// (refine.cu:6557-6567 with the off-center constant kept in double).
static InstPar inst_par(double theta, int offcenter, int mode) {
  double ga = std::cos(theta * 3.141592653589793238462643383279502884197169399375105820974944592308 / 180.0);
  InstPar q;
  q.goodangle = ga * ga;
  q.lens = (2.0 * q.goodangle - 1.0) * (2.0 * q.goodangle - 1.0);
  q.offconst = (ga == 1.0) ? 0.0 : 0.475 * std::sqrt((1.0 + ga) / (1.0 - ga));
  q.offcenter = offcenter ? 1 : 0;
  q.ruppert = mode == 1;
  return q;
}

// Split one wave's union output into its instances.  Input points of
// instance i are the wave's points [pin[i], pin[i+1]); Steiner points belong
// to the instance of a triangle that uses them.  Every instance keeps its
// input points first, in order, then its Steiner points in creation order;
// triangles and subsegments keep their relative order.
struct WaveIndex {
  std::vector<int> pinst, plocal, tlocal, cp, ct, cs;
};

static void index_wave(const gcdt_batch *b, int w, WaveIndex &x) {
  const MeshHost &h = b->out[w];
  int i0 = b->wave_first[w], i1 = b->wave_first[w + 1], k = i1 - i0;
  std::vector<int> pin(k + 1, 0);
  for (int i = 0; i < k; i++) pin[i + 1] = pin[i] + b->np[i0 + i];
  int np0 = pin[k];
  x.pinst.assign((size_t)h.np, -1);
  for (int i = 0; i < k; i++)
    for (int p = pin[i]; p < pin[i + 1]; p++) x.pinst[p] = i;
  for (int t = 0; t < h.nt; t++) {
    if (h.tinst[t] < 0 || h.tinst[t] >= k) throw std::runtime_error("batch split: bad triangle instance");
    for (int j = 0; j < 3; j++) {
      int v = h.T[3 * (size_t)t + j];
      if (v >= np0) x.pinst[v] = h.tinst[t];
    }
  }
  x.cp.assign(k, 0);
  x.ct.assign(k, 0);
  x.cs.assign(k, 0);
  x.plocal.assign((size_t)h.np, 0);
  x.tlocal.assign((size_t)h.nt, 0);
  for (int i = 0; i < k; i++) x.cp[i] = b->np[i0 + i];
  for (int p = 0; p < np0; p++) x.plocal[p] = p - pin[x.pinst[p]];
  for (int p = np0; p < h.np; p++) {
    if (x.pinst[p] < 0) throw std::runtime_error("batch split: a Steiner point is in no triangle");
    x.plocal[p] = x.cp[x.pinst[p]]++;
  }
  for (int t = 0; t < h.nt; t++) x.tlocal[t] = x.ct[h.tinst[t]]++;
  for (int s = 0; s < h.ns; s++) x.cs[x.pinst[h.S[2 * (size_t)s]]]++;
}

static void fill_wave(gcdt_batch *b, int w, const WaveIndex &x) {
  const MeshHost &h = b->out[w];
  int i0 = b->wave_first[w];
  for (int p = 0; p < h.np; p++) {
    size_t q = (size_t)(b->op[i0 + x.pinst[p]] + x.plocal[p]);
    b->P[2 * q] = h.P[2 * (size_t)p];
    b->P[2 * q + 1] = h.P[2 * (size_t)p + 1];
    b->ps[q] = h.pstat[p];
  }
  for (int t = 0; t < h.nt; t++) {
    size_t q = (size_t)(b->ot[i0 + h.tinst[t]] + x.tlocal[t]);
    for (int j = 0; j < 3; j++) {
      b->T[3 * q + j] = x.plocal[h.T[3 * (size_t)t + j]];
      int nb = h.N[3 * (size_t)t + j];
      b->N[3 * q + j] = nb < 0 ? -1 : x.tlocal[nb];
    }
    b->ts[q] = h.tstat[t];
  }
  std::vector<int> cs(x.cs.size(), 0);
  for (int s = 0; s < h.ns; s++) {
    int a = h.S[2 * (size_t)s], c = h.S[2 * (size_t)s + 1];
    int i = x.pinst[a];
    size_t q = (size_t)(b->os[i0 + i] + cs[i]++);
    b->S[2 * q] = x.plocal[a];
    b->S[2 * q + 1] = x.plocal[c];
  }
}

static void do_split(gcdt_batch *b) {
  if (b->split) return;
  if (!b->ran) throw std::runtime_error("gcdt_batch_run has not been called");
  int nw = (int)b->out.size();
  if (!b->last_downloaded) {  // the last wave's output is still on the device
    b->r->download(b->out[nw - 1]);
    b->last_downloaded = true;
  }
  b->onp.assign(b->n_inst, 0);
  b->ont.assign(b->n_inst, 0);
  b->ons.assign(b->n_inst, 0);
  std::vector<WaveIndex> idx(nw);
  for (int w = 0; w < nw; w++) {
    index_wave(b, w, idx[w]);
    int i0 = b->wave_first[w];
    for (size_t i = 0; i < idx[w].cp.size(); i++) {
      b->onp[i0 + i] = idx[w].cp[i];
      b->ont[i0 + i] = idx[w].ct[i];
      b->ons[i0 + i] = idx[w].cs[i];
    }
  }
  b->op.assign(b->n_inst + 1, 0);
  b->ot.assign(b->n_inst + 1, 0);
  b->os.assign(b->n_inst + 1, 0);
  for (int i = 0; i < b->n_inst; i++) {
    b->op[i + 1] = b->op[i] + b->onp[i];
    b->ot[i + 1] = b->ot[i] + b->ont[i];
    b->os[i + 1] = b->os[i] + b->ons[i];
  }
  b->P.assign(2 * (size_t)b->op[b->n_inst], 0.0);
  b->ps.assign((size_t)b->op[b->n_inst], 0);
  b->T.assign(3 * (size_t)b->ot[b->n_inst], 0);
  b->N.assign(3 * (size_t)b->ot[b->n_inst], 0);
  b->ts.assign((size_t)b->ot[b->n_inst], 0);
  b->S.assign(2 * (size_t)b->os[b->n_inst], 0);
  for (int w = 0; w < nw; w++) fill_wave(b, w, idx[w]);
  b->split = true;
}

extern "C" {

const char *gcdt_batch_last_error(void) { return g_batch_err.c_str(); }

int gcdt_batch_create(int n_inst, const int *np, const int *nt, const int *ns,
                      const double *points, const int *tris, const int *nbrs,
                      const int *segs, const double *theta_deg,
                      const int *offcenter, const int *mode,
                      const gcdt_options *opt, long long wave_triangles,
                      gcdt_batch **out) {
  g_batch_err.clear();
  *out = nullptr;
  gcdt_batch *b = nullptr;
  try {
    if (n_inst < 1) throw std::runtime_error("a batch needs at least one instance");
    std::string oerr;
    if (!gcdt::options_ok(opt, oerr)) throw std::runtime_error(oerr);
    if (wave_triangles <= 0) wave_triangles = 1LL << 25;
    if (wave_triangles > (1LL << 27)) wave_triangles = 1LL << 27;
    b = new gcdt_batch;
    b->n_inst = n_inst;
    b->np.assign(np, np + n_inst);
    b->nt.assign(nt, nt + n_inst);
    b->ns.assign(ns, ns + n_inst);
    // input offsets and per-instance index validation
    std::vector<long long> pi(n_inst + 1, 0), ti(n_inst + 1, 0), si(n_inst + 1, 0);
    for (int i = 0; i < n_inst; i++) {
      if (np[i] < 3 || nt[i] < 1 || ns[i] < 0) throw std::runtime_error("instance " + std::to_string(i) + " is empty");
      if (!(theta_deg[i] >= 0.0 && theta_deg[i] < 60.0)) throw std::runtime_error("instance theta out of range [0, 60)");
      pi[i + 1] = pi[i] + np[i];
      ti[i + 1] = ti[i] + nt[i];
      si[i + 1] = si[i] + ns[i];
      for (long long q = 3 * ti[i]; q < 3 * ti[i + 1]; q++) {
        if (tris[q] < 0 || tris[q] >= np[i]) throw std::runtime_error("instance triangle vertex index out of range");
        if (nbrs[q] < -1 || nbrs[q] >= nt[i]) throw std::runtime_error("instance neighbour index out of range");
      }
      for (long long q = 2 * si[i]; q < 2 * si[i + 1]; q++)
        if (segs[q] < 0 || segs[q] >= np[i]) throw std::runtime_error("instance segment vertex index out of range");
    }
    // waves of consecutive instances, at most wave_triangles input triangles
    // each (a larger instance is a wave by itself)
    b->wave_first.push_back(0);
    long long acc = 0;
    for (int i = 0; i < n_inst; i++) {
      if (acc > 0 && acc + nt[i] > wave_triangles) {
        b->wave_first.push_back(i);
        acc = 0;
      }
      acc += nt[i];
    }
    b->wave_first.push_back(n_inst);
    b->r = new Refiner(*opt);
    int nw = (int)b->wave_first.size() - 1;
    for (int w = 0; w < nw; w++) {
      int i0 = b->wave_first[w], i1 = b->wave_first[w + 1];
      long long P0 = pi[i0], T0 = ti[i0], S0 = si[i0];
      long long wnp = pi[i1] - P0, wnt = ti[i1] - T0, wns = si[i1] - S0;
      std::vector<int> U((size_t)(3 * wnt)), V((size_t)(3 * wnt)), W((size_t)(2 * wns));
      std::vector<int> I((size_t)wnt);
      std::vector<InstPar> ip(i1 - i0);
      for (int i = i0; i < i1; i++) {
        int poff = (int)(pi[i] - P0), toff = (int)(ti[i] - T0);
        for (long long q = 3 * ti[i]; q < 3 * ti[i + 1]; q++) {
          U[q - 3 * T0] = tris[q] + poff;
          V[q - 3 * T0] = nbrs[q] < 0 ? -1 : nbrs[q] + toff;
        }
        for (long long q = 2 * si[i]; q < 2 * si[i + 1]; q++) W[q - 2 * S0] = segs[q] + poff;
        for (long long t = ti[i]; t < ti[i + 1]; t++) I[t - T0] = i - i0;
        ip[i - i0] = inst_par(theta_deg[i], offcenter[i], mode[i]);
      }
      b->r->add_input(points + 2 * P0, (int)wnp, U.data(), V.data(), (int)wnt, W.data(),
                      (int)wns, I.data(), ip.data(), i1 - i0);
    }
    *out = b;
    return 0;
  } catch (const std::exception &e) {
    g_batch_err = e.what();
    if (b) {
      delete b->r;
      delete b;
    }
    return -2;
  }
}

int gcdt_batch_waves(const gcdt_batch *b) { return (int)b->wave_first.size() - 1; }

// Host wall time (ms) of the last gcdt_batch_run's refinement, without the
// downloads between waves.
double gcdt_batch_run_ms(const gcdt_batch *b) {
  double t = 0.0;
  for (double x : b->run_ms) t += x;
  return t;
}

// Severity order of the status codes: converged, converged with given-up
// elements, round cap, stalled.
static int status_rank(int s) { return s == 0 ? 0 : s == 3 ? 1 : s == 1 ? 2 : 3; }

// Refine every wave.  stats: sums over waves (rounds and final_cav_cap: the
// largest; status: the most severe code).
int gcdt_batch_run(gcdt_batch *b, gcdt_stats *stats) {
  g_batch_err.clear();
  try {
    int nw = (int)b->wave_first.size() - 1;
    b->out.assign(nw, MeshHost());
    b->split = false;
    gcdt_stats tot{};
    b->run_ms.assign(nw, 0.0);
    for (int w = 0; w < nw; w++) {
      auto t0 = std::chrono::steady_clock::now();
      b->r->select_input(w);
      b->r->run();
      const gcdt_stats &s = b->r->stats();
      tot.rounds = std::max(tot.rounds, s.rounds);
      tot.seg_passes += s.seg_passes;
      tot.bad_passes += s.bad_passes;
      tot.inserted += s.inserted;
      tot.seg_splits += s.seg_splits;
      tot.rej_encroach += s.rej_encroach;
      tot.rej_blocked += s.rej_blocked;
      tot.rej_overflow += s.rej_overflow;
      tot.lost += s.lost;
      tot.giveups += s.giveups;
      tot.final_cav_cap = std::max(tot.final_cav_cap, s.final_cav_cap);
      if (status_rank(s.status) > status_rank(tot.status)) tot.status = s.status;
      tot.giveup_tris_left += s.giveup_tris_left;
      tot.giveup_segs_left += s.giveup_segs_left;
      tot.device_ms += s.device_ms;
      tot.pred_fp32 += s.pred_fp32;
      tot.pred_fp64 += s.pred_fp64;
      tot.pred_exact += s.pred_exact;
      tot.pred_dd += s.pred_dd;
      tot.n_points += s.n_points;
      tot.n_triangles += s.n_triangles;
      tot.n_subsegs += s.n_subsegs;
      tot.cand_evals += s.cand_evals;
      tot.big_cavities += s.big_cavities;
      tot.tri_examined_seg += s.tri_examined_seg;
      tot.tri_examined_bad += s.tri_examined_bad;
      tot.ws_ints_seg += s.ws_ints_seg;
      tot.ws_ints_bad += s.ws_ints_bad;
      tot.deleted += s.deleted;
      tot.flips += s.flips;
      tot.growths += s.growths;
      tot.grow_in_place = s.grow_in_place;
      b->run_ms[w] = std::chrono::duration<double, std::milli>(
                         std::chrono::steady_clock::now() - t0).count();
      // the Refiner's buffers are reused by the next wave: fetch this one
      if (w + 1 < nw) b->r->download(b->out[w]);
    }
    b->ran = true;
    b->last_downloaded = false;
    b->st = tot;
    if (stats) *stats = tot;
    return 0;
  } catch (const std::exception &e) {
    g_batch_err = e.what();
    return -2;
  }
}

int gcdt_batch_counts(gcdt_batch *b, int *np, int *nt, int *ns) {
  g_batch_err.clear();
  try {
    do_split(b);
    for (int i = 0; i < b->n_inst; i++) {
      if (np) np[i] = b->onp[i];
      if (nt) nt[i] = b->ont[i];
      if (ns) ns[i] = b->ons[i];
    }
    return 0;
  } catch (const std::exception &e) {
    g_batch_err = e.what();
    return -2;
  }
}

int gcdt_batch_copy(gcdt_batch *b, double *points, int *tris, int *nbrs,
                    int *segs, unsigned char *pstatus, unsigned char *tstatus) {
  g_batch_err.clear();
  try {
    do_split(b);
    if (points) std::memcpy(points, b->P.data(), sizeof(double) * b->P.size());
    if (tris) std::memcpy(tris, b->T.data(), sizeof(int) * b->T.size());
    if (nbrs) std::memcpy(nbrs, b->N.data(), sizeof(int) * b->N.size());
    if (segs) std::memcpy(segs, b->S.data(), sizeof(int) * b->S.size());
    if (pstatus) std::memcpy(pstatus, b->ps.data(), b->ps.size());
    if (tstatus) std::memcpy(tstatus, b->ts.data(), b->ts.size());
    return 0;
  } catch (const std::exception &e) {
    g_batch_err = e.what();
    return -2;
  }
}

int gcdt_batch_set_priority_model(gcdt_batch *b, const float *params) {
  g_batch_err.clear();
  try {
    b->r->set_priority_model(params, gcdt_mlp_nparam());
    return 0;
  } catch (const std::exception &e) {
    g_batch_err = e.what();
    return -2;
  }
}

int gcdt_batch_autotune(gcdt_batch *b, int reps, gcdt_options *best,
                        gcdt_tune_record *rec, int max_rec) {
  g_batch_err.clear();
  try {
    gcdt_options o = b->r->opt;
    int n = gcdt::autotune(o, reps, [b](const gcdt_options &x) {
      b->r->opt = x;
      gcdt_stats st;
      if (gcdt_batch_run(b, &st) != 0) throw std::runtime_error(g_batch_err);
      return gcdt_batch_run_ms(b);
    }, rec, max_rec);
    b->r->opt = o;
    b->ran = false;
    b->split = false;
    if (best) *best = o;
    return n;
  } catch (const std::exception &e) {
    g_batch_err = e.what();
    return -2;
  }
}

void gcdt_batch_destroy(gcdt_batch *b) {
  if (!b) return;
  delete b->r;
  delete b;
}

}  // extern "C"
