// C ABI of libgcdt.so and the reference-compatible GPU_Refine_Quality.
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <string>
#include <vector>
#include <algorithm>

#include "gcdt_internal.h"
#include "gpu_cdt_refine.h"
#include "priority_mlp_weights.h"

using gcdt::MeshHost;
using gcdt::Refiner;

static thread_local std::string g_err;

struct gcdt_mesh {
  MeshHost h;
};

struct gcdt_session {
  Refiner *r;
};

extern "C" {

const char *gcdt_last_error(void) { return g_err.c_str(); }

int gcdt_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

void gcdt_default_options(gcdt_options *o) {
  std::memset(o, 0, sizeof(*o));
  o->theta_deg = 25.0;
  o->offcenter = 1;
  o->mode = 1;
  o->cav_cap0 = 32;
  o->cav_cap_max = 2048;
  o->cand_budget = 1LL << 30;
  o->max_rounds = 100000;
  o->priority = 1;
  o->pred_mode = 0;
  o->profile = 0;
  o->device = 0;
  o->cand_factor = 4;
  o->cand_min = 16384;
  o->claim_iters = 1;
  o->log_training = 0;
  o->graph = 0;
  o->graph_seg_passes = 2;
  o->graph_rounds = 4;
  o->grid_cavity = 8192;
  o->grid_claim = 2048;
  o->grid_apply = 4096;
  o->cavity_lanes = 16;
  o->cav_lds_cap = 32;
  o->learned_shift = 2;
  o->mixed = 0;
  o->shared_ring = 0;
  o->mix_below = 0;
  o->deletion = 1;
  o->seg_claim_iters = 0;
  o->vmm = 1;
}

}  // extern "C"

namespace gcdt {
bool options_ok(const gcdt_options *o, std::string &g_err) {
  if (!o) { g_err = "options is NULL"; return false; }
  if (!(o->theta_deg >= 0.0 && o->theta_deg < 60.0)) { g_err = "theta out of range [0, 60)"; return false; }
  if (o->cav_cap0 < 4 || o->cav_cap_max < o->cav_cap0 || o->cav_cap_max > 2048) {
    g_err = "cavity capacity must satisfy 4 <= cav_cap0 <= cav_cap_max <= 2048";
    return false;
  }
  if (o->cand_budget < 6LL * o->cav_cap_max + 16) { g_err = "cand_budget smaller than one candidate at cav_cap_max"; return false; }
  if (o->priority < 0 || o->priority > 2) { g_err = "priority must be 0, 1 or 2"; return false; }
  if (o->claim_iters < 1 || o->claim_iters > 64) { g_err = "claim_iters must be in [1, 64]"; return false; }
  if (o->seg_claim_iters < 0 || o->seg_claim_iters > 64) { g_err = "seg_claim_iters must be in [0, 64]"; return false; }
  if (o->graph_seg_passes < 1 || o->graph_seg_passes > 16 || o->graph_rounds < 1 || o->graph_rounds > 64) {
    g_err = "graph_seg_passes must be in [1, 16] and graph_rounds in [1, 64]";
    return false;
  }
  if (o->grid_cavity < 1 || o->grid_cavity > 16384 || o->grid_claim < 1 || o->grid_claim > 65536 ||
      o->grid_apply < 1 || o->grid_apply > 65536) {
    g_err = "grid sizes must be in [1, 16384] (cavity) and [1, 65536] (claim, apply)";
    return false;
  }
  if (o->cavity_lanes != 8 && o->cavity_lanes != 16 && o->cavity_lanes != 32 && o->cavity_lanes != 64) {
    g_err = "cavity_lanes must be 8, 16, 32 or 64";
    return false;
  }
  if (o->learned_shift < 0 || o->learned_shift > 12) {
    g_err = "learned_shift must be in [0, 12]";
    return false;
  }
  if (o->cav_lds_cap < 4 || o->cav_lds_cap > 2048) {
    g_err = "cav_lds_cap must be in [4, 2048]";
    return false;
  }
  // The graph loop resets the 11-bit claim epoch only between graph
  // launches; one launch uses claim_iters + 2 epochs per pass (claims, the
  // winners' stamp, the vertex deletion step).
  if (o->graph && (long long)o->graph_rounds * (o->graph_seg_passes + 1) *
                          (std::max(o->claim_iters, o->seg_claim_iters) + 2) + 2 >
                      2047) {
    g_err = "graph_rounds * (graph_seg_passes + 1) * (claim_iters + 2) + 2 must not exceed 2047 "
            "(claim epochs of one graph launch)";
    return false;
  }
  if (o->deletion < 0 || o->deletion > 2) { g_err = "deletion must be 0, 1 or 2"; return false; }
  if (o->vmm != 0 && o->vmm != 1) { g_err = "vmm must be 0 or 1"; return false; }
  if (o->graph && o->log_training) { g_err = "log_training needs the host-driven loop (graph = 0)"; return false; }
  if (o->mixed != 0 && o->mixed != 1) { g_err = "mixed must be 0 or 1"; return false; }
  if (o->shared_ring != 0 && o->shared_ring != 1) { g_err = "shared_ring must be 0 or 1"; return false; }
  if (o->mix_below < 0) { g_err = "mix_below must be >= 0"; return false; }
  if (o->mix_below > 0 && o->graph) { g_err = "mix_below needs the host-driven loop (graph = 0)"; return false; }
  if (o->mixed && o->graph) { g_err = "mixed rounds run in the host-driven loop (graph = 0)"; return false; }
  if (o->mixed && o->log_training) { g_err = "log_training needs separate segment and bad passes (mixed = 0)"; return false; }
  return true;
}
}  // namespace gcdt

static bool valid_options(const gcdt_options *o) { return gcdt::options_ok(o, g_err); }

extern "C" {

int gcdt_refine_host(const double *points, int np, const int *tris,
                     const int *nbrs, int nt, const int *segs, int ns,
                     const gcdt_options *opt, gcdt_mesh **out,
                     gcdt_stats *stats) {
  g_err.clear();
  if (!valid_options(opt)) return -1;
  try {
    Refiner r(*opt);
    r.upload(points, np, tris, nbrs, nt, segs, ns);
    r.run();
    gcdt_mesh *m = new gcdt_mesh;
    r.download(m->h);
    if (stats) *stats = r.stats();
    *out = m;
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -2;
  }
}

void gcdt_mesh_counts(const gcdt_mesh *m, int *np, int *nt, int *ns) {
  *np = m->h.np;
  *nt = m->h.nt;
  *ns = m->h.ns;
}

void gcdt_mesh_copy(const gcdt_mesh *m, double *points, int *tris, int *nbrs,
                    int *segs, unsigned char *pstatus, unsigned char *tstatus) {
  const MeshHost &h = m->h;
  if (points) std::memcpy(points, h.P.data(), sizeof(double) * h.P.size());
  if (tris) std::memcpy(tris, h.T.data(), sizeof(int) * h.T.size());
  if (nbrs) std::memcpy(nbrs, h.N.data(), sizeof(int) * h.N.size());
  if (segs) std::memcpy(segs, h.S.data(), sizeof(int) * h.S.size());
  if (pstatus) std::memcpy(pstatus, h.pstat.data(), h.pstat.size());
  if (tstatus) std::memcpy(tstatus, h.tstat.data(), h.tstat.size());
}

void gcdt_mesh_free(gcdt_mesh *m) { delete m; }

gcdt_session *gcdt_session_create(const double *points, int np,
                                  const int *tris, const int *nbrs, int nt,
                                  const int *segs, int ns,
                                  const gcdt_options *opt) {
  g_err.clear();
  if (!valid_options(opt)) return nullptr;
  try {
    gcdt_session *s = new gcdt_session;
    s->r = new Refiner(*opt);
    s->r->upload(points, np, tris, nbrs, nt, segs, ns);
    return s;
  } catch (const std::exception &e) {
    g_err = e.what();
    return nullptr;
  }
}

int gcdt_session_run(gcdt_session *s, gcdt_stats *stats) {
  g_err.clear();
  try {
    s->r->run();
    if (stats) *stats = s->r->stats();
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -2;
  }
}

int gcdt_session_check(gcdt_session *s, long long *viol, long long *exact_signs,
                       long long *n_exempt, double *min_angle_deg) {
  g_err.clear();
  try {
    if (!s || !viol) throw std::runtime_error("session and viol must be non-NULL");
    if (s->r->input_count() != 1) throw std::runtime_error("session holds no single refined mesh");
    gcdt::CheckExtra ex;
    gcdt::check_device_mesh(s->r->device_mesh(), s->r->stream, s->r->opt.theta_deg, viol,
                            exact_signs, &ex);
    if (n_exempt) *n_exempt = ex.n_exempt;
    if (min_angle_deg) *min_angle_deg = ex.min_angle_deg;
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -2;
  }
}

int gcdt_build_cdt_device(const double *points, int np, const int *segs, int ns,
                          gcdt_cdt **out, double *stats) {
  g_err.clear();
  try {
    if (!out || (np > 0 && !points) || (ns > 0 && !segs) || np < 0 || ns < 0)
      throw std::runtime_error("invalid arrays");
    gcdt_cdt *c = new gcdt_cdt;
    c->np = np;
    try {
      gcdt::build_cdt_device(points, np, segs, ns, c->T, c->N, c->S, c->stats);
    } catch (...) {
      delete c;
      throw;
    }
    if (stats)
      for (int i = 0; i < 4; i++) stats[i] = c->stats[i];
    *out = c;
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -2;
  }
}

int gcdt_check_mesh_device_noalloc(const double *d_points, int np, const int *d_tris,
                                   const int *d_nbrs, int nt, const int *d_segs, int ns,
                                   const double *d_points_in, int np_in, const int *d_segs_in,
                                   int ns_in, double theta_deg, long long *d_out,
                                   void *stream) {
  g_err.clear();
  try {
    if (!d_out || np < 0 || nt < 0 || ns < 0 || np_in < 0 || ns_in < 0 ||
        (np > 0 && !d_points) || (nt > 0 && (!d_tris || !d_nbrs)) || (ns > 0 && !d_segs) ||
        (np_in > 0 && !d_points_in))
      throw std::runtime_error("invalid mesh arrays");
    if (d_segs_in && !(theta_deg > 0.0 && theta_deg < 60.0))
      throw std::runtime_error("theta must lie in (0, 60) degrees for the quality check");
    gcdt::Refiner::DeviceMesh dm{(const double2 *)d_points, np, d_tris, d_nbrs, nt,
                                 (const int2 *)d_segs, ns, (const double2 *)d_points_in, np_in,
                                 d_segs_in ? (const int2 *)d_segs_in : nullptr,
                                 d_segs_in ? ns_in : 0};
    gcdt::check_device_mesh_noalloc(dm, (hipStream_t)stream, theta_deg, d_out);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -2;
  }
}

int gcdt_check_mesh(const double *points, int np, const int *tris, const int *nbrs, int nt,
                    const int *segs, int ns, const double *points_in, int np_in,
                    const int *segs_in, int ns_in, double theta_deg, long long *viol,
                    long long *exact_signs, long long *n_exempt, double *min_angle_deg) {
  g_err.clear();
  void *d[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  hipStream_t stream = nullptr;
  try {
    if (!viol || np < 0 || nt < 0 || ns < 0 || np_in < 0 || ns_in < 0 || (np > 0 && !points) ||
        (nt > 0 && (!tris || !nbrs)) || (ns > 0 && !segs) || (np_in > 0 && !points_in))
      throw std::runtime_error("invalid mesh arrays");
    if (segs_in && !(theta_deg > 0.0 && theta_deg < 60.0))
      throw std::runtime_error("theta must lie in (0, 60) degrees for the quality check");
    auto up = [&](int k, const void *src, size_t bytes) {
      if (hipMalloc(&d[k], bytes ? bytes : 1) != hipSuccess) throw std::runtime_error("hipMalloc failed");
      if (bytes && hipMemcpy(d[k], src, bytes, hipMemcpyHostToDevice) != hipSuccess)
        throw std::runtime_error("upload failed");
    };
    up(0, points, sizeof(double) * 2 * (size_t)np);
    up(1, tris, sizeof(int) * 3 * (size_t)nt);
    up(2, nbrs, sizeof(int) * 3 * (size_t)nt);
    up(3, segs, sizeof(int) * 2 * (size_t)ns);
    up(4, points_in, sizeof(double) * 2 * (size_t)np_in);
    if (segs_in) up(5, segs_in, sizeof(int) * 2 * (size_t)ns_in);
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess)
      throw std::runtime_error("stream creation failed");
    gcdt::Refiner::DeviceMesh dm{(const double2 *)d[0], np, (const int *)d[1], (const int *)d[2], nt,
                                 (const int2 *)d[3], ns, (const double2 *)d[4], np_in,
                                 segs_in ? (const int2 *)d[5] : nullptr, segs_in ? ns_in : 0};
    gcdt::CheckExtra ex;
    gcdt::check_device_mesh(dm, stream, theta_deg, viol, exact_signs, &ex);
    if (n_exempt) *n_exempt = ex.n_exempt;
    if (min_angle_deg) *min_angle_deg = ex.min_angle_deg;
  } catch (const std::exception &e) {
    g_err = e.what();
    for (void *p : d) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
    return -2;
  }
  for (void *p : d) (void)hipFree(p);
  (void)hipStreamDestroy(stream);
  return 0;
}

int gcdt_session_download(gcdt_session *s, gcdt_mesh **out) {
  g_err.clear();
  try {
    gcdt_mesh *m = new gcdt_mesh;
    s->r->download(m->h);
    *out = m;
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -2;
  }
}

int gcdt_session_set_priority_model(gcdt_session *s, const float *params) {
  g_err.clear();
  try {
    s->r->set_priority_model(params, gcdt_mlp_nparam());
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -2;
  }
}

long long gcdt_session_logs(gcdt_session *s, float *X, float *t, long long max_n) {
  long long n = (long long)s->r->log_t.size();
  long long k = std::min(n, max_n);
  if (X && k > 0) std::memcpy(X, s->r->log_X.data(), sizeof(float) * 16 * (size_t)k);
  if (t && k > 0) std::memcpy(t, s->r->log_t.data(), sizeof(float) * (size_t)k);
  return n;
}

int gcdt_session_kernel_times(gcdt_session *s, gcdt_kernel_time *rec,
                              int max_rec) {
  return s->r->kernel_times(rec, max_rec);
}

int gcdt_session_autotune(gcdt_session *s, int reps, gcdt_options *best,
                          gcdt_tune_record *rec, int max_rec) {
  g_err.clear();
  try {
    Refiner *r = s->r;
    gcdt_options o = r->opt;
    int n = gcdt::autotune(o, reps, [r](const gcdt_options &x) {
      r->opt = x;
      r->run();
      return r->stats().device_ms;
    }, rec, max_rec);
    r->opt = o;
    if (best) *best = o;
    return n;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -2;
  }
}

void gcdt_session_destroy(gcdt_session *s) {
  if (!s) return;
  delete s->r;
  delete s;
}

int gcdt_device_predicates(const double *pts8, int n, int pred_mode,
                           int *orient_out, int *incircle_out,
                           long long *stage_counts) {
  g_err.clear();
  try {
    gcdt::device_predicates(pts8, n, pred_mode, orient_out, incircle_out,
                            stage_counts);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -2;
  }
}

int gcdt_host_constructions(const double *tri6, int n, double theta_deg, int offcenter,
                            double *ccw_out, double *steiner_out, int *ok_out) {
  g_err.clear();
  if (n < 0 || (n > 0 && (!tri6 || !ccw_out || !steiner_out || !ok_out))) {
    g_err = "bad arguments";
    return -1;
  }
  gcdt::host_constructions(tri6, n, theta_deg, offcenter, ccw_out, steiner_out, ok_out);
  return 0;
}

int gcdt_device_constructions(const double *tri6, int n, double theta_deg, int offcenter,
                              double *ccw_out, double *steiner_out, int *ok_out) {
  g_err.clear();
  if (n < 0 || (n > 0 && (!tri6 || !ccw_out || !steiner_out || !ok_out))) {
    g_err = "bad arguments";
    return -1;
  }
  try {
    gcdt::device_constructions(tri6, n, theta_deg, offcenter, ccw_out, steiner_out, ok_out);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -2;
  }
}

}  // extern "C"

// Drop-in entry point with the reference's signature (refine.h:385-386).
// ---------------------------------------------------------------------------
// The drop-in's process-wide state: its options (the chosen form unless a
// caller sets others), the priority network, and one cached refiner whose
// device buffers serve every call with the same options.
// ---------------------------------------------------------------------------
namespace {
struct DropIn {
  std::mutex mu;
  bool configured = false;
  gcdt_options opt{};
  std::vector<float> model;
  std::unique_ptr<Refiner> r;
  gcdt_stats last{};
};
// Never destroyed: freeing device memory from a static destructor could run
// after the HIP runtime has shut down (gcdt_dropin_release frees it early).
DropIn &dropin() {
  static DropIn *d = new DropIn;
  return *d;
}
// The chosen form: FP32 filter, three claim iterations with shared rings,
// a bad pass taking up to 3x the previous pass's winners, the learned
// priority (coarsened by one half-octave), segment phases that hand their
// last few (at most 4) encroached subsegments to the bad pass; launch shapes
// the autotuner picked on the 10^6-point workload (they change timing only,
// never the mesh).
void chosen_defaults(gcdt_options *o) {
  gcdt_default_options(o);
  o->pred_mode = 1;
  o->claim_iters = 3;
  o->seg_claim_iters = 2;
  o->mix_below = 4;
  o->shared_ring = 1;
  o->cand_factor = 3;
  o->priority = 2;
  o->learned_shift = 1;
  o->grid_cavity = 4096;
  o->grid_claim = 2048;
  o->grid_apply = 2048;
  o->cavity_lanes = 16;
  o->cav_lds_cap = 32;
}
void dropin_configure(DropIn &d) {
  if (d.configured) return;
  chosen_defaults(&d.opt);
  d.model.assign(gcdt::kEmbeddedMlp, gcdt::kEmbeddedMlp + gcdt::kEmbeddedMlpCount);
  d.configured = true;
}
// same options apart from the per-call angle, policy and mode
bool same_run_options(const gcdt_options &a, const gcdt_options &b) {
  gcdt_options x = a, y = b;
  x.theta_deg = y.theta_deg = 0.0;
  x.offcenter = y.offcenter = 0;
  x.mode = y.mode = 0;
  return std::memcmp(&x, &y, sizeof(x)) == 0;
}
}  // namespace

extern "C" {

int gcdt_dropin_set_options(const gcdt_options *opt) {
  g_err.clear();
  DropIn &d = dropin();
  std::lock_guard<std::mutex> lk(d.mu);
  dropin_configure(d);
  gcdt_options o;
  if (opt) o = *opt;
  else chosen_defaults(&o);
  o.theta_deg = 25.0;  // validated with a representative angle; calls set their own
  if (!valid_options(&o)) return -1;
  d.opt = o;
  d.r.reset();
  return 0;
}

int gcdt_dropin_set_priority_model(const float *params) {
  g_err.clear();
  DropIn &d = dropin();
  std::lock_guard<std::mutex> lk(d.mu);
  dropin_configure(d);
  if (params) d.model.assign(params, params + gcdt::kEmbeddedMlpCount);
  else d.model.assign(gcdt::kEmbeddedMlp, gcdt::kEmbeddedMlp + gcdt::kEmbeddedMlpCount);
  if (d.r && d.r->opt.priority == 2) d.r->set_priority_model(d.model.data(), (int)d.model.size());
  return 0;
}

int gcdt_dropin_last_stats(gcdt_stats *stats) {
  if (!stats) return -1;
  DropIn &d = dropin();
  std::lock_guard<std::mutex> lk(d.mu);
  *stats = d.last;
  return 0;
}

void gcdt_dropin_release(void) {
  DropIn &d = dropin();
  std::lock_guard<std::mutex> lk(d.mu);
  d.r.reset();
}

void gcdt_dropin_free(struct triangulateio *result, PStatus *ps_debug, TStatus *ts_debug) {
  if (result) {
    delete[] result->pointlist;
    delete[] result->trianglelist;
    delete[] result->neighborlist;
    delete[] result->segmentlist;
    result->pointlist = nullptr;
    result->trianglelist = nullptr;
    result->neighborlist = nullptr;
    result->segmentlist = nullptr;
  }
  delete[] ps_debug;
  delete[] ts_debug;
}

}  // extern "C"

void GPU_Refine_Quality(struct triangulateio *input,
                        struct triangulateio *result, double theta,
                        InsertPolicy insertpolicy, DeletePolicy deletepolicy,
                        int mode, int debug_iter, PStatus **ps_debug,
                        TStatus **ts_debug) {
  (void)debug_iter;
  DropIn &d = dropin();
  std::lock_guard<std::mutex> lk(d.mu);
  dropin_configure(d);
  gcdt_options opt = d.opt;
  opt.theta_deg = theta;
  opt.offcenter = insertpolicy == Offcenter ? 1 : 0;
  opt.mode = mode;
  if (deletepolicy == noDeletion) opt.deletion = 0;
  else if (deletepolicy == DeleteFromDT) opt.deletion = 2;
  g_err.clear();
  if (!input || !result || !valid_options(&opt)) {
    std::fprintf(stderr, "GPU_Refine_Quality: %s\n",
                 (!input || !result) ? "input and result must be non-NULL" : g_err.c_str());
    std::exit(EXIT_FAILURE);
  }
  try {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("no HIP device");
    if (d.r && (!same_run_options(d.r->opt, opt) || d.r->opt.device != dev)) d.r.reset();
    if (!d.r) {
      opt.device = dev;
      d.r.reset(new Refiner(opt));
      if (opt.priority == 2) d.r->set_priority_model(d.model.data(), (int)d.model.size());
    }
    Refiner &r = *d.r;
    r.set_angle(opt.theta_deg, opt.offcenter, opt.mode);
    r.upload(input->pointlist, input->numberofpoints, input->trianglelist, input->neighborlist,
             input->numberoftriangles, input->segmentlist, input->numberofsegments);
    r.run();
    d.last = r.stats();
    const int np = r.out_np(), nt = r.out_nt(), ns = r.out_ns();
    result->numberofpoints = np;
    result->numberoftriangles = nt;
    result->numberofcorners = 3;
    result->numberofsegments = ns;
    result->numberofflips = (int)std::min<long long>(d.last.flips, INT_MAX);
    result->pointlist = new double[2 * (size_t)np];
    result->trianglelist = new int[3 * (size_t)nt];
    result->neighborlist = new int[3 * (size_t)nt];
    result->segmentlist = new int[2 * (size_t)ns];
    Refiner::HostDst dst;
    dst.P = result->pointlist;
    dst.T = result->trianglelist;
    dst.N = result->neighborlist;
    dst.S = result->segmentlist;
    dst.P_in_host = input->pointlist;
    dst.np_in_host = input->numberofpoints;
    if (ps_debug) {
      *ps_debug = new PStatus[(size_t)np];
      dst.pstat = reinterpret_cast<unsigned char *>(*ps_debug);
    }
    if (ts_debug) {
      *ts_debug = new TStatus[(size_t)nt];
      dst.tstat = reinterpret_cast<unsigned char *>(*ts_debug);
    }
    r.download_into(dst);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "GPU_Refine_Quality: %s\n", e.what());
    std::exit(EXIT_FAILURE);
  }
}
