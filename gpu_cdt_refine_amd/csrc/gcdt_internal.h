// Host-side refinement driver (internal to libgcdt.so).
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "gcdt_device.h"
#include "gpu_cdt_refine.h"
#include "host_copy.h"
#include "vm_range.h"

#include "gcdt_cdt_result.h"

namespace gcdt {

// Device CDT construction (gcdt_cdt.hip); stats: insertion rounds, segment
// passes, segment pieces, device milliseconds.
void build_cdt_device(const double *points, int np, const int *segs, int ns,
                      std::vector<int> &T_out, std::vector<int> &N_out, std::vector<int> &S_out,
                      double *stats);

struct MeshHost {
  int np = 0, nt = 0, ns = 0;
  std::vector<double> P;
  std::vector<int> T, N, S;
  std::vector<unsigned char> pstat, tstat;
  std::vector<int> tinst;  // batched refinement: instance of every triangle
};

// One device-resident input CDT (a single mesh, or the union of a batch).
struct DelD;  // vertex deletion step of a pass (gcdt_refine.hip)

struct InputSet {
  int np = 0, nt = 0, ns = 0, n_inst = 0;
  int capP = 0, capT = 0, capS = 0;  // allocated sizes (reuse on upload)
  double2 *P = nullptr;
  int *T = nullptr, *N = nullptr, *S = nullptr;
  int *tinst = nullptr;     // instance of every input triangle (batches)
  InstPar *ipar = nullptr;  // per-instance parameters (batches)
  bool any_chew = false;    // a batch instance refines in Chew mode
};

enum KernelId {
  K_INIT = 0,
  K_MARKBAD,
  K_COMPACT,
  K_SEGCAV,
  K_BADCAV,
  K_CLAIM,
  K_WIN,
  K_APPLY,
  K_EXPORT,
  K_MLP,
  K_DEL_SCAN,
  K_DEL_APPLY,
  K_NKERN
};

class Refiner {
 public:
  explicit Refiner(const gcdt_options &o);
  ~Refiner();
  void upload(const double *points, int np, const int *tris, const int *nbrs,
              int nt, const int *segs, int ns);
  // Resident inputs: add_input uploads one (for a batch, the union of
  // independent meshes with tri_inst[nt] the instance of every triangle and
  // ipar[n_inst] their parameters); run() refines the selected one.
  int add_input(const double *points, int np, const int *tris, const int *nbrs,
                int nt, const int *segs, int ns, const int *tri_inst,
                const InstPar *ipar, int n_inst);
  void select_input(int k);
  // theta / insertion policy / mode of the next run (a cached refiner serves
  // calls with different angles)
  void set_angle(double theta_deg, int offcenter, int mode);
  int input_count() const { return (int)inputs.size(); }
  // device copies of the last run's output (for batches kept across waves)
  const double2 *out_points() const { return m.P; }
  int out_np() const { return np; }
  // the last run's exported mesh in device memory (triangles / neighbours in
  // the output numbering, points, subsegments) and the first input's points
  struct DeviceMesh {
    const double2 *P;
    int np;
    const int *T, *N;
    int nt;
    const int2 *S;
    int ns;
    const double2 *P_in;
    int np_in;
    const int2 *S_in;  // input segments (coverage and quality); may be null
    int ns_in;
  };
  DeviceMesh device_mesh() const;
  void run();
  void download(MeshHost &out);
  // The last run's mesh straight into caller arrays (any may be null):
  // points 2*np doubles, triangles / neighbours 3*nt ints, subsegments 2*ns
  // ints, point / triangle status bytes, triangle instances.  The first
  // np_in_host points are the input points and are copied from the host
  // array P_in_host when it is given.
  struct HostDst {
    double *P = nullptr;
    int *T = nullptr, *N = nullptr, *S = nullptr;
    unsigned char *pstat = nullptr, *tstat = nullptr;
    int *tinst = nullptr;
    const double *P_in_host = nullptr;
    int np_in_host = 0;
  };
  void download_into(const HostDst &d);
  int out_nt() const { return nt_out; }
  int out_ns() const { return ns; }
  int kernel_times(gcdt_kernel_time *rec, int max_rec) const;
  const gcdt_stats &stats() const { return st; }

  gcdt_options opt;
  ParamsD pr;
  MeshD m{};
  hipStream_t stream = nullptr;
  int np = 0, nt = 0, ns = 0;
  // learned-priority extension
  float *mlp_params = nullptr;      // device FP32 master weights (priority 2)
  void set_priority_model(const float *host_params, int nparam);
  // training logs (options.log_training): features and conflict-degree labels
  std::vector<float> log_X, log_t;

 private:
  void ensure_stage();
  void upload_staged(void *dst, const void *src, size_t bytes, size_t &k);
  void copy_input(InputSet &in, const double *points, int np, const int *tris, const int *nbrs,
                  int nt, const int *segs, int ns, const int *tri_inst, const InstPar *ipar,
                  int n_inst);
  template <class Pred>
  int compact(Pred pred, int n, int *out_list, int *out_rank, const int *src,
              unsigned long long *total64 = nullptr, bool counted = false);
  void ensure_cb(int nb);
  void flush_counters();
  int run_pass(const int *ids, int nc, int nseg, long long *progress, bool *overflowed,
               int *nwin_seg, bool list_enc = false);
  DelD del_base() const;
  void enqueue_deletion(const DelD &dd);
  void compact_points();
  PassD pass_base(int Kc) const;
  void enqueue_claims(PassD ps, int gc, int iters);
  int iters_for(bool seg_only) const;
  void enqueue_seg_pass();
  void enqueue_bad_pass();
  void capture_graph();
  void loop_graph();
  void loop_host();
  void loop_mixed();
  bool bad_pass(int ne, int *prev_win);
  void ensure_mix(long long n);
  void size_pass_buffers();
  unsigned long long *ctl_cnt_ptr() const;
  int cavity_lanes(int Kc) const;
  void launch_cavity(bool is_seg, int grid, const PassD &ps);
  int select_by_priority(int *ids, int nb, int cap, int **out_ids);
  void ensure_capacity(long long needP, long long needT, long long needS);
  bool vm_setup();
  void vm_release_all();
  void vm_map_mesh(long long cp, long long ct, long long cs);
  void ensure_cand(int nc);
  void ensure_ws(long long ints);
  void ensure_tmp(int n);
  int cands_per_pass(int K) const;
  void read_counters();
  void free_mesh();
  void tbegin(int id);
  void tend(int id, long long items);
  void collect_times();
  static const char *kname(int id);

  gcdt_stats st{};
  std::vector<InputSet> inputs;
  int np0 = 0, nt0 = 0, ns0 = 0;
  double2 *in_P = nullptr;
  int *in_T = nullptr, *in_N = nullptr, *in_S = nullptr;
  int seg_in_cap = 0;
  int2 *seg_in = nullptr;
  long long capP = 0, capT = 0, capS = 0;
  // growth in place (options.vmm on a device with virtual memory
  // management): every array a pass can outgrow lives in a reserved range
  bool use_vm = false, vm_checked = false;
  VmRange vm_P, vm_pflag, vm_ppar, vm_T, vm_tflag, vm_tbucket, vm_claim, vm_SV, vm_SP, vm_ST,
      vm_senc, vm_tmp, vm_sel_prio, vm_sel_tmp, vm_sel_list;
  int K = 32;
  int np_round = 0;  // live points when the last round opened
  // encroached subsegments listed in tmp_list behind the last pass's apply
  // (-1: not listed; the next segment phase compacts them itself)
  int listed_enc = -1;
  unsigned long long epoch = 0;
  // candidates
  int c_cap = 0;
  double2 *c_pt = nullptr;
  int *c_status = nullptr, *c_n = nullptr, *c_r = nullptr, *c_prio = nullptr,
      *c_win = nullptr, *c_rank = nullptr, *c_act = nullptr, *c_slow = nullptr;
  int *c_ws = nullptr;
  long long ws_ints = 0;
  double *c_pws = nullptr;
  int pws_blocks = 0;
  int *tmp_list = nullptr;
  int tmp_list_cap = 0;
  int *mix_list = nullptr;  // mixed pass: subsegment ids, then bad triangle ids
  long long mix_cap = 0;
  int *sel_prio = nullptr, *sel_tmp = nullptr, *sel_list = nullptr;
  int sel_cap = 0;
  unsigned *d_hist = nullptr;
  float *c_feat = nullptr, *c_score = nullptr;
  int *c_deg = nullptr, *t_cnt = nullptr;
  long long t_cnt_cap = 0;
  int feat_cap = 0;
  int pass_index = 0;
  // compaction scratch
  int *cb_cnt = nullptr, *cb_off = nullptr;
  int cb_cap = 0;
  int *d_total = nullptr;
  int *h_total = nullptr;       // pinned, host-coherent: compaction totals
  int *d_host_total = nullptr;  // its device address
  unsigned long long *d_cnt = nullptr;
  int *h_scalar = nullptr;
  unsigned long long *h_cnt = nullptr;
  unsigned long long *d_host_cnt = nullptr;  // h_cnt as the device sees it (mapped)
  bool cnt_zero = false;  // d_cnt is known to be all zero (flushed by the last pass)
  // segment hash
  unsigned long long *h_key = nullptr;
  int *h_val = nullptr;
  long long h_cap = 0;
  // export
  int *newid = nullptr;
  int newid_cap = 0;
  int nt_out = 0;
  int *o_T = nullptr, *o_N = nullptr;
  uint8_t *o_tstat = nullptr, *o_pstat = nullptr;
  int o_cap = 0, o_pcap = 0;
  // graph-driven loop
  Ctl *d_ctl = nullptr, *h_ctl = nullptr;
  int *g_cnt = nullptr, *g_off = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int g_K = 0;
  bool capturing = false;
  // vertex deletion: on for this run; per-nomination workspace, vertex,
  // star size
  bool del_on = false, in_any_chew = false;
  int *del_ws = nullptr, *del_dv = nullptr, *del_dk = nullptr;
  // export with deleted vertices: new point ids and compacted copies
  int *p_new = nullptr;
  double2 *p_tmp = nullptr;
  uint8_t *pf_tmp = nullptr;
  int *pp_tmp = nullptr;
  long long p_tmp_cap = 0;
  // batched refinement
  int *in_tinst = nullptr;
  InstPar *d_ipar = nullptr;
  int n_inst = 0;
  int *o_inst = nullptr;
  hipEvent_t ev0, ev1;
  // input validation counters
  unsigned long long *d_val = nullptr, *h_val4 = nullptr;
  // download staging: two pinned buffers, filled by the copy engine while
  // host threads drain the other one into the caller's arrays
  static constexpr int NSTAGE = 4;  // pinned staging buffers of the download
  char *h_stage[NSTAGE] = {};
  size_t stage_bytes = 0;
  hipEvent_t ev_stage[NSTAGE] = {};
  CopyPool *copier = nullptr;
  struct KT {
    long long launches = 0;
    double total_ms = 0;
    long long items = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  };
  KT ktimes[K_NKERN];
};

void device_predicates(const double *pts8, int n, int pred_mode, int *o,
                       int *ic, long long *stage_counts);
void device_constructions(const double *tri6, int n, double theta_deg, int offcenter,
                          double *ccw, double *stp, int *ok);
void host_constructions(const double *tri6, int n, double theta_deg, int offcenter, double *ccw,
                        double *stp, int *ok);
void mlp_forward_device(const float *dX, int n, const float *dP, float *dy,
                        hipStream_t stream);
void mlp_forward_device_n(const float *dX, const int *dn, const int *dgo,
                          const float *dP, float *dy, int blocks,
                          hipStream_t stream);
constexpr int FEAT = 16;  // features per candidate (learned priority)
// Validates options; on failure err holds the reason.
bool options_ok(const gcdt_options *o, std::string &err);

// Device invariant checks of a mesh in device memory (gcdt_check.hip);
// viol[10] receives the per-category counts of gcdt_check_mesh, exact_signs
// (if non-null) the number of signs left to the exact stage, extra (if
// non-null) the MPW-exempt count and the smallest angle.  Coverage and
// quality (categories 6 and 8) are checked when dm.S_in is set.
struct CheckExtra {
  long long n_exempt = 0;
  double min_angle_deg = 180.0;
};
void check_device_mesh(const Refiner::DeviceMesh &dm, hipStream_t stream, double theta_deg,
                       long long *viol, long long *exact_signs, CheckExtra *extra);
// The same checks in the library's static scratch array (no allocation, no
// host synchronisation): results to the device array out[13] (viol[10],
// exact signs, exempt count, smallest angle as the bits of a double).
void check_device_mesh_noalloc(const Refiner::DeviceMesh &dm, hipStream_t stream,
                               double theta_deg, long long *out);
size_t check_noalloc_bytes(int np, int nt, int ns);  // scratch a mesh needs
size_t check_noalloc_limit();                        // scratch the array holds

}  // namespace gcdt

namespace gcdt {
// Coordinate-descent search over launch shapes and graph structure; `run`
// refines once with the given options and returns its device time in ms.
// Returns the number of trials written to rec (all trials if rec is NULL).
int autotune(gcdt_options &opt, int reps,
             const std::function<double(const gcdt_options &)> &run,
             gcdt_tune_record *rec, int max_rec);
}  // namespace gcdt
