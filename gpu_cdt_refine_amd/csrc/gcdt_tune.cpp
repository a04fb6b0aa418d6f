// Autotuner for the launch shapes of the persistent kernels and the
// structure of the captured round graph.  None of these settings changes the
// refined mesh (every kernel grid-strides over device-resident counts, and
// the graph loop's gates skip passes exactly as the host loop would), so the
// search only measures time.
#include <algorithm>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "gcdt_internal.h"
#include "gpu_cdt_refine.h"

namespace gcdt {

namespace {
struct Knob {
  int gcdt_options::*field;
  std::vector<int> values;
  bool graph_only;
};

double median_ms(const gcdt_options &o, int reps,
                 const std::function<double(const gcdt_options &)> &run) {
  std::vector<double> t;
  for (int r = 0; r < reps; r++) t.push_back(run(o));
  std::sort(t.begin(), t.end());
  return t[t.size() / 2];
}
}  // namespace

int autotune(gcdt_options &opt, int reps,
             const std::function<double(const gcdt_options &)> &run,
             gcdt_tune_record *rec, int max_rec) {
  if (reps < 1) reps = 1;
  const std::vector<Knob> knobs = {
      {&gcdt_options::graph, {0, 1}, false},
      {&gcdt_options::cavity_lanes, {8, 16, 32, 64}, false},
      {&gcdt_options::cav_lds_cap, {16, 32, 64}, false},
      {&gcdt_options::grid_cavity, {2048, 4096, 8192, 16384}, false},
      {&gcdt_options::grid_apply, {1024, 2048, 4096, 8192}, false},
      {&gcdt_options::grid_claim, {1024, 2048, 4096}, false},
      {&gcdt_options::graph_rounds, {1, 2, 4, 8}, true},
      {&gcdt_options::graph_seg_passes, {1, 2, 3}, true},
  };
  int n = 0;
  auto log = [&](const gcdt_options &o, double ms) {
    if (rec && n < max_rec) {
      rec[n].graph_rounds = o.graph_rounds;
      rec[n].graph_seg_passes = o.graph_seg_passes;
      rec[n].grid_cavity = o.grid_cavity;
      rec[n].grid_claim = o.grid_claim;
      rec[n].grid_apply = o.grid_apply;
      rec[n].cavity_lanes = o.cavity_lanes;
      rec[n].cav_lds_cap = o.cav_lds_cap;
      rec[n].graph = o.graph;
      rec[n].ms = ms;
    }
    n++;
  };
  gcdt_options best = opt;
  double best_ms = median_ms(best, reps, run);
  log(best, best_ms);
  for (const Knob &k : knobs) {
    if (k.graph_only && !best.graph) continue;
    if (k.field == &gcdt_options::graph && (opt.log_training || opt.mixed || opt.mix_below))
      continue;  // host loop only
    int keep = best.*(k.field);
    for (int v : k.values) {
      if (v == keep) continue;
      gcdt_options o = best;
      o.*(k.field) = v;
      std::string why;
      if (!options_ok(&o, why)) continue;  // e.g. too many claim epochs per graph launch
      double ms = median_ms(o, reps, run);
      log(o, ms);
      if (ms < best_ms) {
        best_ms = ms;
        best = o;
      }
    }
  }
  opt = best;
  return n;
}

}  // namespace gcdt
