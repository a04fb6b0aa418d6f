/*
 * gpu_cdt_refine.h -- MI355X (gfx950) constrained-Delaunay quality refinement.
 *
 * Two layers:
 *
 * 1. The drop-in entry point, with the reference's exact C++ signature and
 *    record layouts (reference GPU_Refine/refine.h:385-386, README.md:11-51):
 *
 *      void GPU_Refine_Quality(triangulateio *input, triangulateio *result,
 *          double theta, InsertPolicy insertpolicy, DeletePolicy deletepolicy,
 *          int mode, int debug_iter, PStatus **ps_debug, TStatus **ts_debug);
 *
 *    `triangulateio` is byte-compatible with the reference's modified
 *    Triangle record, including its trailing `numberofflips`
 *    (GPU_Refine/triangle.h:253-284).  Define GCDT_HAVE_TRIANGULATEIO before
 *    including this header when triangle.h is already included.
 *
 * 2. A C ABI (extern "C", plain pointers and sizes) that the drop-in is built
 *    on, used by the Python bindings and by callers that keep their mesh in
 *    device memory.
 */
#ifndef GPU_CDT_REFINE_H
#define GPU_CDT_REFINE_H

#include <stddef.h>
#include <stdint.h>

#ifndef GCDT_HAVE_TRIANGULATEIO
#define GCDT_HAVE_TRIANGULATEIO
/* Layout of the reference's triangulateio (GPU_Refine/triangle.h:253-284). */
struct triangulateio {
  double *pointlist;
  double *pointattributelist;
  int *pointmarkerlist;
  int numberofpoints;
  int numberofpointattributes;
  int *trianglelist;
  double *triangleattributelist;
  double *trianglearealist;
  int *neighborlist;
  int numberoftriangles;
  int numberofcorners;
  int numberoftriangleattributes;
  int *segmentlist;
  int *segmentmarkerlist;
  int numberofsegments;
  double *holelist;
  int numberofholes;
  double *regionlist;
  int numberofregions;
  int *edgelist;
  int *edgemarkerlist;
  double *normlist;
  int numberofedges;
  int numberofflips;
};
#endif

/* Enumerations of the reference (GPU_Refine/refine.h:32-33), same values. */
typedef enum { Circumcenter = 0, Offcenter = 1 } InsertPolicy;
typedef enum {
  Strict = 0,
  Connected = 1,
  Orthogonal = 2,
  noDeletion = 3,
  DeleteFromDT = 4,
  noPriority = 5
} DeletePolicy;

/* One-byte debug status records (GPU_Refine/refine.h:36-205).
 * PStatus bits: 0 deleted, 1 Steiner, 2 segment split, 3 new, 4 redundant.
 * TStatus bits: 0 non-null, 1 bad, 2 check, 3 flip, 4-7 flip orientation. */
typedef struct PStatus { unsigned char _status; } PStatus;
typedef struct TStatus { unsigned char _status; } TStatus;

#ifdef __cplusplus
/* Drop-in replacement for the reference's GPU_Refine_Quality
 * (GPU_Refine/refine.h:385-386, refine.cu:6500-7139).  C++ linkage, like the
 * reference.  `input` is a CDT in Triangle's convention (CCW triangles,
 * neighbour j opposite corner j, every convex-hull edge in segmentlist).
 * `result` receives pointlist/trianglelist/neighborlist/segmentlist
 * allocated with new[] (free with delete[]) and the four counts; the first
 * numberofpoints input points are the input points, unmoved.
 * `debug_iter` is accepted and ignored, as in the reference.  `deletepolicy`
 * (inert in the reference) selects vertex deletion: noDeletion turns it
 * off, DeleteFromDT deletes in every mode (the reference's rule,
 * refine.cu:604-677), any other value keeps the configured rule (default:
 * in Chew mode, Triangle's).  numberofflips receives gcdt_stats.flips.  On a device error the function prints it and exits, as the
 * reference's checkCUDAError does (refine.cu:20-29).
 * It runs the chosen form (FP32-filtered exact predicates, three claim
 * iterations with shared rings, the learned priority compiled into the
 * library, mix_below 4) unless gcdt_dropin_set_options chose other options; the device
 * buffers of one cached refiner serve every call (gcdt_dropin_release frees
 * them).  Calls are serialised by a process-wide lock. */
void GPU_Refine_Quality(struct triangulateio *input,
                        struct triangulateio *result, double theta,
                        InsertPolicy insertpolicy, DeletePolicy deletepolicy,
                        int mode, int debug_iter, PStatus **ps_debug,
                        TStatus **ts_debug);
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Options of the refinement.  gcdt_default_options() fills the defaults. */
typedef struct gcdt_options {
  double theta_deg;      /* minimum angle in degrees (refine.cu:6500 `theta`) */
  int offcenter;         /* 1 = off-center (Offcenter), 0 = circumcenter */
  int mode;              /* 1 = Ruppert (diametral circle), else Chew (lens) */
  int cav_cap0;          /* initial cavity capacity K (triangles), default 32 */
  int cav_cap_max;       /* largest cavity capacity, default and limit 2048
                            (one wave's LDS); a larger cavity gives up */
  long long cand_budget; /* cavity workspace in ints; candidates per pass =
                            cand_budget / (6K + 16); default 2^30.  When more
                            triangles are bad, the smallest (priority, index)
                            ones are taken */
  int max_rounds;        /* refinement round cap, default 100000 */
  int priority;          /* 1 = shortest-edge bucket then index (default);
                            0 = lowest index only (refine.cu:4159);
                            2 = learned priority (gcdt_session_set_priority_model):
                            coarsened bucket, then the MLP's predicted
                            conflict degree, higher first (learned_shift) */
  int pred_mode;         /* 0 = FP64 filter + exact; 1 = FP32 semi-static,
                            then FP64, then exact */
  int profile;           /* 1 = time every kernel launch with HIP events */
  int device;            /* HIP device ordinal */
  int cand_factor;       /* a bad pass evaluates at most max(cand_min,
                            cand_factor * winners of the previous bad pass)
                            candidates, the smallest (priority, index) ones;
                            0 = every bad triangle (default 4) */
  int cand_min;          /* default 16384 */
  int log_training;      /* 1 = record per-candidate features and conflict
                            degrees of every bad pass (gcdt_session_logs) */
  int claim_iters;       /* conflict-resolution iterations per pass (default 1,
                            the reference's single atomicMin claim); more
                            iterations let candidates that touch no winner
                            compete again, enlarging each independent set */
  int graph;             /* 1 = device-driven round loop: every kernel reads
                            its work count from a device control block, a
                            block of rounds is one hipGraph launch and the
                            host synchronises once per launch (same result
                            as 0); 0 = host-driven passes (default) */
  int graph_seg_passes;  /* segment passes captured per round (default 2) */
  int graph_rounds;      /* rounds per graph launch (default 4) */
  int grid_cavity;       /* wave64 blocks of the cavity kernels (default 8192,
                            limit 16384) */
  int grid_claim;        /* 256-lane blocks of the claim kernels (default 2048) */
  int grid_apply;        /* 128-lane blocks of the apply kernel (default 4096) */
  int cavity_lanes;      /* lanes per candidate in the cavity kernel: 8, 16,
                            32 or 64 (a wave evaluates 64/cavity_lanes
                            candidates at once; default 16; widened when the
                            groups' LDS would not fit) */
  int cav_lds_cap;       /* cavities of up to this many triangles are built
                            in the filtered kernel's LDS; larger ones (and
                            undecided signs) go to the exact kernel, which
                            holds cav_cap triangles (default 32) */
  int learned_shift;     /* learned priority: claims compare shortest-edge
                            buckets coarsened by 2^learned_shift half-octaves,
                            then the predicted conflict degree (higher first);
                            0 keeps Triangle's half-octave order (default 2) */
  int mixed;             /* 0: a round is segment-split passes until no
                            subsegment is encroached, then one bad-triangle
                            pass (the reference's round, refine.cu:6700-7100);
                            1: a round is ONE pass in which the encroached
                            subsegments and the selected bad triangles compete
                            for their regions, subsegments first in candidate
                            order (they win every tie).  Host-driven loop only
                            (graph = 0).  Default 0 */
  int shared_ring;       /* 0: the winners of a pass own disjoint cavity +
                            ring regions (the reference's rule); 1: a winner's
                            cavity may not meet another winner's cavity or
                            ring, but rings may overlap -- the condition under
                            which simultaneous star insertions equal a
                            sequential order.  Default 0 */
  int mix_below;         /* the segment phase of a round ends once at most
                            this many subsegments are encroached; those join
                            the round's bad pass, first in candidate order.
                            0 is the reference's round.  Host-driven loop
                            only (graph = 0).  Default 0 */
  int deletion;          /* before an encroached subsegment splits, free
                            vertices (inserted off segments) strictly inside
                            its diametral circle that are apexes of its
                            triangles are deleted and their star
                            retriangulated: 0 never; 1 in Chew mode (mode !=
                            1), Triangle's rule (triangle.cpp:13319-13357,
                            default); 2 in every mode, the reference's
                            (refine.cu:604-677) */
  int seg_claim_iters;   /* claim iterations of passes that split only
                            subsegments (0: claim_iters; default 0) */
  int vmm;               /* 1: arrays a pass can outgrow live in reserved
                            address ranges and grow in place by mapping more
                            memory (no copy; the graph loop keeps its captured
                            graph) when the device supports it (default);
                            0: they are reallocated and copied */
} gcdt_options;

typedef struct gcdt_stats {
  int rounds, seg_passes, bad_passes;
  long long inserted, seg_splits, rej_encroach, rej_blocked, rej_overflow,
      lost, giveups;
  int final_cav_cap;
  int status;            /* 0 converged: no bad triangle and no encroached
                            subsegment is left; 1 round cap; 2 stalled (no
                            progress possible); 3 converged except for
                            given-up elements (giveup_tris_left triangles
                            below theta, giveup_segs_left encroached
                            subsegments whose split point or cavity could not
                            be formed in FP64) */
  double device_ms;      /* HIP-event time of the refinement on the device */
  long long pred_fp32, pred_fp64, pred_exact; /* predicate stage hits */
  long long n_points, n_triangles, n_subsegs;
  long long cand_evals;  /* bad-pass candidates evaluated */
  long long big_cavities; /* cavity evaluations that left the LDS fast path */
  long long pred_dd;     /* signs decided by the double-double stage */
  long long tri_examined_seg, tri_examined_bad; /* triangle records the cavity
                            kernels examined (walk steps, seeds, conflict
                            tests), segment / bad-triangle candidates */
  long long ws_ints_seg, ws_ints_bad; /* ints written to candidate workspaces
                            (cavity + ring lists, ordered star boundary) */
  long long giveup_tris_left, giveup_segs_left; /* given-up elements still in
                            the refined mesh (see status 3) */
  long long deleted;     /* free vertices deleted (options.deletion) */
  long long flips;       /* edge flips Lawson's algorithm would make for the
                            same insertions and deletions: per inserted
                            vertex its cavity's triangles less one (less two
                            on a subsegment with triangles on both sides), per
                            deleted vertex its degree less three (the drop-in
                            returns it as numberofflips, triangle.h:283) */
  long long growths;     /* times the mesh arrays grew during the run */
  long long grow_in_place; /* 1: they grew in place (options.vmm) */
} gcdt_stats;

typedef struct gcdt_mesh gcdt_mesh;       /* host-side result */
typedef struct gcdt_session gcdt_session; /* device-resident input + state */

/* Per-kernel timing record (filled when options.profile = 1). */
typedef struct gcdt_kernel_time {
  char name[32];
  long long launches;
  double total_ms;
  long long items; /* work items processed by those launches */
} gcdt_kernel_time;

void gcdt_default_options(gcdt_options *opt);

/* Options and model of the drop-in GPU_Refine_Quality (no reference
 * counterpart: the reference's entry point has one fixed method).
 * gcdt_dropin_set_options: every field but theta_deg / offcenter / mode
 * (those come from each call); NULL restores the chosen form.  Returns -1
 * (gcdt_last_error) on invalid options.  gcdt_dropin_set_priority_model:
 * the priority network's gcdt_mlp_nparam() FP32 weights, NULL = the
 * trained weights compiled into the library.  gcdt_dropin_last_stats: the
 * counters of the last call.  gcdt_dropin_release: frees the cached device
 * buffers.  gcdt_dropin_free: delete[]s the arrays a call allocated, for
 * callers (FFI) that cannot. */
int gcdt_dropin_set_options(const gcdt_options *opt);
int gcdt_dropin_set_priority_model(const float *params);
int gcdt_dropin_last_stats(gcdt_stats *stats);
void gcdt_dropin_release(void);
void gcdt_dropin_free(struct triangulateio *result, PStatus *ps_debug, TStatus *ts_debug);

/* Refine a CDT given in host memory; returns 0 on success.  On success
 * *out holds the refined mesh (free with gcdt_mesh_free). */
int gcdt_refine_host(const double *points, int np, const int *tris,
                     const int *nbrs, int nt, const int *segs, int ns,
                     const gcdt_options *opt, gcdt_mesh **out,
                     gcdt_stats *stats);

void gcdt_mesh_counts(const gcdt_mesh *m, int *np, int *nt, int *ns);
/* Any output pointer may be NULL.  pstatus/tstatus use the PStatus /
 * TStatus bit layouts. */
void gcdt_mesh_copy(const gcdt_mesh *m, double *points, int *tris, int *nbrs,
                    int *segs, unsigned char *pstatus, unsigned char *tstatus);
void gcdt_mesh_free(gcdt_mesh *m);

/* Device-resident sessions: upload once, refine many times (each run
 * starts again from the uploaded CDT), download the last result.  The
 * timed work of a run starts with the input already in HBM. */
gcdt_session *gcdt_session_create(const double *points, int np,
                                  const int *tris, const int *nbrs, int nt,
                                  const int *segs, int ns,
                                  const gcdt_options *opt);
int gcdt_session_run(gcdt_session *s, gcdt_stats *stats);
int gcdt_session_download(gcdt_session *s, gcdt_mesh **out);
int gcdt_session_kernel_times(gcdt_session *s, gcdt_kernel_time *rec,
                              int max_rec);
void gcdt_session_destroy(gcdt_session *s);

/* Invariant checks on the GPU with exact predicates (no reference
 * counterpart: they replace the reference's host checks checkNeighbors
 * main.cpp:79-113, checkIncircle main.cpp:153-242, checkQuality
 * main.cpp:280-328 and the range checks of checkResult main.cpp:336-362).
 * viol[10] receives counts per category: 0 index out of range, 1 neighbour
 * links not symmetric, 2 triangle not counterclockwise, 3 boundary edge that
 * is not a subsegment, 4 exact incircle > 0 across a non-subsegment edge,
 * 5 subsegment that is not a mesh edge, 6 input segment not covered end to
 * end by a chain of output subsegments on it (plus the number of output
 * subsegments no chain used), 7 input point moved or missing, 8 smallest
 * angle below theta_deg outside the Miller-Pav-Walkington exemption
 * (triangle.cpp:7331-7378), 9 edge shared by three or more triangles.  A
 * range violation skips the other checks.  Categories 6 and 8 are checked
 * only when segs_in (the input segments, 2 ints each) is given; gcdt_session_check
 * takes them and theta from the session.  gcdt_check_mesh takes host arrays
 * in the triangulateio convention (neighbours as triangle indices, -1 on the
 * boundary); gcdt_session_check checks the session's last result in HBM.
 * exact_signs receives how many incircle signs the filters left to the
 * exact expansion stage, n_exempt the MPW-exempt triangles and min_angle_deg
 * the smallest angle of the triangles that are not exempt; any of the three
 * may be NULL. */
int gcdt_check_mesh(const double *points, int np, const int *tris, const int *nbrs, int nt,
                    const int *segs, int ns, const double *points_in, int np_in,
                    const int *segs_in, int ns_in, double theta_deg, long long *viol,
                    long long *exact_signs, long long *n_exempt, double *min_angle_deg);
int gcdt_session_check(gcdt_session *s, long long *viol, long long *exact_signs,
                       long long *n_exempt, double *min_angle_deg);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The checks of gcdt_check_mesh on a mesh already in
 * device memory (device pointers, same layouts and meaning), enqueued on
 * `stream` (a hipStream_t; NULL = the null stream) with no allocation, no
 * memset node and no host synchronisation.  Scratch comes from a static
 * 64 MiB device array per GPU, held until the process ends; a mesh whose
 * scratch would not fit (a subsegment hash table of 8 B x the power of two
 * >= 2 ns + 16, an edge hash table of 12 B x the power of two >= 6 nt + 16,
 * 12 B per triangle, 12 B per point, 8 B per subsegment and a 4 MiB
 * exact-stage arena: meshes of up to about 350,000 triangles fit) is
 * rejected with -2 and a message in gcdt_last_error before anything is
 * enqueued.  Results go to
 * the device array d_out[13]: viol[0..9] as in gcdt_check_mesh, [10] signs
 * left to the exact stage, [11] MPW-exempt triangles, [12] the smallest
 * non-exempt angle in degrees as the bits of a double (180 if none).  Calls
 * share the array: calls that overlap on one GPU race; each call clears what
 * it uses. */
int gcdt_check_mesh_device_noalloc(const double *d_points, int np, const int *d_tris,
                                   const int *d_nbrs, int nt, const int *d_segs, int ns,
                                   const double *d_points_in, int np_in, const int *d_segs_in,
                                   int ns_in, double theta_deg, long long *d_out,
                                   void *stream);

/* Batched refinement (no counterpart in the reference, which refines one
 * CDT per GPU_Refine_Quality call, refine.cu:6500-7139): n_inst independent
 * CDTs with their own theta (degrees), insertion policy (1 off-center) and
 * mode (1 Ruppert), refined together.  Inputs are concatenated per instance
 * with instance-local indices: np/nt/ns[n_inst] counts, points[2*sum np],
 * tris and nbrs[3*sum nt], segs[2*sum ns] (each instance in the input
 * contract of GPU_Refine_Quality).  Instances are packed into waves of at
 * most wave_triangles input triangles (<= 0: 2^25); a wave is refined as the
 * union of its instances, with every instance keeping its own parameters.
 * All inputs stay resident on the device; gcdt_batch_run refines every wave
 * (graph-driven when options.graph = 1).  Results are per instance, in
 * instance order, with instance-local indices: input points first, then
 * Steiner points. */
typedef struct gcdt_batch gcdt_batch;
int gcdt_batch_create(int n_inst, const int *np, const int *nt, const int *ns,
                      const double *points, const int *tris, const int *nbrs,
                      const int *segs, const double *theta_deg,
                      const int *offcenter, const int *mode,
                      const gcdt_options *opt, long long wave_triangles,
                      gcdt_batch **out);
int gcdt_batch_waves(const gcdt_batch *b);
int gcdt_batch_run(gcdt_batch *b, gcdt_stats *stats);
/* host wall time of the last run's refinement (ms), transfers excluded */
double gcdt_batch_run_ms(const gcdt_batch *b);
int gcdt_batch_set_priority_model(gcdt_batch *b, const float *params);
/* Per-instance output counts (arrays of n_inst; any may be NULL). */
int gcdt_batch_counts(gcdt_batch *b, int *np, int *nt, int *ns);
/* Concatenated per-instance outputs (sizes from gcdt_batch_counts). */
int gcdt_batch_copy(gcdt_batch *b, double *points, int *tris, int *nbrs,
                    int *segs, unsigned char *pstatus, unsigned char *tstatus);
void gcdt_batch_destroy(gcdt_batch *b);
const char *gcdt_batch_last_error(void);

/* Autotuner (no counterpart in the reference, whose launch shapes are fixed
 * at 256-thread blocks, refine.h:19).  Coordinate descent over the launch
 * shapes of the persistent kernels (cavity_lanes, cav_lds_cap, grid_cavity,
 * grid_apply, grid_claim), the round driver (graph 0 / 1) and, for the graph
 * loop, the captured graph's structure (graph_rounds, graph_seg_passes), timing `reps` refinements of the resident input per
 * setting (median device time).  These settings never change the refined
 * mesh.  The best setting is kept for later runs and returned in *best;
 * rec (may be NULL) receives up to max_rec trials; returns the number of
 * trials, or a negative error code. */
typedef struct gcdt_tune_record {
  int graph_rounds, graph_seg_passes, grid_cavity, grid_claim, grid_apply,
      cavity_lanes, cav_lds_cap, graph;
  double ms;
} gcdt_tune_record;
int gcdt_session_autotune(gcdt_session *s, int reps, gcdt_options *best,
                          gcdt_tune_record *rec, int max_rec);
int gcdt_batch_autotune(gcdt_batch *b, int reps, gcdt_options *best,
                        gcdt_tune_record *rec, int max_rec);

/* Learned priority (options.priority = 2): load the MLP parameters
 * (gcdt_mlp_nparam() floats: W1[16][64] b1[64] W2[64][64] b2[64] W3[64] b3).
 * The model predicts log2(1 + conflict degree) from 16 candidate features and
 * orders candidates inside a shortest-edge bucket. */
int gcdt_session_set_priority_model(gcdt_session *s, const float *params);
/* After a run with options.log_training = 1: copy up to max_n logged samples
 * (features X[max_n][16], targets t[max_n]); returns the number available. */
long long gcdt_session_logs(gcdt_session *s, float *X, float *t, long long max_n);

/* MLP kernels (BF16 MFMA, FP32 accumulate) on the device. */
int gcdt_mlp_nparam(void);
int gcdt_mlp_forward(const float *X, int n, const float *params, float *y);
int gcdt_mlp_grad(const float *X, const float *t, int n, const float *params,
                  float *grad, float *loss);
int gcdt_mlp_train(const float *X, const float *t, int n, int epochs, int batch,
                   float lr, unsigned long long seed, float *params,
                   float *losses);
int gcdt_mfma_probe(const float *A, const float *B, float *D);
/* D(16x16) = A(16x32) B(32x16) through gfx950's 16x16x32 BF16 MFMA (the
 * inference kernel's operand layout, checked against a host product). */
int gcdt_mfma_probe32(const float *A, const float *B, float *D);
const char *gcdt_mlp_last_error(void);

/* Last error message of this thread ("" if none). */
const char *gcdt_last_error(void);

/* Number of visible HIP devices (0 if none). */
int gcdt_device_count(void);

/* Host-side input preparation (no device use).
 * gcdt_generate_pslg: seeded synthetic PSLG -- `dist` 0 uniform in
 * [0,L]^2, 1 Gaussian clusters (sigma = 0.05 L), 2 the precision-stress mix
 * (jittered lattice, exactly cocircular groups, uniform background), and
 * the reference generator's remaining families (mesh.cpp:131-193): 3 disk
 * (radius 0.45 L), 4 circle (ring of radii 0.45 L..0.49 L), 5 grid (nodes of
 * a lattice with 10 n cells a side), 6 ellipse (on the curve with half-axes
 * 0.25 L and 0.5 L); every family is centred in [0,L]^2 and its points are
 * distinct.  nseg random non-crossing
 * segments between point pairs, each no longer than maxlen (<= 0: no cap),
 * whose angle with any segment sharing an endpoint is at least
 * min_angle_deg.  Returns the number of segments produced.
 * gcdt_build_cdt: constrained Delaunay triangulation of points + segments
 * with the convex hull enclosed by segments (Triangle's "pzQnc" contract),
 * written as triangles (CCW), neighbours (opposite corner j, -1 on the
 * boundary) and the final segment list (input segments, split where they
 * pass through points, plus hull edges).  Returns 0 on success. */
int gcdt_generate_pslg(int dist, int n, int nseg, double L, double maxlen,
                       double min_angle_deg, unsigned long long seed,
                       double *points, int *segs);
typedef struct gcdt_cdt gcdt_cdt;
int gcdt_build_cdt(const double *points, int np, const int *segs, int ns,
                   gcdt_cdt **out);
void gcdt_cdt_counts(const gcdt_cdt *c, int *np, int *nt, int *ns);
void gcdt_cdt_copy(const gcdt_cdt *c, int *tris, int *nbrs, int *segs);
void gcdt_cdt_free(gcdt_cdt *c);
/* The same CDT built on the GPU (no reference counterpart: the reference
 * builds its input with Triangle's triangulate("pzQnc") on the host,
 * mesh.cpp:397, main.cpp:454): parallel Bowyer-Watson insertion with a ghost
 * vertex, then segment recovery by retriangulating the crossed triangles'
 * two sides.  Same output convention and the same triangles as Triangle
 * where no four points are cocircular; the final segment list holds each
 * input segment's pieces in order, then the hull edges that are not
 * segments.  stats (may be NULL) receives insertion rounds, segment passes,
 * segment pieces and device milliseconds.  Returns 0, or -2
 * (gcdt_last_error) for invalid input or a failed build. */
int gcdt_build_cdt_device(const double *points, int np, const int *segs, int ns,
                          gcdt_cdt **out, double *stats);

/* CDT files (reference GPU_Refine/mesh.cpp:402-532, saveCDT / readCDT).
 * gcdt_save_cdt_text writes the reference's text layout (counts of points,
 * triangles, segments; then points, triangles, neighbours, segments one per
 * line) with 17 significant digits per coordinate, so the reference's readCDT
 * reads it and nothing is lost; gcdt_save_cdt_binary writes "GCDTBIN1", three
 * int64 counts and the raw arrays.  gcdt_load_cdt reads either (by the magic
 * bytes), including files from the reference's saveCDT.
 * gcdt_save_triangle_files writes <base>.node/.ele/.neigh/.poly in Triangle's
 * formats with zero-based numbering (Triangle's -z).  Return 0 or -1 with
 * gcdt_io_last_error(). */
typedef struct gcdt_file_mesh gcdt_file_mesh;
int gcdt_save_cdt_text(const char *path, const double *points, int np, const int *tris,
                       const int *nbrs, int nt, const int *segs, int ns);
int gcdt_save_cdt_binary(const char *path, const double *points, int np, const int *tris,
                         const int *nbrs, int nt, const int *segs, int ns);
int gcdt_load_cdt(const char *path, gcdt_file_mesh **out);
void gcdt_file_mesh_counts(const gcdt_file_mesh *m, int *np, int *nt, int *ns);
void gcdt_file_mesh_copy(const gcdt_file_mesh *m, double *points, int *tris, int *nbrs,
                         int *segs);
void gcdt_file_mesh_free(gcdt_file_mesh *m);
int gcdt_save_triangle_files(const char *basename, const double *points, int np,
                             const int *tris, const int *nbrs, int nt, const int *segs, int ns);
const char *gcdt_io_last_error(void);

/* The refinement's predicates evaluated on the host (same source as the
 * kernels): signs of orient2d(a,b,c) and incircle(a,b,c,d) for n queries
 * (ax, ay, bx, by, cx, cy, dx, dy).  No device use. */
int gcdt_host_predicates(const double *pts8, int n, int pred_mode,
                         int *orient_out, int *incircle_out);

/* Exact predicates on the device, for tests: evaluates n queries;
 * stage_counts (4 entries, may be NULL): signs decided by the FP32, FP64,
 * exact-expansion and double-double stages. */
int gcdt_device_predicates(const double *pts8, int n, int pred_mode,
                           int *orient_out, int *incircle_out,
                           long long *stage_counts);

/* The refinement's constructions, for tests (no reference counterpart as
 * entry points): for n triangles tri6 = (p0, p1, p2), ccw_out[i] receives the
 * value of Triangle's adaptive counterclockwise(p0, p1, p2)
 * (triangle.cpp:5155-5291, what findcircumcenter divides by) and
 * steiner_out[2i..2i+1] the circumcenter / off-center of the triangle with
 * apex p0, org p1, dest p2 as the refinement computes it
 * (triangle.cpp:6523-6618, refine.cu:3780-3868); ok_out[i] is 0 where the
 * triangle is degenerate.  The host version runs the kernels' source on the
 * CPU; the device version on HIP device 0. */
int gcdt_host_constructions(const double *tri6, int n, double theta_deg, int offcenter,
                            double *ccw_out, double *steiner_out, int *ok_out);
int gcdt_device_constructions(const double *tri6, int n, double theta_deg, int offcenter,
                              double *ccw_out, double *steiner_out, int *ok_out);

#ifdef __cplusplus
}
#endif
#endif
