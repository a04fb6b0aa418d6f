"""MI355X-native constrained-Delaunay quality refinement.

Python bindings (ctypes) over libgcdt.so, the HIP/gfx950 library whose C ABI
is declared in include/gpu_cdt_refine.h.  The refinement itself always runs
on the GPU; there is no CPU fallback: if the library or a GPU is missing the
calls raise.

Main entry points
  refine(points, triangles, neighbors, segments, theta=25, ...) -> RefineResult
      the reference's GPU_Refine_Quality (refine.h:385-386) on numpy arrays
  dropin_refine(...)  the C++ drop-in GPU_Refine_Quality itself (host arrays in
                      and out, triangulateio records), with its wall time
  Session(...)        device-resident input; run() repeatedly (benchmarks)
  generate_pslg(...)  seeded synthetic PSLG (points + non-crossing segments)
  build_cdt(...)      constrained Delaunay triangulation with hull segments
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgcdt.so")

Circumcenter, Offcenter = 0, 1
RUPPERT, CHEW = 1, 0

# The chosen form (what the drop-in runs by default): FP32-filtered exact
# predicates, three claim iterations with shared rings, a bad pass of up to
# 3x the previous winners, the learned priority coarsened by one half-octave,
# and the launch shapes the autotuner picked on the 10^6-point workload.
CHOSEN_FORM = dict(pred_mode=1, claim_iters=3, seg_claim_iters=2, mix_below=4, shared_ring=1,
                   cand_factor=3,
                   priority=2,
                   learned_shift=1, grid_cavity=4096, grid_claim=2048, grid_apply=2048,
                   cavity_lanes=16, cav_lds_cap=32)


class Options(C.Structure):
    _fields_ = [
        ("theta_deg", C.c_double),
        ("offcenter", C.c_int),
        ("mode", C.c_int),
        ("cav_cap0", C.c_int),
        ("cav_cap_max", C.c_int),
        ("cand_budget", C.c_longlong),
        ("max_rounds", C.c_int),
        ("priority", C.c_int),
        ("pred_mode", C.c_int),
        ("profile", C.c_int),
        ("device", C.c_int),
        ("cand_factor", C.c_int),
        ("cand_min", C.c_int),
        ("log_training", C.c_int),
        ("claim_iters", C.c_int),
        ("graph", C.c_int),
        ("graph_seg_passes", C.c_int),
        ("graph_rounds", C.c_int),
        ("grid_cavity", C.c_int),
        ("grid_claim", C.c_int),
        ("grid_apply", C.c_int),
        ("cavity_lanes", C.c_int),
        ("cav_lds_cap", C.c_int),
        ("learned_shift", C.c_int),
        ("mixed", C.c_int),
        ("shared_ring", C.c_int),
        ("mix_below", C.c_int),
        ("deletion", C.c_int),
        ("seg_claim_iters", C.c_int),
        ("vmm", C.c_int),
    ]


class TuneRecord(C.Structure):
    _fields_ = [
        ("graph_rounds", C.c_int),
        ("graph_seg_passes", C.c_int),
        ("grid_cavity", C.c_int),
        ("grid_claim", C.c_int),
        ("grid_apply", C.c_int),
        ("cavity_lanes", C.c_int),
        ("cav_lds_cap", C.c_int),
        ("graph", C.c_int),
        ("ms", C.c_double),
    ]


TUNE_KEYS = ("graph_rounds", "graph_seg_passes", "grid_cavity", "grid_claim", "grid_apply",
             "cavity_lanes", "cav_lds_cap", "graph")


def _tune_result(n, best, rec):
    trials = [dict({k: getattr(rec[i], k) for k in TUNE_KEYS}, ms=rec[i].ms)
              for i in range(min(n, len(rec)))]
    return {k: getattr(best, k) for k in TUNE_KEYS}, trials


class Stats(C.Structure):
    _fields_ = [
        ("rounds", C.c_int),
        ("seg_passes", C.c_int),
        ("bad_passes", C.c_int),
        ("inserted", C.c_longlong),
        ("seg_splits", C.c_longlong),
        ("rej_encroach", C.c_longlong),
        ("rej_blocked", C.c_longlong),
        ("rej_overflow", C.c_longlong),
        ("lost", C.c_longlong),
        ("giveups", C.c_longlong),
        ("final_cav_cap", C.c_int),
        ("status", C.c_int),
        ("device_ms", C.c_double),
        ("pred_fp32", C.c_longlong),
        ("pred_fp64", C.c_longlong),
        ("pred_exact", C.c_longlong),
        ("n_points", C.c_longlong),
        ("n_triangles", C.c_longlong),
        ("n_subsegs", C.c_longlong),
        ("cand_evals", C.c_longlong),
        ("big_cavities", C.c_longlong),
        ("pred_dd", C.c_longlong),
        ("tri_examined_seg", C.c_longlong),
        ("tri_examined_bad", C.c_longlong),
        ("ws_ints_seg", C.c_longlong),
        ("ws_ints_bad", C.c_longlong),
        ("giveup_tris_left", C.c_longlong),
        ("giveup_segs_left", C.c_longlong),
        ("deleted", C.c_longlong),
        ("flips", C.c_longlong),
        ("growths", C.c_longlong),
        ("grow_in_place", C.c_longlong),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_longlong),
                ("total_ms", C.c_double), ("items", C.c_longlong)]


_lib = None


def lib():
    """Load libgcdt.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libgcdt.so is not built; run `python -m gpu_cdt_refine_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.gcdt_default_options.argtypes = [C.POINTER(Options)]
        L.gcdt_refine_host.argtypes = [vp, C.c_int, vp, vp, C.c_int, vp, C.c_int,
                                       C.POINTER(Options), C.POINTER(vp), C.POINTER(Stats)]
        L.gcdt_refine_host.restype = C.c_int
        L.gcdt_mesh_counts.argtypes = [vp, ip, ip, ip]
        L.gcdt_mesh_copy.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.gcdt_mesh_free.argtypes = [vp]
        L.gcdt_session_create.argtypes = [vp, C.c_int, vp, vp, C.c_int, vp, C.c_int,
                                          C.POINTER(Options)]
        L.gcdt_session_create.restype = vp
        L.gcdt_session_run.argtypes = [vp, C.POINTER(Stats)]
        L.gcdt_session_run.restype = C.c_int
        L.gcdt_session_download.argtypes = [vp, C.POINTER(vp)]
        L.gcdt_session_download.restype = C.c_int
        L.gcdt_session_kernel_times.argtypes = [vp, C.POINTER(KernelTime), C.c_int]
        L.gcdt_session_kernel_times.restype = C.c_int
        L.gcdt_session_destroy.argtypes = [vp]
        for name in ("gcdt_save_cdt_text", "gcdt_save_cdt_binary", "gcdt_save_triangle_files"):
            getattr(L, name).argtypes = [C.c_char_p, vp, C.c_int, vp, vp, C.c_int, vp, C.c_int]
            getattr(L, name).restype = C.c_int
        L.gcdt_load_cdt.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.gcdt_load_cdt.restype = C.c_int
        L.gcdt_file_mesh_counts.argtypes = [vp, ip, ip, ip]
        L.gcdt_file_mesh_copy.argtypes = [vp, vp, vp, vp, vp]
        L.gcdt_file_mesh_free.argtypes = [vp]
        L.gcdt_io_last_error.restype = C.c_char_p
        L.gcdt_session_check.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                         C.POINTER(C.c_longlong), C.POINTER(C.c_double)]
        L.gcdt_session_check.restype = C.c_int
        L.gcdt_check_mesh.argtypes = [vp, C.c_int, vp, vp, C.c_int, vp, C.c_int, vp, C.c_int,
                                      vp, C.c_int, C.c_double, C.POINTER(C.c_longlong),
                                      C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                      C.POINTER(C.c_double)]
        L.gcdt_check_mesh.restype = C.c_int
        L.gcdt_check_mesh_device_noalloc.argtypes = [vp, C.c_int, vp, vp, C.c_int, vp, C.c_int,
                                                     vp, C.c_int, vp, C.c_int, C.c_double, vp,
                                                     vp]
        L.gcdt_check_mesh_device_noalloc.restype = C.c_int
        L.gcdt_build_cdt_device.argtypes = [vp, C.c_int, vp, C.c_int, vp, vp]
        L.gcdt_build_cdt_device.restype = C.c_int
        L.gcdt_dropin_set_options.argtypes = [vp]
        L.gcdt_dropin_set_options.restype = C.c_int
        L.gcdt_dropin_set_priority_model.argtypes = [vp]
        L.gcdt_dropin_set_priority_model.restype = C.c_int
        L.gcdt_dropin_last_stats.argtypes = [vp]
        L.gcdt_dropin_last_stats.restype = C.c_int
        L.gcdt_dropin_release.argtypes = []
        L.gcdt_dropin_release.restype = None
        L.gcdt_dropin_free.argtypes = [vp, vp, vp]
        L.gcdt_dropin_free.restype = None
        getattr(L, DROPIN_SYMBOL).restype = None
        L.gcdt_last_error.restype = C.c_char_p
        L.gcdt_device_predicates.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp]
        L.gcdt_device_predicates.restype = C.c_int
        L.gcdt_session_set_priority_model.argtypes = [vp, vp]
        L.gcdt_session_set_priority_model.restype = C.c_int
        L.gcdt_session_logs.argtypes = [vp, vp, vp, C.c_longlong]
        L.gcdt_session_logs.restype = C.c_longlong
        L.gcdt_mlp_nparam.restype = C.c_int
        L.gcdt_mlp_forward.argtypes = [vp, C.c_int, vp, vp]
        L.gcdt_mlp_forward.restype = C.c_int
        L.gcdt_mlp_grad.argtypes = [vp, vp, C.c_int, vp, vp, vp]
        L.gcdt_mlp_grad.restype = C.c_int
        L.gcdt_mlp_train.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, C.c_float,
                                     C.c_ulonglong, vp, vp]
        L.gcdt_mlp_train.restype = C.c_int
        L.gcdt_mfma_probe.argtypes = [vp, vp, vp]
        L.gcdt_mfma_probe.restype = C.c_int
        L.gcdt_mfma_probe32.argtypes = [vp, vp, vp]
        L.gcdt_mfma_probe32.restype = C.c_int
        L.gcdt_mlp_last_error.restype = C.c_char_p
        L.gcdt_batch_create.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                        C.POINTER(Options), C.c_longlong, C.POINTER(vp)]
        L.gcdt_batch_create.restype = C.c_int
        L.gcdt_batch_waves.argtypes = [vp]
        L.gcdt_batch_waves.restype = C.c_int
        L.gcdt_batch_run.argtypes = [vp, C.POINTER(Stats)]
        L.gcdt_batch_run.restype = C.c_int
        L.gcdt_batch_run_ms.argtypes = [vp]
        L.gcdt_batch_run_ms.restype = C.c_double
        L.gcdt_batch_set_priority_model.argtypes = [vp, vp]
        L.gcdt_batch_set_priority_model.restype = C.c_int
        L.gcdt_batch_counts.argtypes = [vp, vp, vp, vp]
        L.gcdt_batch_counts.restype = C.c_int
        L.gcdt_batch_copy.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.gcdt_batch_copy.restype = C.c_int
        L.gcdt_batch_destroy.argtypes = [vp]
        L.gcdt_session_autotune.argtypes = [vp, C.c_int, C.POINTER(Options),
                                            C.POINTER(TuneRecord), C.c_int]
        L.gcdt_session_autotune.restype = C.c_int
        L.gcdt_batch_autotune.argtypes = [vp, C.c_int, C.POINTER(Options),
                                          C.POINTER(TuneRecord), C.c_int]
        L.gcdt_batch_autotune.restype = C.c_int
        L.gcdt_batch_last_error.restype = C.c_char_p
        if hasattr(L, "gcdt_generate_pslg"):
            L.gcdt_generate_pslg.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double,
                                             C.c_double, C.c_double, C.c_ulonglong, vp, vp]
            L.gcdt_generate_pslg.restype = C.c_int
            L.gcdt_build_cdt.argtypes = [vp, C.c_int, vp, C.c_int, C.POINTER(vp)]
            L.gcdt_build_cdt.restype = C.c_int
            L.gcdt_cdt_counts.argtypes = [vp, ip, ip, ip]
            L.gcdt_cdt_copy.argtypes = [vp, vp, vp, vp]
            L.gcdt_cdt_free.argtypes = [vp]
        _lib = L
    return _lib


def _err(rc, what):
    msg = lib().gcdt_last_error().decode()
    raise RuntimeError(f"{what} failed ({rc}): {msg}")


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def make_options(theta=25.0, insert=Offcenter, mode=RUPPERT, **kw):
    o = Options()
    lib().gcdt_default_options(C.byref(o))
    o.theta_deg = float(theta)
    o.offcenter = 1 if insert == Offcenter else 0
    o.mode = int(mode)
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


class RefineResult:
    """Refined mesh in the reference's triangulateio convention."""

    def __init__(self, points, triangles, neighbors, segments, pstatus,
                 tstatus, stats):
        self.points = points          # (P, 2) float64
        self.triangles = triangles    # (T, 3) int32, CCW
        self.neighbors = neighbors    # (T, 3) int32, opposite corner j, -1
        self.segments = segments      # (S, 2) int32 subsegments
        self.pstatus = pstatus        # (P,) uint8 PStatus bits
        self.tstatus = tstatus        # (T,) uint8 TStatus bits
        self.stats = stats            # dict


def _as_inputs(points, triangles, neighbors, segments):
    P = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    T = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)
    N = np.ascontiguousarray(neighbors, dtype=np.int32).reshape(-1, 3)
    S = np.ascontiguousarray(segments, dtype=np.int32).reshape(-1, 2)
    return P, T, N, S


def _fetch_mesh(m):
    L = lib()
    np_, nt, ns = C.c_int(), C.c_int(), C.c_int()
    L.gcdt_mesh_counts(m, C.byref(np_), C.byref(nt), C.byref(ns))
    P = np.empty((np_.value, 2), np.float64)
    T = np.empty((nt.value, 3), np.int32)
    N = np.empty((nt.value, 3), np.int32)
    S = np.empty((ns.value, 2), np.int32)
    ps = np.empty(np_.value, np.uint8)
    ts = np.empty(nt.value, np.uint8)
    L.gcdt_mesh_copy(m, _ptr(P), _ptr(T), _ptr(N), _ptr(S), _ptr(ps), _ptr(ts))
    L.gcdt_mesh_free(m)
    return P, T, N, S, ps, ts


def refine(points, triangles, neighbors, segments, theta=25.0,
           insert=Offcenter, mode=RUPPERT, **options):
    """Refine a CDT on the GPU (GPU_Refine_Quality, refine.cu:6500-7139)."""
    P, T, N, S = _as_inputs(points, triangles, neighbors, segments)
    o = make_options(theta, insert, mode, **options)
    st = Stats()
    m = C.c_void_p()
    rc = lib().gcdt_refine_host(_ptr(P), len(P), _ptr(T), _ptr(N), len(T),
                                _ptr(S), len(S), C.byref(o), C.byref(m), C.byref(st))
    if rc != 0:
        _err(rc, "gcdt_refine_host")
    return RefineResult(*_fetch_mesh(m), st.as_dict())


class TriangulateIO(C.Structure):
    """The reference's triangulateio record (GPU_Refine/triangle.h:253-284),
    as declared in include/gpu_cdt_refine.h."""
    _fields_ = [
        ("pointlist", C.POINTER(C.c_double)), ("pointattributelist", C.POINTER(C.c_double)),
        ("pointmarkerlist", C.POINTER(C.c_int)), ("numberofpoints", C.c_int),
        ("numberofpointattributes", C.c_int), ("trianglelist", C.POINTER(C.c_int)),
        ("triangleattributelist", C.POINTER(C.c_double)),
        ("trianglearealist", C.POINTER(C.c_double)), ("neighborlist", C.POINTER(C.c_int)),
        ("numberoftriangles", C.c_int), ("numberofcorners", C.c_int),
        ("numberoftriangleattributes", C.c_int), ("segmentlist", C.POINTER(C.c_int)),
        ("segmentmarkerlist", C.POINTER(C.c_int)), ("numberofsegments", C.c_int),
        ("holelist", C.POINTER(C.c_double)), ("numberofholes", C.c_int),
        ("regionlist", C.POINTER(C.c_double)), ("numberofregions", C.c_int),
        ("edgelist", C.POINTER(C.c_int)), ("edgemarkerlist", C.POINTER(C.c_int)),
        ("normlist", C.POINTER(C.c_double)), ("numberofedges", C.c_int),
        ("numberofflips", C.c_int)]


# GPU_Refine_Quality keeps the reference's C++ linkage (refine.h:385-386)
DROPIN_SYMBOL = ("_Z18GPU_Refine_QualityP13triangulateioS0_d12InsertPolicy12DeletePolicyiiPP7"
                 "PStatusPP7TStatus")


def dropin_set_options(theta=25.0, insert=Offcenter, mode=RUPPERT, plain=False, **options):
    """Options of the drop-in (gcdt_dropin_set_options).  plain=True selects
    the library defaults (the plain GPU path); otherwise the chosen form with
    the given overrides.  theta / insert / mode come from each call."""
    if plain:
        o = make_options(theta, insert, mode, **options)
    else:
        o = make_options(theta, insert, mode, **dict(CHOSEN_FORM, **options))
    if lib().gcdt_dropin_set_options(C.byref(o)) != 0:
        _err(-1, "gcdt_dropin_set_options")


def dropin_reset():
    """Back to the chosen form with the compiled-in priority network; frees
    the drop-in's cached device buffers."""
    L = lib()
    if L.gcdt_dropin_set_options(None) != 0 or L.gcdt_dropin_set_priority_model(None) != 0:
        _err(-1, "gcdt_dropin_set_options")
    L.gcdt_dropin_release()


def dropin_refine(points, triangles, neighbors, segments, theta=25.0, insert=Offcenter,
                  mode=RUPPERT, debug=False):
    """Calls the C++ drop-in GPU_Refine_Quality on triangulateio records built
    from the arrays and returns (RefineResult, seconds): the wall time of the
    call alone (host arrays in, host arrays out, the reference's timing,
    main.cpp:1860-1868).  The result arrays are copied out and the call's
    new[] arrays freed."""
    P, T, N, S = _as_inputs(points, triangles, neighbors, segments)
    L = lib()
    f = getattr(L, DROPIN_SYMBOL)
    inp, res = TriangulateIO(), TriangulateIO()
    inp.pointlist = P.ctypes.data_as(C.POINTER(C.c_double))
    inp.numberofpoints = len(P)
    inp.trianglelist = T.ctypes.data_as(C.POINTER(C.c_int))
    inp.neighborlist = N.ctypes.data_as(C.POINTER(C.c_int))
    inp.numberoftriangles = len(T)
    inp.numberofcorners = 3
    inp.segmentlist = S.ctypes.data_as(C.POINTER(C.c_int))
    inp.numberofsegments = len(S)
    ps, ts = C.c_void_p(), C.c_void_p()
    import time
    t0 = time.perf_counter()
    f(C.byref(inp), C.byref(res), C.c_double(theta), C.c_int(int(insert)), C.c_int(1),
      C.c_int(int(mode)), C.c_int(-1), C.byref(ps) if debug else None,
      C.byref(ts) if debug else None)
    dt = time.perf_counter() - t0
    np_, nt, ns = res.numberofpoints, res.numberoftriangles, res.numberofsegments
    Po = np.ctypeslib.as_array(res.pointlist, shape=(np_, 2)).copy()
    To = np.ctypeslib.as_array(res.trianglelist, shape=(nt, 3)).copy()
    No = np.ctypeslib.as_array(res.neighborlist, shape=(nt, 3)).copy()
    So = (np.ctypeslib.as_array(res.segmentlist, shape=(ns, 2)).copy() if ns
          else np.zeros((0, 2), np.int32))
    pst = tst = None
    if debug:
        pst = np.ctypeslib.as_array(C.cast(ps, C.POINTER(C.c_uint8)), shape=(np_,)).copy()
        tst = np.ctypeslib.as_array(C.cast(ts, C.POINTER(C.c_uint8)), shape=(nt,)).copy()
    L.gcdt_dropin_free(C.byref(res), ps if debug else None, ts if debug else None)
    st = Stats()
    L.gcdt_dropin_last_stats(C.byref(st))
    d = st.as_dict()
    d["numberofflips"] = res.numberofflips
    return RefineResult(Po, To, No, So, pst, tst, d), dt


class Batch:
    """Many independent CDTs refined together (gcdt_batch_*): `instances` is
    a list of (points, triangles, neighbors, segments) with instance-local
    indices; theta / insert / mode are per instance (scalars broadcast)."""

    def __init__(self, instances, theta=25.0, insert=Offcenter, mode=RUPPERT,
                 wave_triangles=0, **options):
        k = len(instances)
        ins = [_as_inputs(*x) for x in instances]
        self.n_inst = k
        np_ = np.array([len(x[0]) for x in ins], np.int32)
        nt_ = np.array([len(x[1]) for x in ins], np.int32)
        ns_ = np.array([len(x[3]) for x in ins], np.int32)
        P = np.ascontiguousarray(np.concatenate([x[0] for x in ins]))
        T = np.ascontiguousarray(np.concatenate([x[1] for x in ins]))
        N = np.ascontiguousarray(np.concatenate([x[2] for x in ins]))
        S = np.ascontiguousarray(np.concatenate([x[3].reshape(-1, 2) for x in ins]).astype(np.int32))
        th = np.ascontiguousarray(np.broadcast_to(np.asarray(theta, np.float64), (k,)))
        off = np.ascontiguousarray(np.broadcast_to(
            np.asarray([1 if v == Offcenter else 0 for v in np.atleast_1d(insert)], np.int32), (k,)))
        md = np.ascontiguousarray(np.broadcast_to(np.asarray(mode, np.int32), (k,)))
        self._keep = (np_, nt_, ns_, P, T, N, S, th, off, md)
        o = make_options(25.0, Offcenter, RUPPERT, **options)
        h = C.c_void_p()
        rc = lib().gcdt_batch_create(k, _ptr(np_), _ptr(nt_), _ptr(ns_), _ptr(P), _ptr(T),
                                     _ptr(N), _ptr(S), _ptr(th), _ptr(off), _ptr(md),
                                     C.byref(o), int(wave_triangles), C.byref(h))
        if rc != 0:
            raise RuntimeError("gcdt_batch_create: " + lib().gcdt_batch_last_error().decode())
        self.h = h

    def waves(self):
        return lib().gcdt_batch_waves(self.h)

    def run(self):
        st = Stats()
        rc = lib().gcdt_batch_run(self.h, C.byref(st))
        if rc != 0:
            raise RuntimeError("gcdt_batch_run: " + lib().gcdt_batch_last_error().decode())
        d = st.as_dict()
        d["run_ms"] = lib().gcdt_batch_run_ms(self.h)
        return d

    def autotune(self, reps=1):
        """Tune launch shapes / graph structure on this batch (gcdt_batch_autotune);
        returns (best settings, trials)."""
        best = Options()
        rec = (TuneRecord * 64)()
        n = lib().gcdt_batch_autotune(self.h, int(reps), C.byref(best), rec, 64)
        if n < 0:
            raise RuntimeError("gcdt_batch_autotune: " + lib().gcdt_batch_last_error().decode())
        return _tune_result(n, best, rec)

    def set_priority_model(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != mlp_nparam():
            raise ValueError("wrong parameter count")
        rc = lib().gcdt_batch_set_priority_model(self.h, _ptr(p))
        if rc != 0:
            raise RuntimeError("gcdt_batch_set_priority_model: " + lib().gcdt_batch_last_error().decode())

    def results(self):
        """List of RefineResult, one per instance."""
        k = self.n_inst
        cp = np.zeros(k, np.int32); ct = np.zeros(k, np.int32); cs = np.zeros(k, np.int32)
        rc = lib().gcdt_batch_counts(self.h, _ptr(cp), _ptr(ct), _ptr(cs))
        if rc != 0:
            raise RuntimeError("gcdt_batch_counts: " + lib().gcdt_batch_last_error().decode())
        P = np.zeros((int(cp.sum()), 2), np.float64)
        T = np.zeros((int(ct.sum()), 3), np.int32)
        N = np.zeros((int(ct.sum()), 3), np.int32)
        S = np.zeros((int(cs.sum()), 2), np.int32)
        ps = np.zeros(int(cp.sum()), np.uint8)
        ts = np.zeros(int(ct.sum()), np.uint8)
        rc = lib().gcdt_batch_copy(self.h, _ptr(P), _ptr(T), _ptr(N), _ptr(S), _ptr(ps), _ptr(ts))
        if rc != 0:
            raise RuntimeError("gcdt_batch_copy: " + lib().gcdt_batch_last_error().decode())
        out = []
        op = np.concatenate([[0], np.cumsum(cp)]); ot = np.concatenate([[0], np.cumsum(ct)])
        os_ = np.concatenate([[0], np.cumsum(cs)])
        for i in range(k):
            out.append(RefineResult(P[op[i]:op[i + 1]], T[ot[i]:ot[i + 1]], N[ot[i]:ot[i + 1]],
                                    S[os_[i]:os_[i + 1]], ps[op[i]:op[i + 1]], ts[ot[i]:ot[i + 1]], {}))
        return out

    def close(self):
        if self.h:
            lib().gcdt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _io(rc, what):
    if rc != 0:
        raise OSError(f"{what}: {lib().gcdt_io_last_error().decode()}")


def save_cdt(path, points, triangles, neighbors, segments, binary=False):
    """Write a CDT in the reference's text layout (saveCDT, mesh.cpp:402-427;
    coordinates with 17 significant digits) or the binary one."""
    P, T, N, S = _as_inputs(points, triangles, neighbors, segments)
    f = lib().gcdt_save_cdt_binary if binary else lib().gcdt_save_cdt_text
    _io(f(os.fsencode(path), _ptr(P), len(P), _ptr(T), _ptr(N), len(T), _ptr(S), len(S)),
        "save_cdt")


def load_cdt(path):
    """Read a CDT file (text -- also the reference's own saveCDT output -- or
    binary); returns (points, triangles, neighbors, segments)."""
    L = lib()
    m = C.c_void_p()
    _io(L.gcdt_load_cdt(os.fsencode(path), C.byref(m)), "load_cdt")
    try:
        np_, nt, ns = C.c_int(), C.c_int(), C.c_int()
        L.gcdt_file_mesh_counts(m, C.byref(np_), C.byref(nt), C.byref(ns))
        P = np.empty((np_.value, 2))
        T = np.empty((nt.value, 3), np.int32)
        N = np.empty((nt.value, 3), np.int32)
        S = np.empty((ns.value, 2), np.int32)
        L.gcdt_file_mesh_copy(m, _ptr(P), _ptr(T), _ptr(N), _ptr(S))
    finally:
        L.gcdt_file_mesh_free(m)
    return P, T, N, S


def save_triangle_files(basename, points, triangles, neighbors, segments):
    """Write <basename>.node/.ele/.neigh/.poly in Triangle's formats
    (zero-based, as with Triangle's -z)."""
    P, T, N, S = _as_inputs(points, triangles, neighbors, segments)
    _io(lib().gcdt_save_triangle_files(os.fsencode(basename), _ptr(P), len(P), _ptr(T), _ptr(N),
                                        len(T), _ptr(S), len(S)), "save_triangle_files")


CHECK_NAMES = ["range", "symmetry", "orientation", "boundary", "delaunay", "subsegment",
               "coverage", "moved", "quality", "manifold"]
# categories that need the input segments (and theta)
SEGMENT_CHECKS = ["coverage", "quality"]


def _check_dict(v, with_segments):
    return {name: int(v[i]) for i, name in enumerate(CHECK_NAMES)
            if with_segments or name not in SEGMENT_CHECKS}


def check_mesh(points, triangles, neighbors, segments, points_in=None, segments_in=None,
               theta=None):
    """Invariant checks on the GPU with exact predicates (gcdt_check_mesh):
    violation counts per category -- index range, neighbour symmetry,
    orientation, boundary edges that are not subsegments, Delaunay across
    non-subsegment edges, subsegments that are not edges, input points moved
    or missing, edges in three or more triangles, and, when the input
    segments and theta are given, input segments not covered by a chain of
    subsegments and triangles below theta outside the Miller-Pav-Walkington
    exemption.  Also "exact_signs" (incircle signs left to the exact stage),
    "exempt" and "min_angle" (none of them violations)."""
    P, T, N, S = _as_inputs(points, triangles, neighbors, segments)
    Pin = np.zeros((0, 2)) if points_in is None else np.ascontiguousarray(points_in, np.float64).reshape(-1, 2)
    with_seg = segments_in is not None
    if with_seg and theta is None:
        raise ValueError("the quality check needs theta")
    Sin = (np.ascontiguousarray(segments_in, np.int32).reshape(-1, 2) if with_seg
           else np.zeros((0, 2), np.int32))
    v = (C.c_longlong * 10)()
    ex = C.c_longlong(0)
    nex = C.c_longlong(0)
    mina = C.c_double(0.0)
    rc = lib().gcdt_check_mesh(_ptr(P), len(P), _ptr(T), _ptr(N), len(T), _ptr(S), len(S),
                               _ptr(Pin), len(Pin), _ptr(Sin) if with_seg else None, len(Sin),
                               float(theta or 0.0), v, C.byref(ex), C.byref(nex), C.byref(mina))
    if rc != 0:
        _err(rc, "gcdt_check_mesh")
    d = _check_dict(v, with_seg)
    d["exact_signs"] = ex.value  # signs the filters left to the exact stage
    if with_seg:
        d["exempt"] = nex.value
        d["min_angle"] = mina.value
    return d


def check_mesh_device_noalloc(d_points, np_, d_tris, d_nbrs, nt, d_segs, ns, d_points_in,
                              np_in, d_segs_in, ns_in, theta, d_out, stream=None):
    """gcdt_check_mesh_device_noalloc: the checks of check_mesh on a mesh in
    device memory (integer device addresses), enqueued on `stream` (a
    hipStream_t handle or None) without allocating; results go to the device
    array d_out (13 int64, see include/gpu_cdt_refine.h).  Raises on
    rejected input (e.g. a mesh larger than the static scratch holds)."""
    rc = lib().gcdt_check_mesh_device_noalloc(d_points, int(np_), d_tris, d_nbrs, int(nt), d_segs,
                                              int(ns), d_points_in, int(np_in), d_segs_in,
                                              int(ns_in), float(theta), d_out, stream)
    if rc != 0:
        _err(rc, "gcdt_check_mesh_device_noalloc")


NON_VIOLATIONS = ("exact_signs", "exempt", "min_angle")


def violations(d):
    """Total of the violation categories of a check_mesh / Session.check
    result (exact_signs, exempt and min_angle are not violations)."""
    return sum(v for k, v in d.items() if k not in NON_VIOLATIONS)


class Session:
    """Device-resident CDT: upload once, refine repeatedly."""

    def __init__(self, points, triangles, neighbors, segments, theta=25.0,
                 insert=Offcenter, mode=RUPPERT, **options):
        P, T, N, S = _as_inputs(points, triangles, neighbors, segments)
        self._keep = (P, T, N, S)
        o = make_options(theta, insert, mode, **options)
        self.h = lib().gcdt_session_create(_ptr(P), len(P), _ptr(T), _ptr(N), len(T),
                                           _ptr(S), len(S), C.byref(o))
        if not self.h:
            _err(-1, "gcdt_session_create")

    def run(self):
        st = Stats()
        rc = lib().gcdt_session_run(self.h, C.byref(st))
        if rc != 0:
            _err(rc, "gcdt_session_run")
        return st.as_dict()

    def autotune(self, reps=1):
        """Tune launch shapes / graph structure on this input
        (gcdt_session_autotune); returns (best settings, trials)."""
        best = Options()
        rec = (TuneRecord * 64)()
        n = lib().gcdt_session_autotune(self.h, int(reps), C.byref(best), rec, 64)
        if n < 0:
            _err(n, "gcdt_session_autotune")
        return _tune_result(n, best, rec)

    def download(self):
        m = C.c_void_p()
        rc = lib().gcdt_session_download(self.h, C.byref(m))
        if rc != 0:
            _err(rc, "gcdt_session_download")
        P, T, N, S, ps, ts = _fetch_mesh(m)
        return RefineResult(P, T, N, S, ps, ts, {})

    def set_priority_model(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == lib().gcdt_mlp_nparam()
        if lib().gcdt_session_set_priority_model(self.h, _ptr(p)) != 0:
            _err(-2, "gcdt_session_set_priority_model")

    def logs(self):
        """(X[n,16], t[n]) recorded during the last run (log_training=1)."""
        L = lib()
        n = L.gcdt_session_logs(self.h, None, None, 0)
        X = np.empty((n, 16), np.float32)
        t = np.empty(n, np.float32)
        if n:
            L.gcdt_session_logs(self.h, _ptr(X), _ptr(t), n)
        return X, t

    def check(self):
        """Invariant counts of the last refined mesh, checked in HBM
        (gcdt_session_check) against the session's input segments and theta;
        see check_mesh."""
        v = (C.c_longlong * 10)()
        ex = C.c_longlong(0)
        nex = C.c_longlong(0)
        mina = C.c_double(0.0)
        rc = lib().gcdt_session_check(self.h, v, C.byref(ex), C.byref(nex), C.byref(mina))
        if rc != 0:
            _err(rc, "gcdt_session_check")
        d = _check_dict(v, True)
        d["exact_signs"] = ex.value
        d["exempt"] = nex.value
        d["min_angle"] = mina.value
        return d

    def kernel_times(self):
        rec = (KernelTime * 32)()
        n = lib().gcdt_session_kernel_times(self.h, rec, 32)
        return {rec[i].name.decode(): {"launches": rec[i].launches,
                                       "total_ms": rec[i].total_ms,
                                       "items": rec[i].items} for i in range(n)}

    def close(self):
        if self.h:
            lib().gcdt_session_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_predicates(queries, pred_mode=0):
    """Evaluate orient2d(a,b,c) and incircle(a,b,c,d) signs on the GPU for
    an (n, 8) array of (ax, ay, bx, by, cx, cy, dx, dy)."""
    q = np.ascontiguousarray(queries, dtype=np.float64).reshape(-1, 8)
    o = np.empty(len(q), np.int32)
    ic = np.empty(len(q), np.int32)
    sc = np.zeros(4, np.int64)
    rc = lib().gcdt_device_predicates(_ptr(q), len(q), pred_mode, _ptr(o), _ptr(ic), _ptr(sc))
    if rc != 0:
        _err(rc, "gcdt_device_predicates")
    return o, ic, sc


def host_predicates(queries, pred_mode=0):
    """The kernels' predicate code evaluated on the host (no device)."""
    q = np.ascontiguousarray(queries, dtype=np.float64).reshape(-1, 8)
    o = np.empty(len(q), np.int32)
    ic = np.empty(len(q), np.int32)
    L = lib()
    L.gcdt_host_predicates.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    L.gcdt_host_predicates(_ptr(q), len(q), pred_mode, _ptr(o), _ptr(ic))
    return o, ic


def _constructions(fn, tris, theta, offcenter):
    q = np.ascontiguousarray(tris, dtype=np.float64).reshape(-1, 6)
    ccw = np.empty(len(q), np.float64)
    stp = np.empty((len(q), 2), np.float64)
    ok = np.empty(len(q), np.int32)
    f = getattr(lib(), fn)
    f.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    rc = f(_ptr(q), len(q), float(theta), int(offcenter), _ptr(ccw), _ptr(stp), _ptr(ok))
    if rc != 0:
        _err(rc, fn)
    return ccw, stp, ok


def host_constructions(tris, theta=25.0, offcenter=1):
    """The kernels' constructions run on the host for an (n, 6) array of
    triangles (p0, p1, p2): the adaptive orient2d value ccw(p0, p1, p2) and the
    Steiner point of the triangle with apex p0, org p1, dest p2 (ok = 0 when
    degenerate)."""
    return _constructions("gcdt_host_constructions", tris, theta, offcenter)


def device_constructions(tris, theta=25.0, offcenter=1):
    """host_constructions evaluated on the GPU."""
    return _constructions("gcdt_device_constructions", tris, theta, offcenter)


def generate_pslg(n, nseg, L=1000.0, maxlen=0.0, dist=0, min_angle=5.0, seed=1):
    """Seeded synthetic PSLG: n distinct points in [0,L]^2 (dist 0 uniform,
    1 Gaussian clusters, 2 precision-stress mix, 3 disk, 4 circle ring,
    5 grid lattice, 6 ellipse curve -- 3 to 6 are the reference generator's
    families, mesh.cpp:131-193) and up to nseg non-crossing segments (host
    code in libgcdt)."""
    P = np.empty((n, 2), np.float64)
    S = np.empty((nseg, 2), np.int32)
    k = lib().gcdt_generate_pslg(int(dist), int(n), int(nseg), float(L), float(maxlen),
                                 float(min_angle), int(seed), _ptr(P), _ptr(S))
    if k < 0:
        _err(k, "gcdt_generate_pslg")
    return P, S[:k].copy()


def build_cdt(points, segments, device=False, stats=None):
    """CDT of points + segments with hull segments (Triangle's "pzQnc"
    contract).  Returns (triangles, neighbors, segments).  device=True builds
    it on the GPU (gcdt_build_cdt_device); stats (a dict) then receives the
    rounds, segment passes, pieces and device milliseconds."""
    P = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    S = np.ascontiguousarray(segments, dtype=np.int32).reshape(-1, 2)
    c = C.c_void_p()
    if device:
        st = (C.c_double * 4)()
        rc = lib().gcdt_build_cdt_device(_ptr(P), len(P), _ptr(S), len(S), C.byref(c), st)
        if rc != 0:
            _err(rc, "gcdt_build_cdt_device")
        if stats is not None:
            stats.update(rounds=int(st[0]), segment_passes=int(st[1]), segment_pieces=int(st[2]),
                         device_ms=float(st[3]))
    else:
        rc = lib().gcdt_build_cdt(_ptr(P), len(P), _ptr(S), len(S), C.byref(c))
        if rc != 0:
            _err(rc, "gcdt_build_cdt")
    np_, nt, ns = C.c_int(), C.c_int(), C.c_int()
    lib().gcdt_cdt_counts(c, C.byref(np_), C.byref(nt), C.byref(ns))
    T = np.empty((nt.value, 3), np.int32)
    N = np.empty((nt.value, 3), np.int32)
    SS = np.empty((ns.value, 2), np.int32)
    lib().gcdt_cdt_copy(c, _ptr(T), _ptr(N), _ptr(SS))
    lib().gcdt_cdt_free(c)
    return T, N, SS


# --------------------------------------------------------------- learned MLP
MLP_IN, MLP_H = 16, 64


def mlp_nparam():
    return lib().gcdt_mlp_nparam()


def mlp_init(seed=0):
    """He-initialised parameters (FP32 master copy)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(mlp_nparam(), np.float32)
    o = 0
    p[o:o + MLP_IN * MLP_H] = rng.normal(0, np.sqrt(2 / MLP_IN), MLP_IN * MLP_H); o += MLP_IN * MLP_H
    o += MLP_H
    p[o:o + MLP_H * MLP_H] = rng.normal(0, np.sqrt(2 / MLP_H), MLP_H * MLP_H); o += MLP_H * MLP_H
    o += MLP_H
    p[o:o + MLP_H] = rng.normal(0, np.sqrt(1 / MLP_H), MLP_H)
    return p


def _mlp_err(rc, what):
    raise RuntimeError(f"{what} failed ({rc}): {lib().gcdt_mlp_last_error().decode()}")


def mlp_forward(X, params):
    X = np.ascontiguousarray(X, np.float32).reshape(-1, MLP_IN)
    p = np.ascontiguousarray(params, np.float32)
    y = np.empty(len(X), np.float32)
    rc = lib().gcdt_mlp_forward(_ptr(X), len(X), _ptr(p), _ptr(y))
    if rc:
        _mlp_err(rc, "gcdt_mlp_forward")
    return y


def mlp_grad(X, t, params):
    X = np.ascontiguousarray(X, np.float32).reshape(-1, MLP_IN)
    t = np.ascontiguousarray(t, np.float32)
    p = np.ascontiguousarray(params, np.float32)
    g = np.empty(mlp_nparam(), np.float32)
    loss = np.zeros(1, np.float32)
    rc = lib().gcdt_mlp_grad(_ptr(X), _ptr(t), len(X), _ptr(p), _ptr(g), _ptr(loss))
    if rc:
        _mlp_err(rc, "gcdt_mlp_grad")
    return g, float(loss[0])


def mlp_train(X, t, params, epochs=10, batch=4096, lr=1e-3, seed=0):
    X = np.ascontiguousarray(X, np.float32).reshape(-1, MLP_IN)
    t = np.ascontiguousarray(t, np.float32)
    p = np.ascontiguousarray(params, np.float32).copy()
    losses = np.zeros(epochs, np.float32)
    rc = lib().gcdt_mlp_train(_ptr(X), _ptr(t), len(X), epochs, batch, lr, seed, _ptr(p),
                              _ptr(losses))
    if rc:
        _mlp_err(rc, "gcdt_mlp_train")
    return p, losses


def mfma_probe(A, B):
    A = np.ascontiguousarray(A, np.float32).reshape(16, 16)
    B = np.ascontiguousarray(B, np.float32).reshape(16, 16)
    D = np.empty((16, 16), np.float32)
    rc = lib().gcdt_mfma_probe(_ptr(A), _ptr(B), _ptr(D))
    if rc:
        _mlp_err(rc, "gcdt_mfma_probe")
    return D


def mfma_probe32(A, B):
    """D = A (16x32) @ B (32x16) through the 16x16x32 BF16 MFMA."""
    A = np.ascontiguousarray(A, np.float32).reshape(16, 32)
    B = np.ascontiguousarray(B, np.float32).reshape(32, 16)
    D = np.empty((16, 16), np.float32)
    rc = lib().gcdt_mfma_probe32(_ptr(A), _ptr(B), _ptr(D))
    if rc:
        _mlp_err(rc, "gcdt_mfma_probe32")
    return D
