"""ctypes bindings for the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, and only as the checker / CPU baseline:
  oracle/liboracle.so       plain-C restatement of the refinement rounds,
                            exact big-integer predicates, invariant checker
  oracle/_ref/libtriref.so  the reference's Triangle 1.6 (GPU_Refine/triangle.cpp)
                            compiled from its own source: the CPU "-q" path
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
TRIREF_SO = os.path.join(HERE, "_ref", "libtriref.so")


def build():
    """Compile liboracle.so and, when the reference tree is present,
    _ref/libtriref.so (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


class TriangulateIO(C.Structure):
    """Layout of the reference's triangulateio (GPU_Refine/triangle.h:253-284)."""
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


class OrcParams(C.Structure):
    _fields_ = [("theta_deg", C.c_double), ("offcenter", C.c_int), ("mode", C.c_int),
                ("cav_cap0", C.c_int), ("cav_cap_max", C.c_int),
                ("cand_budget", C.c_longlong), ("max_rounds", C.c_int),
                ("priority", C.c_int), ("cand_factor", C.c_int), ("cand_min", C.c_int),
                ("claim_iters", C.c_int), ("mixed", C.c_int), ("shared_ring", C.c_int),
                ("mix_below", C.c_int), ("seg_claim_iters", C.c_int), ("deletion", C.c_int)]


class OrcStats(C.Structure):
    _fields_ = [("rounds", C.c_int), ("seg_passes", C.c_int), ("bad_passes", C.c_int),
                ("inserted", C.c_longlong), ("seg_splits", C.c_longlong),
                ("rej_encroach", C.c_longlong), ("rej_blocked", C.c_longlong),
                ("rej_overflow", C.c_longlong), ("lost", C.c_longlong),
                ("giveups", C.c_longlong), ("final_cav_cap", C.c_int), ("status", C.c_int),
                ("cand_evals", C.c_longlong), ("giveup_tris_left", C.c_longlong),
                ("giveup_segs_left", C.c_longlong), ("deleted", C.c_longlong),
                ("flips", C.c_longlong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_orc = None
_tri = None


def orc():
    global _orc
    if _orc is None:
        L = C.CDLL(ORACLE_SO)
        vp = C.c_void_p
        L.orc_refine.restype = vp
        L.orc_refine.argtypes = [C.c_int, vp, C.c_int, vp, vp, C.c_int, vp,
                                 C.POINTER(OrcParams), C.POINTER(OrcStats)]
        L.orc_refine_batch.restype = vp
        L.orc_refine_batch.argtypes = [C.c_int, vp, C.c_int, vp, vp, C.c_int, vp, vp, C.c_int,
                                       vp, vp, vp, C.POINTER(OrcParams), C.POINTER(OrcStats)]
        L.orc_result_inst.argtypes = [vp, vp]
        L.orc_cavity.argtypes = [C.c_int, vp, C.c_int, vp, vp, C.c_int, vp, C.c_int, vp,
                                 C.c_int, vp]
        L.orc_cavity.restype = C.c_int
        L.orc_result_counts.argtypes = [vp, vp, vp, vp]
        L.orc_result_copy.argtypes = [vp, vp, vp, vp, vp, vp]
        L.orc_result_free.argtypes = [vp]
        L.orc_check_mesh.restype = C.c_longlong
        L.orc_check_mesh.argtypes = [C.c_int, vp, C.c_int, vp, vp, C.c_int, vp, C.c_int, vp,
                                     C.c_int, vp, C.c_double, C.c_double, vp, vp, vp]
        for f in ("orc_orient2d", "orc_incircle", "orc_orient2d_exact_sign",
                  "orc_incircle_exact_sign"):
            getattr(L, f).restype = C.c_int
        L.orc_orient2d.argtypes = [vp, vp, vp]
        L.orc_incircle.argtypes = [vp, vp, vp, vp]
        L.orc_steiner_point.argtypes = [vp, vp, vp, C.c_double, C.c_int, vp]
        L.orc_encroaches_xy.argtypes = [vp, vp, vp, C.c_double, C.c_int]
        L.orc_encroaches_xy.restype = C.c_int
        L.orc_split_point_xy.argtypes = [vp, vp, C.c_int, C.c_int, vp]
        L.orc_split_point_xy.restype = C.c_int
        L.orc_bad_angle_xy.argtypes = [vp, vp, vp, C.c_double]
        L.orc_bad_angle_xy.restype = C.c_int
        L.orc_ccw_value.restype = C.c_double
        L.orc_ccw_value.argtypes = [vp, vp, vp]
        _orc = L
    return _orc


def triref():
    global _tri
    if _tri is None:
        if not os.path.exists(TRIREF_SO):
            raise RuntimeError("oracle/_ref/libtriref.so not built (needs the reference tree)")
        L = C.CDLL(TRIREF_SO)
        vp = C.c_void_p
        L.triref_orient2d.restype = C.c_double
        L.triref_orient2d.argtypes = [vp, vp, vp]
        L.triref_incircle.restype = C.c_double
        L.triref_incircle.argtypes = [vp, vp, vp, vp]
        L.triref_findcircumcenter.argtypes = [vp, vp, vp, C.c_double, C.c_int, vp]
        L.triref_triangulate.argtypes = [C.c_char_p, vp, vp]
        L.triref_free.argtypes = [vp]
        _tri = L
    return _tri


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _take(o):
    P = np.ctypeslib.as_array(o.pointlist, (o.numberofpoints * 2,)).copy().reshape(-1, 2)
    T = np.ctypeslib.as_array(o.trianglelist, (o.numberoftriangles * 3,)).copy().reshape(-1, 3)
    N = (np.ctypeslib.as_array(o.neighborlist, (o.numberoftriangles * 3,)).copy().reshape(-1, 3)
         if o.neighborlist else None)
    S = (np.ctypeslib.as_array(o.segmentlist, (o.numberofsegments * 2,)).copy().reshape(-1, 2)
         if o.numberofsegments else np.zeros((0, 2), np.int32))
    L = triref()
    for f in ("pointlist", "trianglelist", "neighborlist", "segmentlist", "pointmarkerlist",
              "segmentmarkerlist", "triangleattributelist", "pointattributelist"):
        ptr = getattr(o, f)
        if ptr:
            L.triref_free(C.cast(ptr, C.c_void_p))
    return P, T, N, S


def triangle_cdt(points, segments):
    """Triangle "pzQnc": CDT with hull segments (mesh.cpp:397)."""
    P = np.ascontiguousarray(points, np.float64).reshape(-1)
    S = np.ascontiguousarray(segments, np.int32).reshape(-1)
    i, o = TriangulateIO(), TriangulateIO()
    i.pointlist = _dp(P)
    i.numberofpoints = len(P) // 2
    i.segmentlist = _ip(S)
    i.numberofsegments = len(S) // 2
    triref().triref_triangulate(b"pzQnc", C.byref(i), C.byref(o))
    return _take(o)


def triangle_refine(points, triangles, segments, theta, mode=1):
    """The reference's CPU_Triangle_Quality (main.cpp:389-407):
    "pzDQnrq<theta>" (Ruppert) or "pzQnrq<theta>" (Chew) on a CDT."""
    P = np.ascontiguousarray(points, np.float64).reshape(-1)
    T = np.ascontiguousarray(triangles, np.int32).reshape(-1)
    S = np.ascontiguousarray(segments, np.int32).reshape(-1)
    i, o = TriangulateIO(), TriangulateIO()
    i.pointlist = _dp(P)
    i.numberofpoints = len(P) // 2
    i.trianglelist = _ip(T)
    i.numberoftriangles = len(T) // 3
    i.numberofcorners = 3
    i.segmentlist = _ip(S)
    i.numberofsegments = len(S) // 2
    sw = ("pzDQnrq" if mode == 1 else "pzQnrq") + ("%g" % theta)
    triref().triref_triangulate(sw.encode(), C.byref(i), C.byref(o))
    return _take(o)


def oracle_refine(points, triangles, neighbors, segments, theta=25.0, offcenter=1,
                  mode=1, cav_cap0=32, cav_cap_max=2048, cand_budget=1 << 30,
                  max_rounds=100000, priority=1, cand_factor=4, cand_min=16384,
                  claim_iters=1, mixed=0, shared_ring=0, mix_below=0, deletion=1,
                  seg_claim_iters=0):
    """Run the CPU restatement; returns (P, T, N, S, pflag, stats).
    deletion: 0 never delete vertices, 1 in Chew mode (Triangle's rule),
    2 in every mode (the reference's)."""
    P = np.ascontiguousarray(points, np.float64).reshape(-1, 2)
    T = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
    N = np.ascontiguousarray(neighbors, np.int32).reshape(-1, 3)
    S = np.ascontiguousarray(segments, np.int32).reshape(-1, 2)
    prm = OrcParams(theta, offcenter, mode, cav_cap0, cav_cap_max, cand_budget,
                    max_rounds, priority, cand_factor, cand_min, claim_iters, mixed, shared_ring,
                    mix_below, seg_claim_iters, deletion)
    st = OrcStats()
    L = orc()
    r = L.orc_refine(len(P), _p(P), len(T), _p(T), _p(N), len(S), _p(S),
                     C.byref(prm), C.byref(st))
    np_, nt, ns = C.c_int(), C.c_int(), C.c_int()
    L.orc_result_counts(r, C.byref(np_), C.byref(nt), C.byref(ns))
    Po = np.empty((np_.value, 2))
    To = np.empty((nt.value, 3), np.int32)
    No = np.empty((nt.value, 3), np.int32)
    So = np.empty((ns.value, 2), np.int32)
    pf = np.empty(np_.value, np.uint8)
    L.orc_result_copy(r, _p(Po), _p(To), _p(No), _p(So), _p(pf))
    L.orc_result_free(r)
    return Po, To, No, So, pf, st.as_dict()


def split_instances(np_in, P, T, N, S, pf, tinst):
    """Split a union mesh into its instances: input points of instance i are
    the union's points [sum(np_in[:i]), sum(np_in[:i+1])); a Steiner point
    belongs to the instance of a triangle that uses it.  Each instance keeps
    its input points first, then its Steiner points in creation order, and
    the relative order of its triangles and subsegments."""
    k = len(np_in)
    pin = np.concatenate([[0], np.cumsum(np_in)]).astype(np.int64)
    np0 = int(pin[-1])
    pinst = np.full(len(P), -1, np.int64)
    for i in range(k):
        pinst[pin[i]:pin[i + 1]] = i
    ti = np.repeat(tinst, 3)
    v = T.reshape(-1)
    st = v >= np0
    pinst[v[st]] = ti[st]
    assert (pinst >= 0).all()
    out = []
    for i in range(k):
        pid = np.nonzero(pinst == i)[0]
        loc = np.full(len(P), -1, np.int64)
        loc[pid] = np.arange(len(pid))
        tid = np.nonzero(tinst == i)[0]
        tl = np.full(len(T), -1, np.int64)
        tl[tid] = np.arange(len(tid))
        Ti = loc[T[tid]].astype(np.int32)
        Ni = np.where(N[tid] < 0, -1, tl[np.maximum(N[tid], 0)]).astype(np.int32)
        sid = np.nonzero(pinst[S[:, 0]] == i)[0]
        Si = loc[S[sid]].astype(np.int32)
        out.append((P[pid], Ti, Ni, Si, pf[pid]))
    return out


def oracle_refine_batch(instances, theta, offcenter, mode, **params):
    """Batched refinement on the CPU: the union of `instances` ((P, T, N, S)
    with local indices) refined with per-instance theta / offcenter / mode;
    returns ([(P, T, N, S, pflag)] per instance, stats)."""
    k = len(instances)
    Ps, Ts, Ns, Ss, I = [], [], [], [], []
    po = to = 0
    for i, (P, T, N, S) in enumerate(instances):
        P = np.asarray(P, np.float64).reshape(-1, 2)
        T = np.asarray(T, np.int32).reshape(-1, 3)
        N = np.asarray(N, np.int32).reshape(-1, 3)
        S = np.asarray(S, np.int32).reshape(-1, 2)
        Ps.append(P); Ts.append(T + po); Ns.append(np.where(N < 0, -1, N + to)); Ss.append(S + po)
        I.append(np.full(len(T), i, np.int32))
        po += len(P); to += len(T)
    P = np.ascontiguousarray(np.concatenate(Ps))
    T = np.ascontiguousarray(np.concatenate(Ts).astype(np.int32))
    N = np.ascontiguousarray(np.concatenate(Ns).astype(np.int32))
    S = np.ascontiguousarray(np.concatenate(Ss).astype(np.int32))
    I = np.ascontiguousarray(np.concatenate(I))
    th = np.ascontiguousarray(np.broadcast_to(np.asarray(theta, np.float64), (k,)))
    off = np.ascontiguousarray(np.broadcast_to(np.asarray(offcenter, np.int32), (k,)))
    md = np.ascontiguousarray(np.broadcast_to(np.asarray(mode, np.int32), (k,)))
    d = dict(cav_cap0=32, cav_cap_max=2048, cand_budget=1 << 30, max_rounds=100000,
             priority=1, cand_factor=4, cand_min=16384, claim_iters=1, mixed=0, shared_ring=0,
             mix_below=0, deletion=1, seg_claim_iters=0)
    d.update(params)
    prm = OrcParams(25.0, 1, 1, d["cav_cap0"], d["cav_cap_max"], d["cand_budget"],
                    d["max_rounds"], d["priority"], d["cand_factor"], d["cand_min"],
                    d["claim_iters"], d["mixed"], d["shared_ring"], d["mix_below"],
                    d["seg_claim_iters"], d["deletion"])
    st = OrcStats()
    L = orc()
    r = L.orc_refine_batch(len(P), _p(P), len(T), _p(T), _p(N), len(S), _p(S), _p(I), k,
                           _p(th), _p(off), _p(md), C.byref(prm), C.byref(st))
    np_, nt, ns = C.c_int(), C.c_int(), C.c_int()
    L.orc_result_counts(r, C.byref(np_), C.byref(nt), C.byref(ns))
    Po = np.empty((np_.value, 2))
    To = np.empty((nt.value, 3), np.int32)
    No = np.empty((nt.value, 3), np.int32)
    So = np.empty((ns.value, 2), np.int32)
    pf = np.empty(np_.value, np.uint8)
    Io = np.empty(nt.value, np.int32)
    L.orc_result_copy(r, _p(Po), _p(To), _p(No), _p(So), _p(pf))
    L.orc_result_inst(r, _p(Io))
    L.orc_result_free(r)
    return split_instances([len(x) for x in Ps], Po, To, No, So, pf, Io), st.as_dict()


def steiner_point(apex, org, dest, theta, offcenter=1):
    """FP64 circumcenter / off-center as the refinement computes it
    (bit-identical to the device construction); None if degenerate."""
    a, o, d = (np.ascontiguousarray(x, np.float64) for x in (apex, org, dest))
    out = np.empty(2)
    ok = orc().orc_steiner_point(_p(a), _p(o), _p(d), float(theta), int(offcenter), _p(out))
    return out if ok else None


def encroaches(a, b, p, theta, ruppert=1):
    """The refinement's encroachment rule for segment a-b and point p."""
    a, b, p = (np.ascontiguousarray(x, np.float64) for x in (a, b, p))
    return bool(orc().orc_encroaches_xy(_p(a), _p(b), _p(p), float(theta), int(ruppert)))


def split_point(a, b, acuteorg=0, acutedest=0):
    """The refinement's split point of segment a-b, or None when it would
    coincide with an endpoint."""
    a, b = (np.ascontiguousarray(x, np.float64) for x in (a, b))
    out = np.zeros(2)
    ok = orc().orc_split_point_xy(_p(a), _p(b), int(acuteorg), int(acutedest), _p(out))
    return out if ok else None


def bad_angle(apex, org, dest, theta):
    """The refinement's quality test (no MPW exemption)."""
    a, o, d = (np.ascontiguousarray(x, np.float64) for x in (apex, org, dest))
    return bool(orc().orc_bad_angle_xy(_p(a), _p(o), _p(d), float(theta)))


def steiner_point_exact(apex, org, dest, theta, offcenter=1):
    """The same construction in exact rational arithmetic (the off-center
    constant is the FP64 value the refinement uses): circumcenter relative to
    org, off-center on the shortest edge's bisector when it is closer
    (triangle.cpp:6523-6618)."""
    from fractions import Fraction as F
    ga = np.cos(theta * np.pi / 180.0)
    offc = F(0.0) if ga == 1.0 else F(float(0.475 * np.sqrt((1.0 + ga) / (1.0 - ga))))
    A = (F(float(apex[0])), F(float(apex[1])))
    O = (F(float(org[0])), F(float(org[1])))
    D = (F(float(dest[0])), F(float(dest[1])))
    xdo, ydo = D[0] - O[0], D[1] - O[1]
    xao, yao = A[0] - O[0], A[1] - O[1]
    dodist = xdo * xdo + ydo * ydo
    aodist = xao * xao + yao * yao
    dadist = (D[0] - A[0]) ** 2 + (D[1] - A[1]) ** 2
    det = xdo * yao - xao * ydo
    den = F(1, 2) / det
    dx = (yao * dodist - ydo * aodist) * den
    dy = (xdo * aodist - xao * dodist) * den
    if offcenter and offc > 0:
        if dodist < aodist and dodist < dadist:
            ox, oy = F(1, 2) * xdo - offc * ydo, F(1, 2) * ydo + offc * xdo
            if ox * ox + oy * oy < dx * dx + dy * dy:
                dx, dy = ox, oy
        elif aodist < dadist:
            ox, oy = F(1, 2) * xao + offc * yao, F(1, 2) * yao - offc * xao
            if ox * ox + oy * oy < dx * dx + dy * dy:
                dx, dy = ox, oy
        else:
            ox = F(1, 2) * (A[0] - D[0]) - offc * (A[1] - D[1])
            oy = F(1, 2) * (A[1] - D[1]) + offc * (A[0] - D[0])
            if ox * ox + oy * oy < (dx - xdo) ** 2 + (dy - ydo) ** 2:
                dx, dy = xdo + ox, ydo + oy
    return (O[0] + dx, O[1] + dy)


def cavity(points, triangles, neighbors, segments, t0, c, cap=4096):
    """Bowyer-Watson cavity of point c (oracle build_cavity, seeded by the
    walk from triangle t0); sorted triangle ids, or None on failure."""
    P = np.ascontiguousarray(points, np.float64).reshape(-1, 2)
    T = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
    N = np.ascontiguousarray(neighbors, np.int32).reshape(-1, 3)
    S = np.ascontiguousarray(segments, np.int32).reshape(-1, 2)
    cc = np.ascontiguousarray(c, np.float64)
    out = np.empty(cap, np.int32)
    n = orc().orc_cavity(len(P), _p(P), len(T), _p(T), _p(N), len(S), _p(S), int(t0), _p(cc),
                         int(cap), _p(out))
    return None if n < 0 else out[:n].copy()


VIOLATION_NAMES = ["range", "neighbors", "orientation", "unsegmented_boundary",
                   "delaunay", "subseg_not_edge", "segment_not_covered",
                   "input_point_moved", "quality", "nonmanifold"]


def check_mesh(points, triangles, neighbors, segments, points_in, segments_in,
               theta, seg_tol=1e-9):
    """Invariant checker; returns (total, {name: count}, min_angle, n_exempt)."""
    P = np.ascontiguousarray(points, np.float64).reshape(-1, 2)
    T = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
    N = np.ascontiguousarray(neighbors, np.int32).reshape(-1, 3)
    S = np.ascontiguousarray(segments, np.int32).reshape(-1, 2)
    Pi = np.ascontiguousarray(points_in, np.float64).reshape(-1, 2)
    Si = np.ascontiguousarray(segments_in, np.int32).reshape(-1, 2)
    viol = np.zeros(10, np.int64)
    mina = C.c_double()
    ex = C.c_longlong()
    tot = orc().orc_check_mesh(len(P), _p(P), len(T), _p(T), _p(N), len(S), _p(S), len(Pi),
                               _p(Pi), len(Si), _p(Si), float(theta), float(seg_tol),
                               _p(viol), C.byref(mina), C.byref(ex))
    return tot, dict(zip(VIOLATION_NAMES, viol.tolist())), mina.value, ex.value


def orient2d(a, b, c):
    A, B, Cc = (np.ascontiguousarray(x, np.float64) for x in (a, b, c))
    return orc().orc_orient2d(_p(A), _p(B), _p(Cc))


def incircle(a, b, c, d):
    A, B, Cc, D = (np.ascontiguousarray(x, np.float64) for x in (a, b, c, d))
    return orc().orc_incircle(_p(A), _p(B), _p(Cc), _p(D))


def random_pslg(n, nseg, L=1000.0, maxlen=None, seed=1, dist="uniform"):
    """Small seeded PSLG for tests (numpy; exact-sign crossing test through
    the oracle's orient2d).  Shared endpoints allowed."""
    rng = np.random.default_rng(seed)
    if dist == "uniform":
        pts = rng.uniform(0, L, size=(n, 2))
    else:
        k = max(1, n // 200)
        centers = rng.uniform(0.1 * L, 0.9 * L, size=(k, 2))
        pts = centers[rng.integers(0, k, n)] + rng.normal(0, 0.05 * L, size=(n, 2))
        pts = np.clip(pts, 0, L)
    segs = []
    tries = 0
    while len(segs) < nseg and tries < nseg * 100:
        tries += 1
        i, j = (int(x) for x in rng.integers(0, n, 2))
        if i == j:
            continue
        if maxlen and np.hypot(*(pts[i] - pts[j])) > maxlen:
            continue
        ok = True
        for a, b in segs:
            if {a, b} & {i, j}:
                if {a, b} == {i, j}:
                    ok = False
                    break
                continue
            o1 = orient2d(pts[i], pts[j], pts[a]) * orient2d(pts[i], pts[j], pts[b])
            o2 = orient2d(pts[a], pts[b], pts[i]) * orient2d(pts[a], pts[b], pts[j])
            if o1 <= 0 and o2 <= 0:
                ok = False
                break
        if ok:
            segs.append((i, j))
    return pts, np.array(segs, dtype=np.int32).reshape(-1, 2)
