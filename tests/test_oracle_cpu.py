"""The CPU oracle of the refinement rounds, pinned by what the mathematics
and the reference's own code fix: exact mesh invariants (constrained
Delaunay, segment conformity, minimum angle with the Miller-Pav-Walkington
exemption), the vertex count of the reference's Triangle -q on the same
PSLG, Triangle's own circumcenter/off-center routine (findcircumcenter,
triangle.cpp:6523), and run-to-run determinism."""
import os

import numpy as np
import pytest


def _cdt(oracle, n, ns, seed, L=1000.0, maxlen=200.0):
    pts, segs = oracle.random_pslg(n, ns, L=L, maxlen=maxlen, seed=seed)
    return oracle.triangle_cdt(pts, segs)


@pytest.mark.parametrize("theta,off,mode", [(25, 1, 1), (25, 1, 0), (30, 1, 1), (25, 0, 1),
                                            (20.7, 0, 0), (30, 0, 0)])
def test_oracle_invariants(oracle, theta, off, mode):
    P, T, N, S = _cdt(oracle, 1500, 75, 3)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, theta=theta, offcenter=off, mode=mode)
    assert st["status"] == 0 and st["giveups"] == 0
    tot, viol, mina, ex = oracle.check_mesh(Po, To, No, So, P, S, theta)
    assert tot == 0, viol
    assert mina >= theta - 1e-7
    assert np.array_equal(Po[:len(P)].view(np.uint64), P.view(np.uint64))


def test_checker_catches_planted_faults(oracle):
    """A mistake planted in the mesh (on the CPU) must be reported."""
    P, T, N, S = _cdt(oracle, 800, 40, 5)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, theta=25.0)
    # flip one interior non-segment edge: breaks Delaunay (and nothing else)
    segset = {tuple(sorted(s)) for s in So.tolist()}
    for t in range(len(To)):
        for j in range(3):
            u = No[t, j]
            a, b = To[t, (j + 1) % 3], To[t, (j + 2) % 3]
            if u > t and tuple(sorted((a, b))) not in segset:
                k = [kk for kk in range(3) if No[u, kk] == t][0]
                c, d = To[t, j], To[u, k]
                T2 = To.copy()
                T2[t] = [c, a, d]
                T2[u] = [d, b, c]
                N2 = No.copy()
                tot, viol, _, _ = oracle.check_mesh(Po, T2, N2, So, P, S, 25.0)
                assert tot > 0
                break
        else:
            continue
        break
    # move an input point
    P2 = Po.copy()
    P2[0, 0] = np.nextafter(P2[0, 0], 1e9)
    tot, viol, _, _ = oracle.check_mesh(P2, To, No, So, P, S, 25.0)
    assert viol["input_point_moved"] == 1
    # drop a subsegment: its input segment is no longer covered
    tot, viol, _, _ = oracle.check_mesh(Po, To, No, So[1:], P, S, 25.0)
    assert viol["segment_not_covered"] > 0


@pytest.mark.parametrize("theta,mode,band", [(25, 1, 0.03), (30, 1, 0.04), (25, 0, 0.03)])
def test_vertex_count_band_vs_triangle(oracle, theta, mode, band):
    """Steiner count within a stated band of the reference's Triangle -q."""
    if not os.path.exists(oracle.TRIREF_SO):
        pytest.skip("reference Triangle not built")
    P, T, N, S = _cdt(oracle, 4000, 200, 11)
    Pt, Tt, Nt, St = oracle.triangle_refine(P, T, S, theta, mode)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, theta=theta, mode=mode)
    ratio = len(Po) / len(Pt)
    assert abs(ratio - 1.0) < band, ratio


def test_steiner_point_matches_triangle_findcircumcenter(oracle):
    """Bit-exact against the reference's Triangle findcircumcenter for
    well-shaped and skinny triangles (org, dest, apex = v1, v2, v0)."""
    if not os.path.exists(oracle.TRIREF_SO):
        pytest.skip("reference Triangle not built")
    import ctypes as C
    L = oracle.orc()
    R = oracle.triref()
    rng = np.random.default_rng(3)
    n_off = 0
    for k in range(2000):
        p = rng.uniform(0, 1000, size=(3, 2))
        if k % 2:
            p[2] = p[0] + (p[1] - p[0]) * rng.uniform(0.2, 0.8) + rng.normal(0, 1e-3, 2)
        # CCW order (apex, org, dest)
        if oracle.orient2d(p[0], p[1], p[2]) < 0:
            p[[1, 2]] = p[[2, 1]]
        for theta, off in ((25.0, 1), (30.0, 1), (25.0, 0)):
            out = np.zeros(2)
            apex, org, dest = (np.ascontiguousarray(x) for x in p)
            ok = L.orc_steiner_point(apex.ctypes.data_as(C.c_void_p), org.ctypes.data_as(C.c_void_p),
                                     dest.ctypes.data_as(C.c_void_p), C.c_double(theta),
                                     off, out.ctypes.data_as(C.c_void_p))
            if not ok:
                continue
            ref = np.zeros(2)
            R.triref_findcircumcenter(org.ctypes.data_as(C.c_void_p), dest.ctypes.data_as(C.c_void_p),
                                      apex.ctypes.data_as(C.c_void_p), C.c_double(theta), off,
                                      ref.ctypes.data_as(C.c_void_p))
            assert np.array_equal(out.view(np.uint64), ref.view(np.uint64)), (p, theta, off)
            n_off += off
    assert n_off > 0


def _near_collinear(rng, k):
    """Triangles whose apex sits within 1e-6 .. 1e-15 (relative) of the
    line through org and dest, at several coordinate scales; some exactly
    collinear, some on integer coordinates (exact differences)."""
    scale = 10.0 ** rng.integers(-6, 7)
    a = rng.uniform(0, 1, 2) * scale
    b = a + rng.uniform(-1, 1, 2) * scale
    f = rng.uniform(0.05, 0.95)
    c = a + f * (b - a)
    if k % 5 == 0:
        pass  # the FP64 point nearest the line: often exactly collinear
    elif k % 5 == 1:
        a, b = np.round(a), np.round(b)
        c = np.round(a + f * (b - a))
    else:
        d = np.array([-(b - a)[1], (b - a)[0]])
        c = c + d * 10.0 ** rng.uniform(-15, -6)
    return np.array([c, a, b])


def test_ccw_value_matches_triangle_counterclockwise(oracle):
    """The adaptive orient2d value the circumcenter divides by, bit-exact
    against the reference's Triangle counterclockwise() on near-collinear
    triples (every adaptive stage reached), and sign-exact against rationals."""
    if not os.path.exists(oracle.TRIREF_SO):
        pytest.skip("reference Triangle not built")
    from fractions import Fraction as F
    L = oracle.orc()
    R = oracle.triref()
    rng = np.random.default_rng(17)
    nonfilter = 0
    for k in range(4000):
        p = _near_collinear(rng, k)
        pa, pb, pc = (np.ascontiguousarray(x) for x in p)
        v = L.orc_ccw_value(oracle._p(pa), oracle._p(pb), oracle._p(pc))
        r = R.triref_orient2d(oracle._p(pa), oracle._p(pb), oracle._p(pc))
        assert np.float64(v).view(np.uint64) == np.float64(r).view(np.uint64), (p, v, r)
        ex = ((F(pa[0]) - F(pc[0])) * (F(pb[1]) - F(pc[1])) -
              (F(pa[1]) - F(pc[1])) * (F(pb[0]) - F(pc[0])))
        assert (v > 0) == (ex > 0) and (v < 0) == (ex < 0), (p, v, ex)
        det = (pa[0] - pc[0]) * (pb[1] - pc[1]) - (pa[1] - pc[1]) * (pb[0] - pc[0])
        nonfilter += v != det
    assert nonfilter > 100  # the adaptive stages were exercised


def test_steiner_point_near_collinear_matches_triangle(oracle):
    """findcircumcenter on near-collinear triangles (where the raw FP64
    determinant loses its digits and Triangle's adaptive denominator takes
    over): bit-exact against the reference's Triangle."""
    if not os.path.exists(oracle.TRIREF_SO):
        pytest.skip("reference Triangle not built")
    import ctypes as C
    L = oracle.orc()
    R = oracle.triref()
    rng = np.random.default_rng(23)
    checked = 0
    for k in range(6000):
        p = _near_collinear(rng, k)
        if oracle.orient2d(p[0], p[1], p[2]) < 0:
            p[[1, 2]] = p[[2, 1]]
        apex, org, dest = (np.ascontiguousarray(x) for x in p)
        for theta, off in ((25.0, 1), (25.0, 0)):
            out = np.zeros(2)
            ok = L.orc_steiner_point(oracle._p(apex), oracle._p(org), oracle._p(dest),
                                     C.c_double(theta), off, oracle._p(out))
            if not ok:
                continue
            ref = np.zeros(2)
            R.triref_findcircumcenter(oracle._p(org), oracle._p(dest), oracle._p(apex),
                                      C.c_double(theta), off, oracle._p(ref))
            assert np.array_equal(out.view(np.uint64), ref.view(np.uint64)), (p, theta, off, out, ref)
            checked += 1
    assert checked > 4000


def test_oracle_give_ups_are_reported(oracle):
    """The precision-stress family (jittered lattice, cocircular groups,
    lattice-line segments) makes the refinement give up on a few elements;
    the run then ends with status 3 and counts them, and every quality
    violation the checker finds is one of those given-up triangles."""
    import gpu_cdt_refine_amd as g
    P, S = g.generate_pslg(20000, 400, L=1000.0 * 0.02 ** 0.5, maxlen=20.0, dist=2,
                           min_angle=5.0, seed=2)
    T, N, SS = g.build_cdt(P, S)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, SS, theta=25.0)
    assert st["giveups"] > 0
    left = st["giveup_tris_left"] + st["giveup_segs_left"]
    assert st["status"] == (3 if left else 0), st
    tot, viol, mina, ex = oracle.check_mesh(Po, To, No, So, P, SS, 25.0)
    assert viol["quality"] <= st["giveup_tris_left"], (viol, st)
    assert tot == viol["quality"], viol


def test_small_input_angles_terminate(oracle):
    """Two segments meeting at 10 degrees: concentric shells + MPW keep the
    refinement finite; the only sub-theta triangles are MPW-exempt."""
    a = np.deg2rad(10.0)
    pts = np.array([[0, 0], [400, 0], [400 * np.cos(a), 400 * np.sin(a)],
                    [-300, -300], [700, -300], [700, 500], [-300, 500]], float)
    segs = np.array([[0, 1], [0, 2]], np.int32)
    P, T, N, S = oracle.triangle_cdt(pts, segs)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, theta=25.0)
    assert st["status"] == 0
    tot, viol, mina, ex = oracle.check_mesh(Po, To, No, So, P, S, 25.0)
    assert tot == 0, viol
    assert ex > 0


def test_oracle_deterministic(oracle):
    P, T, N, S = _cdt(oracle, 1200, 60, 9)
    r1 = oracle.oracle_refine(P, T, N, S, theta=28.0)
    r2 = oracle.oracle_refine(P, T, N, S, theta=28.0)
    for a, b in zip(r1[:4], r2[:4]):
        assert np.array_equal(a, b)


def test_cavity_equals_brute_force_conflict_set(oracle):
    """In a Delaunay triangulation of a convex point set, the Bowyer-Watson
    cavity of a point is exactly the set of triangles whose circumcircle
    strictly contains it (exact incircle over every triangle)."""
    rng = np.random.default_rng(5)
    pts = rng.uniform(0, 100, size=(400, 2))
    P, T, N, S = oracle.triangle_cdt(pts, np.zeros((0, 2), np.int32))
    checked = 0
    for k in range(60):
        c = rng.uniform(20, 80, size=2)
        if k % 3 == 0:  # on a circumcircle-heavy spot: near a vertex
            c = P[rng.integers(len(P))] + rng.normal(scale=1e-3, size=2)
        cav = oracle.cavity(P, T, N, S, int(rng.integers(len(T))), c)
        if cav is None:
            continue
        brute = [t for t in range(len(T))
                 if oracle.incircle(P[T[t, 0]], P[T[t, 1]], P[T[t, 2]], c) > 0]
        assert cav.tolist() == brute, (k, c)
        checked += 1
    assert checked >= 50


def test_steiner_error_against_exact_rationals(oracle):
    """FP64 Steiner points vs the same construction in exact rational
    arithmetic, on the CDT of the precision-stress input (jittered lattice,
    cocircular points): the error stays within a few ulps of the
    coordinates' magnitude for well-shaped triangles and bounded relative to
    the triangle's size for skinny ones."""
    from fractions import Fraction
    import gpu_cdt_refine_amd as g
    P, S = g.generate_pslg(6000, 120, L=1000.0, maxlen=20.0, dist=2, min_angle=5.0, seed=3)
    T, N, SS = g.build_cdt(P, S)
    rng = np.random.default_rng(0)
    worst_rel = 0.0
    n = 0
    for t in rng.choice(len(T), 400, replace=False):
        a, o, d = P[T[t, 0]], P[T[t, 1]], P[T[t, 2]]
        c = oracle.steiner_point(a, o, d, 25.0, 1)
        if c is None:
            continue
        ex = oracle.steiner_point_exact(a, o, d, 25.0, 1)
        size = max(np.hypot(*(o - a)), np.hypot(*(d - a)), np.hypot(*(d - o)))
        err = max(abs(float(Fraction(c[0]) - ex[0])), abs(float(Fraction(c[1]) - ex[1])))
        worst_rel = max(worst_rel, err / size)
        n += 1
    assert n > 300
    assert worst_rel < 1e-9, worst_rel


@pytest.mark.parametrize("theta,off,mode,ci", [(25, 1, 1, 1), (25, 1, 1, 3), (30, 0, 0, 2),
                                               (20.7, 1, 0, 1)])
def test_mixed_rounds_invariants(oracle, theta, off, mode, ci):
    """The mixed form (encroached subsegments and bad triangles in one pass per
    round) meets the same invariants, needs fewer passes than segment phases +
    bad passes, and its Steiner count stays within 5% of the classic round's."""
    P, T, N, S = _cdt(oracle, 3000, 150, 8)
    kw = dict(theta=theta, offcenter=off, mode=mode, claim_iters=ci)
    Pc, Tc, Nc, Sc, pfc, stc = oracle.oracle_refine(P, T, N, S, **kw)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, mixed=1, **kw)
    assert st["status"] == 0 and st["giveups"] == 0
    tot, viol, mina, ex = oracle.check_mesh(Po, To, No, So, P, S, theta)
    assert tot == 0, viol
    assert mina >= theta - 1e-7
    assert np.array_equal(Po[:len(P)].view(np.uint64), P.view(np.uint64))
    assert st["rounds"] < stc["seg_passes"] + stc["bad_passes"]
    assert abs((len(Po) - len(Pc)) / (len(Pc) - len(P))) < 0.05
    assert st["seg_splits"] + len(S) == len(So)
    assert st["inserted"] + st["seg_splits"] - st["deleted"] + len(P) == len(Po)


@pytest.mark.parametrize("theta,off,mode,ci", [(25, 1, 1, 1), (25, 1, 1, 3), (30, 0, 0, 2)])
def test_shared_ring_invariants(oracle, theta, off, mode, ci):
    """Winners whose rings overlap (cavities disjoint from every other
    winner's cavity and ring): the mesh meets every invariant, needs no more
    passes than disjoint regions, and its Steiner count stays within 5%."""
    P, T, N, S = _cdt(oracle, 3000, 150, 9)
    kw = dict(theta=theta, offcenter=off, mode=mode, claim_iters=ci)
    Pc, Tc, Nc, Sc, pfc, stc = oracle.oracle_refine(P, T, N, S, **kw)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, shared_ring=1, **kw)
    assert st["status"] == 0 and st["giveups"] == 0
    tot, viol, mina, ex = oracle.check_mesh(Po, To, No, So, P, S, theta)
    assert tot == 0, viol
    assert mina >= theta - 1e-7
    assert st["seg_passes"] + st["bad_passes"] <= stc["seg_passes"] + stc["bad_passes"]
    assert abs((len(Po) - len(Pc)) / (len(Pc) - len(P))) < 0.05


@pytest.mark.parametrize("mix_below", [4, 20, 200])
def test_mix_below_invariants(oracle, mix_below):
    """Rounds whose segment phase hands its last few encroached subsegments
    to the bad pass: same invariants, fewer passes, Steiner count within 5%."""
    P, T, N, S = _cdt(oracle, 3000, 300, 10, maxlen=400.0)
    kw = dict(theta=25.0, claim_iters=3, shared_ring=1)
    Pc, Tc, Nc, Sc, pfc, stc = oracle.oracle_refine(P, T, N, S, **kw)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, mix_below=mix_below, **kw)
    assert st["status"] == 0
    tot, viol, mina, ex = oracle.check_mesh(Po, To, No, So, P, S, 25.0)
    assert tot == 0, viol
    assert mina >= 25.0 - 1e-7
    assert abs((len(Po) - len(Pc)) / (len(Pc) - len(P))) < 0.05


def test_oracle_gives_up_triangles_stuck_on_given_up_subsegments(oracle):
    """A batched-family instance (Gaussian clusters, 69,161 points) whose
    input has two segment chains ~1e-7 apart near coordinate 2000: four
    segment splits there cannot form a positively oriented star in FP64 and
    are given up; the bad triangles that then encroach (or walk into) only
    those subsegments are given up as well, so the run converges with status
    3 instead of stalling, and every quality violation left is a given-up
    triangle."""
    import gpu_cdt_refine_amd as g
    n = 69161
    L = 1000.0 * (n / 10000) ** 0.5
    P, S = g.generate_pslg(n, n // 20, L=L, maxlen=0.05 * L, dist=1, min_angle=5.0,
                           seed=7000 + 4009)
    T, N, SS = g.build_cdt(P, S)
    theta = 21.33356473295036
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, SS, theta=theta, claim_iters=3,
                                                  shared_ring=1, priority=1)
    assert st["status"] == 3 and st["giveup_segs_left"] == 4, st
    tot, viol, mina, ex = oracle.check_mesh(Po, To, No, So, P, SS, theta)
    assert 0 < viol["quality"] <= st["giveup_tris_left"], (viol, st)
    assert tot == viol["quality"], viol


def _apex_angle_deg(a, b, p):
    """Angle a-p-b in degrees from atan2 (independent of the dot formula)."""
    ux, uy = a[0] - p[0], a[1] - p[1]
    vx, vy = b[0] - p[0], b[1] - p[1]
    return np.degrees(abs(np.arctan2(ux * vy - uy * vx, ux * vx + uy * vy)))


def test_encroachment_is_the_apex_angle_rule(oracle):
    """Closed form of the encroachment rule (triangle.cpp:7106-7203,
    refine.cu:115-139): Ruppert's diametral circle holds p iff the angle
    a-p-b exceeds 90 degrees; Chew's lens iff it is at least 180 - 2 theta.
    Angles within 1e-7 degrees of a threshold are skipped."""
    rng = np.random.default_rng(41)
    seen = {(r, e): 0 for r in (0, 1) for e in (False, True)}
    for k in range(20000):
        scale = 10.0 ** rng.integers(-3, 6)
        a = rng.uniform(-1, 1, 2) * scale
        b = rng.uniform(-1, 1, 2) * scale
        mid, half = 0.5 * (a + b), 0.5 * np.linalg.norm(b - a)
        p = mid + rng.uniform(-1.3, 1.3, 2) * half
        theta = float(rng.choice([20.7, 25.0, 28.0, 30.0, 33.0]))
        ang = _apex_angle_deg(a, b, p)
        for ruppert in (1, 0):
            thr = 90.0 if ruppert else 180.0 - 2.0 * theta
            if abs(ang - thr) < 1e-7:
                continue
            want = ang > thr
            got = oracle.encroaches(a, b, p, theta, ruppert)
            assert got == want, (a, b, p, theta, ruppert, ang)
            seen[(ruppert, want)] += 1
    assert min(seen.values()) > 500, seen


def test_quality_test_is_the_smallest_angle_rule(oracle):
    """Closed form of testtriangle's angle test (triangle.cpp:7244-7291): a
    triangle is bad iff its smallest angle is below theta (angles within
    1e-7 degrees of theta skipped)."""
    rng = np.random.default_rng(43)
    bad = good = 0
    for k in range(20000):
        scale = 10.0 ** rng.integers(-3, 6)
        t = rng.uniform(-1, 1, (3, 2)) * scale
        if k % 3 == 0:  # skinny ones
            t[2] = t[0] + rng.uniform(0.05, 0.95) * (t[1] - t[0]) + rng.normal(0, 0.02, 2) * scale
        angs = [_apex_angle_deg(t[(i + 1) % 3], t[(i + 2) % 3], t[i]) for i in range(3)]
        theta = float(rng.choice([20.7, 25.0, 30.0]))
        if abs(min(angs) - theta) < 1e-7 or min(angs) < 1e-6:
            continue
        got = oracle.bad_angle(t[0], t[1], t[2], theta)
        assert got == (min(angs) < theta), (t, angs, theta)
        bad += got
        good += not got
    assert bad > 2000 and good > 2000


def test_split_point_rule(oracle):
    """Closed form of the segment split (triangle.cpp:13368-13415,
    refine.cu:2686-2725): the midpoint when no end is shared; otherwise a
    point at a power-of-two distance d from the shared end (the destination
    when both are shared) with len/3 <= d <= len/1.5 (concentric shells).
    The point is on the segment's line within 2 ulps."""
    from fractions import Fraction as F
    rng = np.random.default_rng(47)
    for k in range(6000):
        scale = 10.0 ** rng.integers(-4, 7)
        a = rng.uniform(-1, 1, 2) * scale
        b = a + rng.uniform(-1, 1, 2) * scale * 10.0 ** rng.uniform(-6, 0)
        ao, ad = int(k % 4 in (1, 3)), int(k % 4 in (2, 3))
        p = oracle.split_point(a, b, ao, ad)
        if p is None:
            continue
        L = float(np.linalg.norm(b - a))
        if not (ao or ad):
            assert np.allclose(p, 0.5 * (a + b), rtol=0, atol=8 * np.spacing(np.abs(a).max() + np.abs(b).max()))
        else:
            end = b if ad else a
            d = float(np.linalg.norm(p - end))
            k2 = np.log2(d)
            assert abs(k2 - round(k2)) < 1e-6, (a, b, ao, ad, d)
            assert L / 3 * (1 - 1e-9) <= d <= L / 1.5 * (1 + 1e-9), (L, d)
        # on the segment's line within rounding: exact perpendicular distance
        # at most 2 ulps of the largest coordinate
        A, B, Q = ((F(x[0]), F(x[1])) for x in (a, b, p))
        cross = abs((A[0] - Q[0]) * (B[1] - Q[1]) - (A[1] - Q[1]) * (B[0] - Q[0]))
        ulp = F(float(np.spacing(max(np.abs(a).max(), np.abs(b).max(), np.abs(p).max()))))
        assert cross * cross <= (2 * ulp) ** 2 * ((B[0] - A[0]) ** 2 + (B[1] - A[1]) ** 2), \
            (a, b, p)


@pytest.mark.parametrize("seed,theta", [(1, 25.0), (2, 30.0), (3, 20.7)])
def test_vertex_deletion_pins(oracle, seed, theta):
    """Vertex deletion (Chew mode, Triangle's rule triangle.cpp:13319-13357):
    the mesh stays a constrained Delaunay triangulation meeting every
    invariant (exact checker, so a wrong retriangulation of a star fails);
    the point count accounts for every insertion and deletion, input points
    keep their indices; and the Steiner count moves towards the reference's
    own Triangle -q (which deletes) compared with no deletion."""
    pts, segs = oracle.random_pslg(5000, 250, L=1000.0, maxlen=200.0, seed=seed)
    P, T, N, S = oracle.triangle_cdt(pts, segs)
    P0, T0, N0, S0, pf0, st0 = oracle.oracle_refine(P, T, N, S, theta=theta, mode=0, deletion=0)
    P1, T1, N1, S1, pf1, st1 = oracle.oracle_refine(P, T, N, S, theta=theta, mode=0, deletion=1)
    assert st0["deleted"] == 0 and st1["deleted"] > 0
    assert st1["status"] == 0
    tot, viol, mina, ex = oracle.check_mesh(P1, T1, N1, S1, P, S, theta)
    assert tot == 0, viol
    assert mina >= theta - 1e-7
    assert np.array_equal(P1[:len(P)].view(np.uint64), P.view(np.uint64))
    assert st1["inserted"] + st1["seg_splits"] - st1["deleted"] + len(P) == len(P1)
    assert st1["seg_splits"] + len(S) == len(S1)
    # every output point is used by a triangle (deleted vertices are gone)
    assert len(np.unique(T1)) == len(P1)
    # flips: per insertion cavity size less one or two, per deletion degree
    # less three -- at least the insertions' count
    assert st1["flips"] > st1["inserted"]
    if os.path.exists(oracle.TRIREF_SO):
        tq = len(oracle.triangle_refine(P, T, S, theta, 0)[0])
        assert abs(len(P1) - tq) < abs(len(P0) - tq)


def test_vertex_deletion_modes(oracle):
    """deletion = 1 acts in Chew mode only; in Ruppert mode no Steiner
    point is ever inserted inside a diametral circle, so deletion = 2 finds
    none there and the mesh equals the run without deletion."""
    P, T, N, S = _cdt(oracle, 3000, 150, 12)
    a = oracle.oracle_refine(P, T, N, S, theta=25.0, mode=1, deletion=0)
    b = oracle.oracle_refine(P, T, N, S, theta=25.0, mode=1, deletion=2)
    assert b[5]["deleted"] == 0
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)
    c = oracle.oracle_refine(P, T, N, S, theta=25.0, mode=0, deletion=2)
    d = oracle.oracle_refine(P, T, N, S, theta=25.0, mode=0, deletion=1)
    assert c[5]["deleted"] == d[5]["deleted"] > 0
    for x, y in zip(c[:4], d[:4]):
        assert np.array_equal(x, y)
