"""Exact predicates, checked on the CPU against what fixes them:
exact rational arithmetic (Python fractions), the reference's own Triangle
predicates (triangle.cpp:5244 counterclockwise, 5892 incircle, compiled from
the reference source with exactinit() called), and degenerate closed cases.
Covers the oracle's big-integer predicates and the product's expansion
predicates evaluated on the host (same source as the kernels)."""
from fractions import Fraction as F

import numpy as np
import pytest


def exact_orient(a, b, c):
    ax, ay, bx, by, cx, cy = (F(v) for v in (*a, *b, *c))
    d = (ax - cx) * (by - cy) - (ay - cy) * (bx - cx)
    return (d > 0) - (d < 0)


def exact_incircle(a, b, c, d):
    ax, ay, bx, by, cx, cy, dx, dy = (F(v) for v in (*a, *b, *c, *d))
    adx, ady, bdx, bdy, cdx, cdy = ax - dx, ay - dy, bx - dx, by - dy, cx - dx, cy - dy
    det = ((adx * adx + ady * ady) * (bdx * cdy - cdx * bdy)
           + (bdx * bdx + bdy * bdy) * (cdx * ady - adx * cdy)
           + (cdx * cdx + cdy * cdy) * (adx * bdy - bdx * ady))
    return (det > 0) - (det < 0)


def near_degenerate_queries(rng, n):
    """Points nudged by a few ulps off exactly collinear / cocircular
    configurations, with coordinates of mixed magnitude so that coordinate
    differences are often inexact (exercises every exact stage)."""
    out = []
    for _ in range(n):
        kind = rng.integers(0, 4)
        s = 10.0 ** rng.integers(-3, 7)
        if kind == 0:  # collinear along a random line, then nudge
            p = rng.uniform(-1, 1, 2) * s
            d = rng.uniform(-1, 1, 2) * s
            t = rng.uniform(-2, 2, 3)
            pts = [p + ti * d for ti in t]
            pts[2] = np.nextafter(pts[2], pts[2] + rng.choice([-1, 1]) * s)
            dpt = rng.uniform(-1, 1, 2) * s
            out.append([*pts[0], *pts[1], *pts[2], *dpt])
        elif kind == 1:  # cocircular on an integer circle (exact), scaled by a power of 2
            sc = 2.0 ** int(rng.integers(-20, 20))
            pts = [(5, 0), (0, 5), (-5, 0), (3, 4), (4, -3), (-3, -4), (0, -5), (-4, 3)]
            idx = rng.permutation(len(pts))[:4]
            q = [np.array(pts[i], float) * sc + 1e6 * sc for i in idx]
            if rng.integers(0, 2):
                q[3] = np.nextafter(q[3], q[3] + rng.choice([-1, 1]))
            out.append([*q[0], *q[1], *q[2], *q[3]])
        elif kind == 2:  # random circle, 4 points rounded from the exact circle
            c = rng.uniform(-1, 1, 2) * s
            r = abs(rng.uniform(0.1, 1)) * s
            a = np.sort(rng.uniform(0, 2 * np.pi, 4))
            q = [c + r * np.array([np.cos(t), np.sin(t)]) for t in a]
            out.append([*q[0], *q[1], *q[2], *q[3]])
        else:  # tiny triangle far from the origin (inexact differences)
            base = rng.uniform(0.5, 2.0, 2) * s * 1e3
            q = [base + rng.uniform(-1, 1, 2) * s * 1e-9 for _ in range(4)]
            out.append([*q[0], *q[1], *q[2], *q[3]])
    return np.array(out)


@pytest.fixture(scope="module")
def queries():
    rng = np.random.default_rng(12345)
    rnd = rng.uniform(-100, 100, size=(300, 8))
    return np.concatenate([rnd, near_degenerate_queries(rng, 700)])


def test_oracle_predicates_exact(oracle, queries):
    for q in queries:
        a, b, c, d = q[0:2], q[2:4], q[4:6], q[6:8]
        assert oracle.orient2d(a, b, c) == exact_orient(a, b, c)
        assert oracle.incircle(a, b, c, d) == exact_incircle(a, b, c, d)


def test_product_host_predicates_exact(queries):
    import gpu_cdt_refine_amd as g
    for mode in (0, 1):
        o, ic = g.host_predicates(queries, pred_mode=mode)
        for k, q in enumerate(queries):
            a, b, c, d = q[0:2], q[2:4], q[4:6], q[6:8]
            assert o[k] == exact_orient(a, b, c), (mode, k, q)
            assert ic[k] == exact_incircle(a, b, c, d), (mode, k, q)


def test_oracle_matches_reference_triangle_predicates(oracle, queries):
    if not __import__("os").path.exists(oracle.TRIREF_SO):
        pytest.skip("reference Triangle not built")
    L = oracle.triref()
    for q in queries:
        a, b, c, d = (np.ascontiguousarray(q[i:i + 2]) for i in (0, 2, 4, 6))
        ro = L.triref_orient2d(a.ctypes.data, b.ctypes.data, c.ctypes.data)
        ri = L.triref_incircle(a.ctypes.data, b.ctypes.data, c.ctypes.data, d.ctypes.data)
        assert oracle.orient2d(a, b, c) == (ro > 0) - (ro < 0)
        assert oracle.incircle(a, b, c, d) == (ri > 0) - (ri < 0)


def test_degenerate_closed_forms(oracle):
    assert oracle.orient2d([0, 0], [1, 1], [2, 2]) == 0
    assert oracle.orient2d([0, 0], [1, 0], [0, 1]) == 1
    assert oracle.orient2d([0, 0], [0, 1], [1, 0]) == -1
    assert oracle.incircle([5, 0], [0, 5], [-5, 0], [3, 4]) == 0      # cocircular
    assert oracle.incircle([5, 0], [0, 5], [-5, 0], [0, 0]) == 1      # centre inside
    assert oracle.incircle([5, 0], [0, 5], [-5, 0], [6, 6]) == -1
    assert oracle.incircle([0, 5], [5, 0], [-5, 0], [0, 0]) == -1     # CW triangle flips sign


def tiny_scale_queries(rng, n):
    """Queries whose coordinates are scaled by 2^-120 (FP32 differences and
    products would be subnormal or zero), and mixed-scale queries with one
    point far away and three within ~2^-70 of each other (FP32 products of
    the close points' differences underflow while the permanent does not)."""
    base = np.concatenate([rng.uniform(-1, 1, size=(n // 2, 8)),
                           near_degenerate_queries(rng, n // 2)])
    scaled = base * 2.0 ** -120
    mixed = []
    for _ in range(n):
        c = rng.uniform(-1, 1, 2) * 2.0 ** -40
        pts = [c + rng.uniform(-1, 1, 2) * 2.0 ** -70 for _ in range(3)]
        far = c + rng.uniform(-1, 1, 2) * 2.0 ** 20
        k = rng.integers(0, 4)
        pts.insert(int(k), far)
        mixed.append(np.concatenate(pts))
    return np.concatenate([scaled, np.array(mixed)])


def test_fp32_stage_exact_at_tiny_scales():
    """The FP32 semi-static stage defers (never decides wrongly) when
    differences or products leave FP32's normal range."""
    import gpu_cdt_refine_amd as g
    rng = np.random.default_rng(99)
    q = tiny_scale_queries(rng, 600)
    o, ic = g.host_predicates(q, pred_mode=1)
    for k, r in enumerate(q):
        a, b, c, d = r[0:2], r[2:4], r[4:6], r[6:8]
        assert o[k] == exact_orient(a, b, c), (k, r)
        assert ic[k] == exact_incircle(a, b, c, d), (k, r)


def near_collinear_triangles(rng, n):
    """(p0, p1, p2) triples: well-shaped ones and apexes within 1e-6 .. 1e-15
    (relative) of the line through the other two, at several scales."""
    out = []
    for k in range(n):
        scale = 10.0 ** rng.integers(-6, 7)
        a = rng.uniform(0, 1, 2) * scale
        b = a + rng.uniform(-1, 1, 2) * scale
        f = rng.uniform(0.05, 0.95)
        c = a + f * (b - a)
        if k % 6 == 0:
            c = rng.uniform(0, 1, 2) * scale
        elif k % 6 == 1:
            a, b = np.round(a), np.round(b)
            c = np.round(a + f * (b - a))
        elif k % 6 > 2:
            c = c + np.array([-(b - a)[1], (b - a)[0]]) * 10.0 ** rng.uniform(-15, -6)
        out.append(np.concatenate([c, a, b]))
    return np.array(out)


def test_kernel_constructions_match_triangle(oracle):
    """The kernels' adaptive orient2d value and Steiner point (their source
    run on the host) bit-exact against the reference's Triangle
    counterclockwise() and findcircumcenter() and the oracle."""
    import gpu_cdt_refine_amd as g
    if not __import__("os").path.exists(oracle.TRIREF_SO):
        pytest.skip("reference Triangle not built")
    import ctypes as C
    R = oracle.triref()
    L = oracle.orc()
    rng = np.random.default_rng(5)
    tris = near_collinear_triangles(rng, 3000)
    for theta, off in ((25.0, 1), (25.0, 0), (30.0, 1)):
        ccw, stp, ok = g.host_constructions(tris, theta, off)
        for k, q in enumerate(tris):
            p0, p1, p2 = (np.ascontiguousarray(q[i:i + 2]) for i in (0, 2, 4))
            r = R.triref_orient2d(oracle._p(p0), oracle._p(p1), oracle._p(p2))
            assert np.float64(ccw[k]).view(np.uint64) == np.float64(r).view(np.uint64), (k, q)
            if ccw[k] <= 0:
                continue  # findcircumcenter is defined on CCW triangles
            out = np.zeros(2)
            okr = L.orc_steiner_point(oracle._p(p0), oracle._p(p1), oracle._p(p2), C.c_double(theta),
                                      off, oracle._p(out))
            assert ok[k] == okr, (k, q)
            if not ok[k]:
                continue
            ref = np.zeros(2)
            R.triref_findcircumcenter(oracle._p(p1), oracle._p(p2), oracle._p(p0), C.c_double(theta),
                                      off, oracle._p(ref))
            assert np.array_equal(stp[k].view(np.uint64), ref.view(np.uint64)), (k, q, theta, off)
            assert np.array_equal(stp[k].view(np.uint64), out.view(np.uint64)), (k, q, theta, off)
