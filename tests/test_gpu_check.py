"""GPU invariant checker (gcdt_check_mesh / gcdt_session_check) against the
CPU checker of the oracle: the same violation counts, category by category,
on refined meshes (all zero) and on meshes with faults planted on the CPU --
a flipped edge (Delaunay), a broken neighbour link (symmetry), a reversed
triangle (orientation), a dropped subsegment (boundary / subsegment edge,
segment coverage), a moved input point, a moved Steiner point (quality), an
index out of range and a duplicated triangle (edge multiplicity) -- and the
same Miller-Pav-Walkington exemptions and smallest angle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# GPU category -> CPU checker category
PAIRS = {"range": "range", "symmetry": "neighbors", "orientation": "orientation",
         "boundary": "unsegmented_boundary", "delaunay": "delaunay",
         "subsegment": "subseg_not_edge", "moved": "input_point_moved",
         "manifold": "nonmanifold", "coverage": "segment_not_covered", "quality": "quality"}


def _refined(gpu, oracle, n=3000, ns=150, seed=101, **kw):
    pts, segs = oracle.random_pslg(n, ns, L=1000.0, maxlen=200.0, seed=seed)
    P, T, N, S = oracle.triangle_cdt(pts, segs)
    r = gpu.refine(P, T, N, S, theta=25.0, **kw)
    return P, S, r


def _same(gpu, oracle, Po, To, No, So, P, S, theta=25.0):
    g = gpu.check_mesh(Po, To, No, So, points_in=P, segments_in=S, theta=theta)
    g.pop("exact_signs")
    tot, cpu, mina, nex = oracle.check_mesh(Po, To, No, So, P, S, theta)
    if cpu["range"] == 0:
        for k, c in PAIRS.items():
            assert g[k] == cpu[c], (k, g[k], cpu[c])
        assert g.pop("exempt") == nex
        assert abs(g.pop("min_angle") - mina) < 1e-6, (mina, g)
    else:
        assert g["range"] == cpu["range"]
        g.pop("exempt")
        g.pop("min_angle")
    return g


def test_refined_mesh_passes(gpu, oracle):
    P, S, r = _refined(gpu, oracle, claim_iters=3, shared_ring=1)
    g = _same(gpu, oracle, r.points, r.triangles, r.neighbors, r.segments, P, S)
    assert all(v == 0 for v in g.values()), g


def test_session_check_in_hbm(gpu, oracle):
    pts, segs = oracle.random_pslg(20000, 1000, L=1000.0, maxlen=200.0, seed=102)
    P, T, N, S = oracle.triangle_cdt(pts, segs)
    s = gpu.Session(P, T, N, S, theta=28.0, claim_iters=3, pred_mode=1)
    try:
        s.run()
        g = s.check()
        r = s.download()
    finally:
        s.close()
    assert gpu.violations(g) == 0, g
    assert g["min_angle"] >= 28.0 - 1e-9
    h = gpu.check_mesh(r.points, r.triangles, r.neighbors, r.segments, points_in=P,
                       segments_in=S, theta=28.0)
    assert h == g


def test_planted_faults(gpu, oracle):
    P, S, r = _refined(gpu, oracle, n=1500, ns=80, seed=103)
    Po, To, No, So = r.points, r.triangles, r.neighbors, r.segments
    segset = {tuple(sorted(s)) for s in So.tolist()}
    # flip an interior non-segment edge
    done = False
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
                g = _same(gpu, oracle, Po, T2, No.copy(), So, P, S)
                assert g["delaunay"] + g["symmetry"] > 0
                done = True
                break
        if done:
            break
    # break one neighbour link
    N2 = No.copy()
    t = int(np.nonzero((No >= 0).all(1))[0][0])
    N2[t, 0] = N2[t, 1]
    g = _same(gpu, oracle, Po, To, N2, So, P, S)
    assert g["symmetry"] > 0
    # reverse one triangle
    T3 = To.copy()
    T3[5] = T3[5, ::-1]
    g = _same(gpu, oracle, Po, T3, No, So, P, S)
    assert g["orientation"] >= 1
    # drop a subsegment on the hull: a boundary edge without a subsegment
    g = _same(gpu, oracle, Po, To, No, So[1:], P, S)
    assert g["boundary"] + g["subsegment"] >= 0
    # a subsegment that is not an edge
    S4 = So.copy()
    S4[0] = [0, len(P) // 2]
    _same(gpu, oracle, Po, To, No, S4, P, S)
    # drop a subsegment from the middle of an input segment's chain: the chain
    # breaks (coverage) and the edge becomes an unsegmented interior edge
    inner = [q for q in range(len(So)) if So[q, 0] >= len(P) and So[q, 1] >= len(P)]
    S5 = np.delete(So, inner[0], axis=0)
    g = _same(gpu, oracle, Po, To, No, S5, P, S)
    assert g["coverage"] >= 1
    # move a Steiner point halfway towards a neighbour: skinny triangles
    t = int(np.nonzero((To >= len(P)).any(1))[0][0])
    v = int(To[t][To[t] >= len(P)][0])
    w = int(To[t][(list(To[t]).index(v) + 1) % 3])
    P6 = Po.copy()
    P6[v] = 0.5 * (Po[v] + Po[w])
    g = _same(gpu, oracle, P6, To, No, So, P, S)
    assert g["quality"] >= 1
    # move an input point by one ulp
    P5 = Po.copy()
    P5[0, 0] = np.nextafter(P5[0, 0], 1e9)
    g = _same(gpu, oracle, P5, To, No, So, P, S)
    assert g["moved"] == 1
    # duplicate a triangle: three triangles on its edges
    T6 = np.concatenate([To, To[:1]])
    N6 = np.concatenate([No, No[:1]])
    g = _same(gpu, oracle, Po, T6, N6, So, P, S)
    assert g["manifold"] == 3
    # an index out of range stops the other checks
    T7 = To.copy()
    T7[0, 0] = len(Po) + 5
    g = gpu.check_mesh(Po, T7, No, So, points_in=P, segments_in=S, theta=25.0)
    assert g["range"] == 1


def test_mpw_exemptions_match(gpu, oracle):
    """Small input angles: the refinement leaves Miller-Pav-Walkington-exempt
    triangles near them; both checkers find the same exemptions and no
    violation."""
    pts = np.array([[0.0, 0.0], [100.0, 0.0], [100.0, 100.0 * np.tan(np.radians(10))],
                    [100.0, -100.0 * np.tan(np.radians(7))], [-40.0, 30.0], [-30.0, -50.0]])
    segs = np.array([[0, 1], [0, 2], [0, 3]], np.int32)
    P, T, N, S = oracle.triangle_cdt(pts, segs)
    r = gpu.refine(P, T, N, S, theta=25.0)
    g = gpu.check_mesh(r.points, r.triangles, r.neighbors, r.segments, points_in=P,
                       segments_in=S, theta=25.0)
    tot, cpu, mina, nex = oracle.check_mesh(r.points, r.triangles, r.neighbors, r.segments,
                                            P, S, 25.0)
    assert nex > 0 and g["exempt"] == nex
    assert tot == 0 and gpu.violations(g) == 0, (cpu, g)


def test_cocircular_lattice_exact_stage(gpu, oracle):
    """A jittered lattice with cocircular groups: Delaunay decisions that the
    filters leave undecided are settled by the exact stage, with the CPU
    checker's result."""
    P, S2 = gpu.generate_pslg(20000, 400, L=1000.0 * 0.02 ** 0.5, maxlen=20.0, dist=2,
                              min_angle=5.0, seed=2)
    T, N, SS = gpu.build_cdt(P, S2)
    g = _same(gpu, oracle, P, T, N, SS, P, SS)
    # the unrefined CDT: only the quality category may count
    g.pop("quality")
    assert all(v == 0 for v in g.values()), g


def test_exact_stage_on_cocircular_grid(gpu, oracle):
    """An integer grid: every interior diagonal has an exactly cocircular
    opposite vertex (incircle = 0), which neither filter can certify; the exact
    stage settles each one as "not inside"."""
    g = 30
    xs, ys = np.meshgrid(np.arange(g, dtype=np.float64), np.arange(g, dtype=np.float64))
    pts = np.stack([xs.ravel(), ys.ravel()], 1)
    T, N, S = gpu.build_cdt(pts, np.zeros((0, 2), np.int32))
    P = pts
    c = gpu.check_mesh(P, T, N, S, points_in=P)
    assert c.pop("exact_signs") >= (g - 1) ** 2 // 2
    tot, cpu, _, _ = oracle.check_mesh(P, T, N, S, P, S, 25.0)
    for k, cc in PAIRS.items():
        if k not in gpu.SEGMENT_CHECKS:
            assert c[k] == cpu[cc], (k, c[k], cpu[cc])
    assert all(v == 0 for v in c.values()), c
