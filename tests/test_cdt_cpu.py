"""Host input preparation: the seeded PSLG generator and the CDT builder,
checked against the reference's Triangle "pzQnc" (mesh.cpp:397), whose CDT
is unique for points in general position."""
import os

import numpy as np
import pytest

import gpu_cdt_refine_amd as g


def tri_set(T):
    return {tuple(sorted(t)) for t in np.asarray(T).tolist()}


@pytest.mark.parametrize("n,ns,L,maxlen,dist,seed", [
    (500, 25, 1000.0, 0.0, 0, 1),
    (5000, 250, 1000.0, 200.0, 0, 2),
    (8000, 400, 1e6, 2e4, 1, 3),
    (3000, 150, 1000.0, 100.0, 3, 4),
    (3000, 150, 1000.0, 100.0, 4, 5),
    (3000, 150, 3e4, 3e3, 5, 6),
    (3000, 150, 1000.0, 100.0, 6, 7),
])
def test_cdt_matches_triangle(oracle, n, ns, L, maxlen, dist, seed):
    if not os.path.exists(oracle.TRIREF_SO):
        pytest.skip("reference Triangle not built")
    P, S = g.generate_pslg(n, ns, L=L, maxlen=maxlen, dist=dist, min_angle=5.0, seed=seed)
    T, N, SS = g.build_cdt(P, S)
    Pt, Tt, Nt, St = oracle.triangle_cdt(P, S)
    assert tri_set(T) == tri_set(Tt)
    assert {tuple(sorted(s)) for s in SS.tolist()} == {tuple(sorted(s)) for s in St.tolist()}
    tot, viol, _, _ = oracle.check_mesh(P, T, N, SS, P, SS, 0.0)
    assert tot == 0, viol


def test_generator_contract(oracle):
    P, S = g.generate_pslg(3000, 200, L=1000.0, maxlen=100.0, dist=0, min_angle=10.0, seed=5)
    assert P.shape == (3000, 2) and (P >= 0).all() and (P <= 1000).all()
    assert len({tuple(p) for p in P.tolist()}) == 3000
    ln = np.hypot(*(P[S[:, 0]] - P[S[:, 1]]).T)
    assert (ln <= 100.0).all()
    # no two segments cross (exact)
    for i in range(len(S)):
        for j in range(i + 1, len(S)):
            a, b = S[i]
            c, d = S[j]
            if len({a, b, c, d}) < 4:
                continue
            o1 = oracle.orient2d(P[a], P[b], P[c]) * oracle.orient2d(P[a], P[b], P[d])
            o2 = oracle.orient2d(P[c], P[d], P[a]) * oracle.orient2d(P[c], P[d], P[b])
            assert not (o1 <= 0 and o2 <= 0)
    # deterministic for a seed
    P2, S2 = g.generate_pslg(3000, 200, L=1000.0, maxlen=100.0, dist=0, min_angle=10.0, seed=5)
    assert np.array_equal(P, P2) and np.array_equal(S, S2)


@pytest.mark.parametrize("dist", [3, 4, 5, 6])
def test_generator_families(dist):
    """The reference generator's disk, circle, grid and ellipse families
    (mesh.cpp:131-193), scaled to [0,L]^2: every point on its shape, all
    distinct, deterministic for a seed."""
    n, L = 4000, 1000.0
    P, S = g.generate_pslg(n, 200, L=L, maxlen=100.0, dist=dist, min_angle=10.0, seed=9)
    assert P.shape == (n, 2) and (P >= 0).all() and (P <= L).all()
    assert len({tuple(p) for p in P.tolist()}) == n
    # on the ellipse every point is a hull vertex, and a chord no longer than
    # maxlen meets the hull edges at its ends at a small angle: few chords
    # keep the minimum angle
    assert len(S) == 200 if dist != 6 else 0 < len(S) < 200
    u, v = P[:, 0] / L - 0.5, P[:, 1] / L - 0.5
    r = np.hypot(u, v)
    if dist == 3:
        assert r.max() <= 0.45 + 1e-12 and r.min() < 0.05
    elif dist == 4:
        assert r.min() >= 0.45 - 1e-12 and r.max() <= 0.49 + 1e-12
        # the radius is drawn uniformly, so the inner half of the ring holds
        # about half of the points (not the 49% an area-uniform draw gives)
        assert abs(np.mean(r < 0.47) - 0.5) < 0.03
    elif dist == 5:
        h = L / (10 * n)
        k = P / h
        assert np.abs(k - np.round(k)).max() < 1e-6
    elif dist == 6:
        assert np.abs((u / 0.25) ** 2 + (v / 0.5) ** 2 - 1).max() < 1e-12
    P2, S2 = g.generate_pslg(n, 200, L=L, maxlen=100.0, dist=dist, min_angle=10.0, seed=9)
    assert np.array_equal(P, P2) and np.array_equal(S, S2)
    with pytest.raises(RuntimeError):
        g.generate_pslg(n, 10, L=L, dist=7)


def test_cdt_segment_through_collinear_points(oracle):
    """A segment passing exactly through input points is split there."""
    pts = np.array([[0, 0], [10, 0], [5, 0], [2.5, 0], [5, 5], [5, -5], [8, 3], [1, -2]], float)
    T, N, SS = g.build_cdt(pts, np.array([[0, 1]], np.int32))
    segs = {tuple(sorted(s)) for s in SS.tolist()}
    assert {(0, 3), (2, 3), (1, 2)} <= segs
    tot, viol, _, _ = oracle.check_mesh(pts, T, N, SS, pts, SS, 0.0)
    assert tot == 0, viol
