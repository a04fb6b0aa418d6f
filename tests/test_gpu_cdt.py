"""GPU construction of the input CDT (gcdt_build_cdt_device) against the
reference's Triangle ("pzQnc", mesh.cpp:397, compiled from its own source in
oracle/_ref): the same triangles and the same segments where no four points
are cocircular, a valid constrained Delaunay triangulation everywhere (exact
checker), identical output run to run, segments split where they pass
through input points, and the result refined by the GPU path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _canon(T):
    T = np.asarray(T).reshape(-1, 3)
    r = np.argmin(T, axis=1)
    i = np.arange(len(T))
    return set(map(tuple, np.stack([T[i, r], T[i, (r + 1) % 3], T[i, (r + 2) % 3]], 1).tolist()))


def _segset(S):
    return set(tuple(sorted(x)) for x in np.asarray(S).reshape(-1, 2).tolist())


def _valid(oracle, P, T, N, S):
    tot, viol, mina, ex = oracle.check_mesh(P, T, N, S, P, S, 1e-6)
    assert tot == 0, viol


@pytest.mark.parametrize("n,ns,dist,seed", [(2000, 100, 0, 1), (20000, 1000, 1, 2),
                                            (60000, 3000, 0, 3), (20000, 1000, 3, 4),
                                            (20000, 1000, 4, 5), (20000, 1000, 5, 6),
                                            (20000, 1000, 6, 7)])
def test_gpu_cdt_equals_triangle(gpu, oracle, n, ns, dist, seed):
    print(f"gpu cdt n={n} dist={dist}", flush=True)
    L = 1000.0 * (n / 10000) ** 0.5
    P, S0 = gpu.generate_pslg(n, ns, L=L, maxlen=0.05 * L, dist=dist, min_angle=5.0, seed=seed)
    st = {}
    T, N, S = gpu.build_cdt(P, S0, device=True, stats=st)
    Pt, Tt, Nt, St = oracle.triangle_cdt(P, S0)
    assert _canon(T) == _canon(Tt)
    assert _segset(S) == _segset(St)
    _valid(oracle, P, T, N, S)
    T2, N2, S2 = gpu.build_cdt(P, S0, device=True)
    assert np.array_equal(T, T2) and np.array_equal(N, N2) and np.array_equal(S, S2)
    # the host builder gives the same triangulation too
    Th, Nh, Sh = gpu.build_cdt(P, S0)
    assert _canon(T) == _canon(Th) and _segset(S) == _segset(Sh)


def test_gpu_cdt_precision_family(gpu, oracle):
    """Jittered lattice, exactly cocircular groups, segments along lattice
    lines: where four points are cocircular any Delaunay choice is valid, so
    the mesh is held to the exact checker and to Triangle's segments and
    triangle count."""
    P, S0 = gpu.generate_pslg(20000, 400, L=1000.0 * 0.02 ** 0.5, maxlen=20.0, dist=2,
                              min_angle=5.0, seed=2)
    T, N, S = gpu.build_cdt(P, S0, device=True)
    Pt, Tt, Nt, St = oracle.triangle_cdt(P, S0)
    assert len(T) == len(Tt)
    assert _segset(S) == _segset(St)
    _valid(oracle, P, T, N, S)


def test_gpu_cdt_segment_through_points(gpu, oracle):
    """A segment through three input points is split into four pieces; a
    grid of points (every quadrilateral cocircular) with segments along its
    lines and diagonals."""
    g = 12
    xs, ys = np.meshgrid(np.arange(g, dtype=np.float64), np.arange(g, dtype=np.float64))
    P = np.stack([xs.ravel(), ys.ravel()], 1)
    idx = lambda x, y: y * g + x
    S0 = np.array([[idx(0, 2), idx(8, 2)], [idx(1, 1), idx(7, 7)], [idx(10, 0), idx(10, 9)],
                   [idx(3, 9), idx(9, 3)]], np.int32)
    T, N, S = gpu.build_cdt(P, S0, device=True)
    Pt, Tt, Nt, St = oracle.triangle_cdt(P, S0)
    assert _segset(S) == _segset(St)
    assert len(T) == len(Tt)
    _valid(oracle, P, T, N, S)
    # the first input segment comes out as its 8 unit pieces, in order
    assert [tuple(x) for x in S[:8].tolist()] == [(idx(k, 2), idx(k + 1, 2)) for k in range(8)]


def test_gpu_cdt_feeds_the_refinement(gpu, oracle):
    P, S0 = gpu.generate_pslg(30000, 1500, L=1000.0 * 3 ** 0.5, maxlen=80.0, dist=1,
                              min_angle=5.0, seed=9)
    T, N, S = gpu.build_cdt(P, S0, device=True)
    r = gpu.refine(P, T, N, S, theta=25.0, claim_iters=3, shared_ring=1)
    assert r.stats["status"] == 0
    c = gpu.check_mesh(r.points, r.triangles, r.neighbors, r.segments, points_in=P,
                       segments_in=S, theta=25.0)
    assert gpu.violations(c) == 0, c
