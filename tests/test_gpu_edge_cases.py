"""Edge cases of the drop-in's input contract (GPU_Refine_Quality,
refine.cu:6500-7139; input rules of README.md:11-51 in the reference): the
smallest meshes, inputs that need no refinement, and inputs that the library
rejects with an error before any kernel runs, after which it keeps working."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hull(k):
    return np.array([[i, (i + 1) % k] for i in range(k)], np.int32)


def _same_as_oracle(gpu, oracle, P, T, N, S, theta, insert, mode):
    r = gpu.refine(P, T, N, S, theta=theta, insert=insert, mode=mode)
    Po, To, No, So, _, st = oracle.oracle_refine(P, T, N, S, theta=theta, offcenter=insert,
                                                 mode=mode)
    assert r.stats["status"] == 0 and st["status"] == 0
    assert r.stats["rounds"] == st["rounds"]
    assert np.array_equal(r.points.view(np.uint64), Po.view(np.uint64))
    assert np.array_equal(r.triangles, To)
    assert np.array_equal(r.neighbors, No)
    assert np.array_equal(r.segments, So)
    tot, viol, mina, _ = oracle.check_mesh(r.points, r.triangles, r.neighbors, r.segments,
                                           P, S, theta)
    assert tot == 0, viol
    assert mina >= theta - 1e-6
    return r


def test_one_good_triangle_is_unchanged(gpu, oracle):
    P = np.array([[0.0, 0.0], [2.0, 0.0], [1.0, 3.0 ** 0.5]])
    T, N, S = gpu.build_cdt(P, _hull(3))
    r = _same_as_oracle(gpu, oracle, P, T, N, S, 30.0, gpu.Offcenter, gpu.RUPPERT)
    assert len(r.points) == 3 and len(r.triangles) == 1
    assert np.array_equal(r.points, P)
    assert r.stats["inserted"] == 0 and r.stats["seg_splits"] == 0


@pytest.mark.parametrize("insert,mode", [(1, 1), (0, 1), (1, 0), (0, 0)])
def test_thin_rectangle(gpu, oracle, insert, mode):
    """Two triangles of aspect 20: every Steiner point lands near the
    boundary, so most candidates encroach a subsegment."""
    P = np.array([[0.0, 0.0], [20.0, 0.0], [20.0, 1.0], [0.0, 1.0]])
    T, N, S = gpu.build_cdt(P, _hull(4))
    r = _same_as_oracle(gpu, oracle, P, T, N, S, 25.0, insert, mode)
    assert len(r.points) > 4


def test_nothing_to_refine_leaves_mesh(gpu, oracle):
    """A 30 x 30 unit lattice with its hull as the only segments: every
    triangle has angles 45-45-90 and no vertex lies inside a hull edge's
    diametral circle, so at theta = 30 degrees the output is the input."""
    xs, ys = np.meshgrid(np.arange(30.0), np.arange(30.0))
    P = np.stack([xs.ravel(), ys.ravel()], 1)
    T, N, S = gpu.build_cdt(P, np.zeros((0, 2), np.int32))
    r = _same_as_oracle(gpu, oracle, P, T, N, S, 30.0, gpu.Offcenter, gpu.RUPPERT)
    assert r.stats["inserted"] == 0 and r.stats["seg_splits"] == 0
    assert np.array_equal(r.points.view(np.uint64), P.view(np.uint64))
    assert len(r.triangles) == len(T) and len(r.segments) == len(S)
    key = lambda A: sorted(map(tuple, np.sort(np.asarray(A), 1).tolist()))
    assert key(r.triangles) == key(T)
    assert key(r.segments) == key(S)


def test_invalid_inputs_raise_and_library_recovers(gpu, oracle):
    pts, segs = oracle.random_pslg(300, 15, L=1000.0, maxlen=200.0, seed=62)
    P, T, N, S = oracle.triangle_cdt(pts, segs)
    T, N, S = (np.asarray(a, np.int32) for a in (T, N, S))
    cases = {
        "no triangles": (P, T[:0], N[:0], S),
        "too few points": (P[:2], T, N, S),
        "vertex index": (P, np.where(np.arange(T.size).reshape(T.shape) == 0, len(P), T), N, S),
        "neighbour index": (P, T, np.where(np.arange(N.size).reshape(N.shape) == 4, len(T), N), S),
        "segment index": (P, T, N, np.where(np.arange(S.size).reshape(S.shape) == 1, -1, S)),
    }
    for name, (p, t, n, s) in cases.items():
        with pytest.raises(RuntimeError):
            gpu.refine(p, t, n, s, theta=25.0)
    for theta in (-1.0, 60.0, float("nan")):
        with pytest.raises(RuntimeError, match="theta"):
            gpu.refine(P, T, N, S, theta=theta)
    _same_as_oracle(gpu, oracle, P, T, N, S, 25.0, gpu.Offcenter, gpu.RUPPERT)
