"""GPU refinement vs the CPU oracle: bit-identical meshes, exact invariants,
run-to-run determinism.  Inputs are seeded PSLGs triangulated by the
reference's Triangle ("pzQnc", mesh.cpp:397)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # (n points, n segments, seed, theta, offcenter, mode)
    (2000, 100, 1, 25.0, 1, 1),
    (2000, 100, 1, 25.0, 1, 0),
    (2000, 100, 2, 30.0, 1, 1),
    (2000, 100, 3, 25.0, 0, 1),
    (2000, 100, 4, 30.0, 0, 0),
    (2000, 100, 5, 20.7, 1, 1),
]


def _input(oracle, n, ns, seed):
    pts, segs = oracle.random_pslg(n, ns, L=1000.0, maxlen=200.0, seed=seed)
    P, T, N, S = oracle.triangle_cdt(pts, segs)
    return P, T, N, S


@pytest.mark.parametrize("n,ns,seed,theta,off,mode", CASES)
def test_gpu_equals_oracle(gpu, oracle, n, ns, seed, theta, off, mode):
    print(f"case n={n} seed={seed} theta={theta} off={off} mode={mode}", flush=True)
    P, T, N, S = _input(oracle, n, ns, seed)
    r = gpu.refine(P, T, N, S, theta=theta, insert=off, mode=mode)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, theta=theta, offcenter=off, mode=mode)
    assert r.stats["status"] == 0 and st["status"] == 0
    assert r.stats["rounds"] == st["rounds"]
    assert np.array_equal(r.points.view(np.uint64), Po.view(np.uint64))
    assert np.array_equal(r.triangles, To)
    assert np.array_equal(r.neighbors, No)
    assert np.array_equal(r.segments, So)
    tot, viol, mina, ex = oracle.check_mesh(r.points, r.triangles, r.neighbors, r.segments,
                                            P, S, theta)
    assert tot == 0, viol
    assert mina >= theta - 1e-6


def test_gpu_deterministic(gpu, oracle):
    P, T, N, S = _input(oracle, 5000, 250, 11)
    a = gpu.refine(P, T, N, S, theta=28.0)
    b = gpu.refine(P, T, N, S, theta=28.0)
    assert np.array_equal(a.points.view(np.uint64), b.points.view(np.uint64))
    assert np.array_equal(a.triangles, b.triangles)
    assert np.array_equal(a.neighbors, b.neighbors)
    assert np.array_equal(a.segments, b.segments)
