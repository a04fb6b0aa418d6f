"""Vertex deletion on the GPU (options.deletion; Triangle's rule before an
encroached subsegment splits, triangle.cpp:13319-13357, and the reference's,
refine.cu:604-677): the refined mesh equals the oracle's bit for bit --
points (deleted vertices dropped, the rest renumbered in order), triangles,
neighbours, subsegments, the deletion and flip counts -- in Chew mode, in
every mode (deletion = 2), in the graph-driven and mixed loops and under the
chosen claim rule; every mesh passes the exact invariant checks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _input(oracle, n, seed):
    pts, segs = oracle.random_pslg(n, n // 20, L=1000.0, maxlen=200.0, seed=seed)
    return oracle.triangle_cdt(pts, segs)


def _same(r, Po, To, No, So):
    assert np.array_equal(r.points.view(np.uint64), Po.view(np.uint64))
    assert np.array_equal(r.triangles, To)
    assert np.array_equal(r.neighbors, No)
    assert np.array_equal(r.segments, So)


CASES = [
    # (n, seed, theta, mode, deletion, extra options)
    (3000, 61, 25.0, 0, 1, {}),
    (3000, 62, 30.0, 0, 1, {}),
    (20000, 63, 28.0, 0, 1, {}),
    (3000, 64, 25.0, 1, 2, {}),
    (3000, 65, 25.0, 0, 0, {}),
    (6000, 66, 27.0, 0, 1, dict(claim_iters=3, shared_ring=1, cand_factor=3, cand_min=512)),
    (6000, 67, 26.0, 0, 1, dict(mixed=1)),
    (6000, 68, 26.0, 0, 1, dict(mix_below=64)),
]


@pytest.mark.parametrize("n,seed,theta,mode,deletion,kw", CASES)
def test_deletion_equals_oracle(gpu, oracle, n, seed, theta, mode, deletion, kw):
    print(f"deletion n={n} seed={seed} theta={theta} mode={mode} deletion={deletion} {kw}",
          flush=True)
    P, T, N, S = _input(oracle, n, seed)
    r = gpu.refine(P, T, N, S, theta=theta, mode=mode, deletion=deletion, **kw)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, theta=theta, mode=mode,
                                                  deletion=deletion, **kw)
    assert r.stats["status"] == st["status"] == 0
    assert r.stats["deleted"] == st["deleted"]
    assert r.stats["flips"] == st["flips"]
    # Chew mode deletes; Ruppert mode rejects every Steiner point inside a
    # diametral circle, so even deletion = 2 rarely finds a vertex there
    if deletion == 0:
        assert st["deleted"] == 0
    elif mode == 0:
        assert st["deleted"] > 0
    assert r.stats["rounds"] == st["rounds"]
    _same(r, Po, To, No, So)
    tot, viol, mina, ex = oracle.check_mesh(r.points, r.triangles, r.neighbors, r.segments,
                                            P, S, theta)
    assert tot == 0, viol
    assert mina >= theta - 1e-6


@pytest.mark.parametrize("mode,deletion", [(0, 1), (1, 2)])
def test_deletion_graph_loop_equals_host_loop(gpu, oracle, mode, deletion):
    """The graph-driven loop runs the deletion step inside the captured
    segment passes: same mesh as the host loop and the oracle."""
    print(f"deletion graph mode={mode} deletion={deletion}", flush=True)
    P, T, N, S = _input(oracle, 8000, 69)
    g = gpu.refine(P, T, N, S, theta=27.0, mode=mode, deletion=deletion, graph=1)
    h = gpu.refine(P, T, N, S, theta=27.0, mode=mode, deletion=deletion)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, theta=27.0, mode=mode,
                                                  deletion=deletion)
    assert g.stats["deleted"] == h.stats["deleted"] == st["deleted"]
    assert st["deleted"] > 0 or mode == 1
    assert g.stats["flips"] == st["flips"]
    _same(g, Po, To, No, So)
    _same(h, Po, To, No, So)


def test_deletion_batch_mixed_modes(gpu, oracle):
    """A batch with Chew- and Ruppert-mode instances: deletion acts in the
    Chew-mode ones only, and every instance equals the oracle's batch."""
    ins = [_input(oracle, n, seed) for n, seed in ((1500, 71), (2500, 72), (1200, 73), (3000, 74))]
    theta = [25.0, 28.0, 22.0, 30.0]
    off = [1, 1, 0, 1]
    mode = [0, 1, 0, 0]
    b = gpu.Batch(ins, theta=theta, insert=off, mode=mode, graph=0)
    try:
        st = b.run()
        res = b.results()
    finally:
        b.close()
    ores, ost = oracle.oracle_refine_batch(ins, theta, off, mode)
    assert st["status"] == ost["status"] == 0
    assert st["deleted"] == ost["deleted"] > 0
    for i in range(len(ins)):
        Po, To, No, So, pf = ores[i]
        _same(res[i], Po, To, No, So)
