"""Mixed rounds (options.mixed = 1): one pass per round in which the
encroached subsegments (first in candidate order) and the selected bad
triangles compete for their regions.  The GPU mesh equals the CPU oracle's
mixed form bit for bit, across the paths where the two kinds of candidate
meet: claim iterations, both priority rules, Chew / circumcenter insertion,
cavity-capacity growth, a workspace that cannot hold every encroached
subsegment, and mesh growth beyond the initial capacity."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ORACLE_KEYS = ["status", "rounds", "seg_passes", "bad_passes", "inserted", "seg_splits",
               "rej_encroach", "rej_blocked", "rej_overflow", "lost", "giveups",
               "final_cav_cap", "cand_evals"]


def _input(oracle, n, ns, seed, maxlen=200.0):
    pts, segs = oracle.random_pslg(n, ns, L=1000.0, maxlen=maxlen, seed=seed)
    return oracle.triangle_cdt(pts, segs)


MIXED_CASES = [
    # (name, input kwargs, refine kwargs)
    ("plain", dict(n=4000, ns=200, seed=71), dict(theta=25.0)),
    ("claim3-fp32", dict(n=4000, ns=200, seed=72), dict(theta=28.0, claim_iters=3, pred_mode=1)),
    ("index-priority", dict(n=4000, ns=200, seed=73),
     dict(theta=25.0, priority=0, cand_min=64, cand_factor=2)),
    ("chew-circumcenter", dict(n=4000, ns=200, seed=74), dict(theta=30.0, insert=0, mode=0)),
    ("segment-cap", dict(n=3000, ns=400, seed=75, maxlen=400.0),
     dict(theta=30.0, cand_budget=(6 * 32 + 16) * 150)),
    ("mesh-growth", dict(n=300, ns=60, seed=76, maxlen=1000.0), dict(theta=30.0)),
    ("large", dict(n=20000, ns=1000, seed=77), dict(theta=25.0, claim_iters=2)),
]


def _oracle_kw(kw):
    okw = {k: v for k, v in kw.items() if k not in ("pred_mode", "insert")}
    if "insert" in kw:
        okw["offcenter"] = kw["insert"]
    return okw


@pytest.mark.parametrize("name,inp,kw", MIXED_CASES, ids=[c[0] for c in MIXED_CASES])
def test_mixed_equals_oracle(gpu, oracle, name, inp, kw):
    print(f"mixed case {name}", flush=True)
    P, T, N, S = _input(oracle, **inp)
    r = gpu.refine(P, T, N, S, mixed=1, **kw)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, mixed=1, **_oracle_kw(kw))
    for k in ORACLE_KEYS:
        assert r.stats[k] == st[k], k
    assert r.stats["status"] == 0
    assert np.array_equal(r.points.view(np.uint64), Po.view(np.uint64))
    assert np.array_equal(r.triangles, To)
    assert np.array_equal(r.neighbors, No)
    assert np.array_equal(r.segments, So)
    theta = kw["theta"]
    tot, viol, mina, ex = oracle.check_mesh(r.points, r.triangles, r.neighbors, r.segments,
                                            P, S, theta)
    assert tot == 0, viol
    if name == "segment-cap":
        # some round held more encroached subsegments than one pass takes
        assert r.stats["seg_passes"] > 0
    if name == "mesh-growth":
        assert len(r.triangles) > 6 * len(T)


def test_mixed_cavity_capacity_growth(gpu, oracle):
    """K starts at 4 and doubles while subsegments and triangles share passes."""
    from test_gpu_parity import circle_pslg
    P, T, N, S = circle_pslg(oracle)
    kw = dict(cav_cap0=4, cav_cap_max=256, theta=25.0)
    r = gpu.refine(P, T, N, S, mixed=1, **kw)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, mixed=1, **kw)
    assert r.stats["final_cav_cap"] >= 64 and r.stats["rej_overflow"] > 0
    for k in ORACLE_KEYS:
        assert r.stats[k] == st[k], k
    assert np.array_equal(r.points.view(np.uint64), Po.view(np.uint64))
    assert np.array_equal(r.triangles, To)
    assert np.array_equal(r.neighbors, No)


def test_mixed_learned_priority_deterministic(gpu, oracle):
    """Mixed rounds with the learned claim priority: run-to-run identical and
    every invariant holds (no CPU restatement of the network's ordering)."""
    import os
    model = os.path.join(os.path.dirname(gpu.__file__), "priority_mlp.npy")
    P, T, N, S = _input(oracle, 8000, 400, 78)
    meshes = []
    for _ in range(2):
        s = gpu.Session(P, T, N, S, theta=25.0, mixed=1, priority=2, claim_iters=3, pred_mode=1)
        try:
            s.set_priority_model(np.load(model))
            st = s.run()
            meshes.append(s.download())
        finally:
            s.close()
    a, b = meshes
    assert st["status"] == 0
    assert np.array_equal(a.points.view(np.uint64), b.points.view(np.uint64))
    assert np.array_equal(a.triangles, b.triangles)
    tot, viol, mina, ex = oracle.check_mesh(a.points, a.triangles, a.neighbors, a.segments,
                                            P, S, 25.0)
    assert tot == 0, viol


def test_mixed_needs_host_loop(gpu, oracle):
    P, T, N, S = _input(oracle, 500, 20, 79)
    with pytest.raises(RuntimeError, match="mixed"):
        gpu.refine(P, T, N, S, mixed=1, graph=1)


MIX_BELOW_CASES = [
    ("mb50", dict(n=4000, ns=300, seed=91), dict(theta=25.0, mix_below=50)),
    ("mb500-shared-ring", dict(n=4000, ns=300, seed=92),
     dict(theta=28.0, mix_below=500, shared_ring=1, claim_iters=3, pred_mode=1)),
    ("mb200-index-priority", dict(n=4000, ns=300, seed=93),
     dict(theta=25.0, mix_below=200, priority=0, cand_min=64, cand_factor=2)),
    ("mb1000-large", dict(n=20000, ns=1000, seed=94),
     dict(theta=25.0, mix_below=1000, shared_ring=1, claim_iters=3)),
    # the chosen form's setting: only the last few encroached subsegments
    # of a segment phase join the bad pass
    ("mb4-chosen-claims", dict(n=12000, ns=600, seed=95),
     dict(theta=25.0, mix_below=4, shared_ring=1, claim_iters=3, seg_claim_iters=2,
          cand_factor=3, pred_mode=1)),
    ("mb4-chew", dict(n=10000, ns=500, seed=96),
     dict(theta=25.0, mode=0, mix_below=4, shared_ring=1, claim_iters=3, seg_claim_iters=2)),
]


@pytest.mark.parametrize("name,inp,kw", MIX_BELOW_CASES, ids=[c[0] for c in MIX_BELOW_CASES])
def test_mix_below_equals_oracle(gpu, oracle, name, inp, kw):
    """Segment phases that stop at mix_below encroached subsegments, which
    then join the round's bad pass."""
    print(f"mix_below case {name}", flush=True)
    P, T, N, S = _input(oracle, **inp)
    r = gpu.refine(P, T, N, S, **kw)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, **_oracle_kw(kw))
    for k in ORACLE_KEYS:
        assert r.stats[k] == st[k], k
    assert r.stats["status"] == 0
    assert np.array_equal(r.points.view(np.uint64), Po.view(np.uint64))
    assert np.array_equal(r.triangles, To)
    assert np.array_equal(r.neighbors, No)
    assert np.array_equal(r.segments, So)
    tot, viol, mina, ex = oracle.check_mesh(r.points, r.triangles, r.neighbors, r.segments,
                                            P, S, kw["theta"])
    assert tot == 0, viol
