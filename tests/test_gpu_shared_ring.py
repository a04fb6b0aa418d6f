"""Shared rings (options.shared_ring = 1): a pass's winners may have
overlapping rings, while no winner's cavity meets another winner's cavity or
ring.  The GPU mesh equals the CPU oracle's bit for bit under both round
loops, with claim iterations and winner stamps, both priority rules, mixed
rounds, cavity-capacity growth and mesh growth."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ORACLE_KEYS = ["status", "rounds", "seg_passes", "bad_passes", "inserted", "seg_splits",
               "rej_encroach", "rej_blocked", "rej_overflow", "lost", "giveups",
               "final_cav_cap", "cand_evals"]


def _input(oracle, n, ns, seed, maxlen=200.0):
    pts, segs = oracle.random_pslg(n, ns, L=1000.0, maxlen=maxlen, seed=seed)
    return oracle.triangle_cdt(pts, segs)


CASES = [
    # (name, input kwargs, refine kwargs)
    ("plain", dict(n=4000, ns=200, seed=81), dict(theta=25.0)),
    ("claim3-fp32", dict(n=4000, ns=200, seed=82), dict(theta=28.0, claim_iters=3, pred_mode=1)),
    ("claim3-graph", dict(n=4000, ns=200, seed=83), dict(theta=25.0, claim_iters=3, graph=1)),
    ("index-priority", dict(n=4000, ns=200, seed=84),
     dict(theta=25.0, priority=0, cand_min=64, cand_factor=2, claim_iters=2)),
    ("chew-circumcenter", dict(n=4000, ns=200, seed=85), dict(theta=30.0, insert=0, mode=0)),
    ("mixed", dict(n=4000, ns=300, seed=86), dict(theta=25.0, mixed=1, claim_iters=3)),
    ("mesh-growth", dict(n=300, ns=60, seed=87, maxlen=1000.0), dict(theta=30.0, claim_iters=2)),
    ("large", dict(n=20000, ns=1000, seed=88), dict(theta=25.0, claim_iters=3)),
]


def _oracle_kw(kw):
    okw = {k: v for k, v in kw.items() if k not in ("pred_mode", "insert", "graph")}
    if "insert" in kw:
        okw["offcenter"] = kw["insert"]
    return okw


@pytest.mark.parametrize("name,inp,kw", CASES, ids=[c[0] for c in CASES])
def test_shared_ring_equals_oracle(gpu, oracle, name, inp, kw):
    print(f"shared-ring case {name}", flush=True)
    P, T, N, S = _input(oracle, **inp)
    r = gpu.refine(P, T, N, S, shared_ring=1, **kw)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, shared_ring=1, **_oracle_kw(kw))
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


def test_shared_ring_cavity_capacity_growth(gpu, oracle):
    from test_gpu_parity import circle_pslg
    P, T, N, S = circle_pslg(oracle)
    kw = dict(cav_cap0=4, cav_cap_max=256, theta=25.0, claim_iters=2)
    r = gpu.refine(P, T, N, S, shared_ring=1, **kw)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, shared_ring=1, **kw)
    assert r.stats["rej_overflow"] > 0
    for k in ORACLE_KEYS:
        assert r.stats[k] == st[k], k
    assert np.array_equal(r.points.view(np.uint64), Po.view(np.uint64))
    assert np.array_equal(r.triangles, To)
    assert np.array_equal(r.neighbors, No)
