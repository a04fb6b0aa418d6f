"""Graph-driven round loop (options.graph = 1): the same mesh, bit for bit,
as the host-driven loop and the CPU oracle, across the paths that make the
device control block decide something -- cavity-capacity growth (graph
re-capture), mesh-capacity growth (suspend + re-capture), priority-ordered
candidate selection, claim iterations, the FP32 filter stage and rounds that
need more segment passes than one graph round holds."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STAT_KEYS = ["status", "rounds", "seg_passes", "bad_passes", "inserted", "seg_splits",
             "rej_encroach", "rej_blocked", "rej_overflow", "lost", "giveups",
             "final_cav_cap", "cand_evals", "n_points", "n_triangles", "n_subsegs"]


def _input(oracle, n, ns, seed, maxlen=200.0):
    pts, segs = oracle.random_pslg(n, ns, L=1000.0, maxlen=maxlen, seed=seed)
    return oracle.triangle_cdt(pts, segs)


def _same(a, b):
    assert np.array_equal(a.points.view(np.uint64), b.points.view(np.uint64))
    assert np.array_equal(a.triangles, b.triangles)
    assert np.array_equal(a.neighbors, b.neighbors)
    assert np.array_equal(a.segments, b.segments)
    assert np.array_equal(a.pstatus, b.pstatus)
    assert np.array_equal(a.tstatus, b.tstatus)


GRAPH_CASES = [
    # (name, input kwargs, refine kwargs)
    ("plain", dict(n=4000, ns=200, seed=51), dict(theta=25.0)),
    ("chew-circumcenter", dict(n=4000, ns=200, seed=52), dict(theta=30.0, insert=0, mode=0)),
    ("claim3-fp32", dict(n=4000, ns=200, seed=53), dict(theta=28.0, claim_iters=3, pred_mode=1)),
    ("index-priority", dict(n=4000, ns=200, seed=54),
     dict(theta=25.0, priority=0, cand_min=64, cand_factor=2)),
    ("selection", dict(n=4000, ns=200, seed=55),
     dict(theta=25.0, cand_min=64, cand_factor=2, cand_budget=(6 * 32 + 16) * 500)),
    ("one-seg-pass", dict(n=3000, ns=300, seed=56),
     dict(theta=30.0, graph_seg_passes=1, graph_rounds=1)),
    ("long-graph", dict(n=3000, ns=300, seed=57),
     dict(theta=30.0, graph_seg_passes=3, graph_rounds=8)),
    ("mesh-growth", dict(n=300, ns=60, seed=58, maxlen=1000.0), dict(theta=30.0)),
    ("small-grids", dict(n=4000, ns=200, seed=59),
     dict(theta=25.0, grid_cavity=7, grid_claim=3, grid_apply=5)),
]


@pytest.mark.parametrize("name,inp,kw", GRAPH_CASES, ids=[c[0] for c in GRAPH_CASES])
def test_graph_equals_host_loop(gpu, oracle, name, inp, kw):
    print(f"graph case {name}", flush=True)
    P, T, N, S = _input(oracle, **inp)
    h = gpu.refine(P, T, N, S, graph=0, **kw)
    g = gpu.refine(P, T, N, S, graph=1, **kw)
    for k in STAT_KEYS:
        assert g.stats[k] == h.stats[k], k
    _same(g, h)
    okw = {k: v for k, v in kw.items()
           if k not in ("graph_seg_passes", "graph_rounds", "grid_cavity", "grid_claim",
                        "grid_apply", "pred_mode", "insert")}
    if "insert" in kw:
        okw["offcenter"] = kw["insert"]
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, **okw)
    assert np.array_equal(g.points.view(np.uint64), Po.view(np.uint64))
    assert np.array_equal(g.triangles, To)
    assert np.array_equal(g.neighbors, No)
    assert np.array_equal(g.segments, So)
    if name == "mesh-growth":
        assert len(g.triangles) > 6 * len(T)  # beyond the initial capacity


def test_graph_cavity_capacity_growth(gpu, oracle):
    """K starts at 4 and doubles: every doubling re-captures the graph."""
    from test_gpu_parity import circle_pslg
    P, T, N, S = circle_pslg(oracle)
    kw = dict(cav_cap0=4, cav_cap_max=256, theta=25.0)
    h = gpu.refine(P, T, N, S, graph=0, **kw)
    g = gpu.refine(P, T, N, S, graph=1, **kw)
    assert g.stats["final_cav_cap"] >= 64 and g.stats["rej_overflow"] > 0
    for k in STAT_KEYS:
        assert g.stats[k] == h.stats[k], k
    _same(g, h)


def test_graph_session_repeatable(gpu, oracle):
    """A session replays its graph loop from the resident input: two runs
    give the same mesh."""
    P, T, N, S = _input(oracle, 5000, 250, 60)
    s = gpu.Session(P, T, N, S, theta=28.0, graph=1, claim_iters=2)
    try:
        s.run()
        a = s.download()
        s.run()
        b = s.download()
    finally:
        s.close()
    _same(a, b)


def test_autotune_keeps_the_mesh(gpu, oracle):
    """The autotuner only changes launch shapes and the graph's structure:
    the refined mesh after tuning is the mesh before."""
    P, T, N, S = _input(oracle, 6000, 300, 61)
    s = gpu.Session(P, T, N, S, theta=26.0, graph=1, claim_iters=2)
    try:
        s.run()
        a = s.download()
        best, trials = s.autotune(reps=1)
        assert len(trials) >= 10
        assert all(t["ms"] > 0 for t in trials)
        s.run()
        b = s.download()
    finally:
        s.close()
    _same(a, b)


@pytest.mark.parametrize("graph", [0, 1])
def test_mesh_growth_in_place_equals_copy(gpu, oracle, graph):
    """A small input whose refined mesh outgrows the initial capacity (6x
    the input) several times: with options.vmm the arrays grow in place
    (reserved ranges, more memory mapped; the graph loop keeps its graph),
    without it they are reallocated and copied -- the same mesh either way,
    equal to the oracle's."""
    print(f"growth graph={graph}", flush=True)
    P, T, N, S = _input(oracle, 300, 60, 58, maxlen=1000.0)
    v = gpu.refine(P, T, N, S, theta=30.0, graph=graph, vmm=1)
    c = gpu.refine(P, T, N, S, theta=30.0, graph=graph, vmm=0)
    assert v.stats["growths"] > 1 and c.stats["growths"] == v.stats["growths"]
    assert v.stats["grow_in_place"] == 1 and c.stats["grow_in_place"] == 0
    _same(v, c)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, theta=30.0)
    assert np.array_equal(v.points.view(np.uint64), Po.view(np.uint64))
    assert np.array_equal(v.triangles, To)
    assert np.array_equal(v.neighbors, No)


def test_graph_equals_host_loop_learned_priority(gpu, oracle):
    """The learned claim priority: the host loop folds the network's scores
    into the priorities inside its first claim iteration, the graph loop in a
    kernel of its own; both give the same mesh."""
    import os
    model = np.load(os.path.join(os.path.dirname(gpu.__file__), "priority_mlp.npy"))
    P, T, N, S = _input(oracle, n=6000, ns=300, seed=60)
    kw = dict(theta=25.0, priority=2, claim_iters=3, shared_ring=1, cand_factor=3,
              pred_mode=1, learned_shift=1)
    out = []
    for graph in (0, 1):
        s = gpu.Session(P, T, N, S, graph=graph, **kw)
        try:
            s.set_priority_model(model)
            st = s.run()
            out.append((st, s.download()))
        finally:
            s.close()
    (sh, h), (sg, g) = out
    assert sh["status"] == 0 and sh["bad_passes"] > 1
    for k in STAT_KEYS:
        assert sg[k] == sh[k], k
    _same(g, h)
