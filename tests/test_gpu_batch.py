"""Batched refinement on the GPU (gcdt_batch_*): per-instance results equal
the CPU oracle's batched refinement bit for bit, and -- where no
cross-instance coupling can act (fixed cavity capacity, every bad triangle in
each pass) -- equal every instance refined alone, also across several waves
and with the graph-driven loop."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIXED_K = dict(cav_cap0=256, cav_cap_max=256, cand_min=1 << 29)


def _instances(oracle, specs):
    out = []
    for n, ns, seed in specs:
        pts, segs = oracle.random_pslg(n, ns, L=1000.0, maxlen=250.0, seed=seed)
        out.append(oracle.triangle_cdt(pts, segs))
    return out


SPECS = [(800, 40, 11), (2500, 120, 12), (300, 20, 13), (1800, 90, 14), (1200, 80, 15),
         (4000, 200, 16)]
THETA = [20.7, 25.0, 30.0, 28.0, 22.5, 26.0]
OFF = [1, 1, 0, 1, 0, 1]
MODE = [1, 0, 1, 1, 0, 1]


def _same(r, o):
    P, T, N, S, pf = o
    assert np.array_equal(r.points.view(np.uint64), P.view(np.uint64))
    assert np.array_equal(r.triangles, T)
    assert np.array_equal(r.neighbors, N)
    assert np.array_equal(r.segments, S)


@pytest.mark.parametrize("graph", [0, 1])
def test_batch_equals_oracle_batch(gpu, oracle, graph):
    """Shared candidate cap and cavity-capacity growth act across instances:
    the union refinement is compared with the oracle's."""
    print(f"batch vs oracle graph={graph}", flush=True)
    ins = _instances(oracle, SPECS)
    kw = dict(cand_min=256, cand_factor=2)
    b = gpu.Batch(ins, theta=THETA, insert=OFF, mode=MODE, graph=graph, **kw)
    try:
        st = b.run()
        res = b.results()
    finally:
        b.close()
    ores, ost = oracle.oracle_refine_batch(ins, THETA, OFF, MODE, **kw)
    assert st["status"] == ost["status"] == 0
    assert st["inserted"] == ost["inserted"] and st["seg_splits"] == ost["seg_splits"]
    for i in range(len(ins)):
        _same(res[i], ores[i])
        P, T, N, S = ins[i]
        tot, viol, mina, ex = oracle.check_mesh(res[i].points, res[i].triangles, res[i].neighbors,
                                                res[i].segments, P, S, THETA[i])
        assert tot == 0, (i, viol)


@pytest.mark.parametrize("wave_triangles,graph", [(0, 0), (6000, 0), (6000, 1)])
def test_batch_instances_equal_solo(gpu, oracle, wave_triangles, graph):
    print(f"batch vs solo waves={wave_triangles} graph={graph}", flush=True)
    ins = _instances(oracle, SPECS)
    b = gpu.Batch(ins, theta=THETA, insert=OFF, mode=MODE, wave_triangles=wave_triangles,
                  graph=graph, **FIXED_K)
    try:
        if wave_triangles:
            assert b.waves() > 1
        b.run()
        res = b.results()
    finally:
        b.close()
    for i, (P, T, N, S) in enumerate(ins):
        r = gpu.refine(P, T, N, S, theta=THETA[i], insert=OFF[i], mode=MODE[i], **FIXED_K)
        _same(res[i], (r.points, r.triangles, r.neighbors, r.segments, r.pstatus))
        assert np.array_equal(res[i].pstatus, r.pstatus)
        assert np.array_equal(res[i].tstatus, r.tstatus)


def test_batch_rerun_is_repeatable(gpu, oracle):
    ins = _instances(oracle, SPECS[:3])
    b = gpu.Batch(ins, theta=THETA[:3], insert=OFF[:3], mode=MODE[:3], graph=1)
    try:
        b.run()
        a = b.results()
        b.run()
        c = b.results()
    finally:
        b.close()
    for x, y in zip(a, c):
        assert np.array_equal(x.points.view(np.uint64), y.points.view(np.uint64))
        assert np.array_equal(x.triangles, y.triangles)
