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


def _family(gpu, specs):
    """Instances of the batched family's kinds: uniform (dist 0) or
    Gaussian-clustered (dist 1, sigma 5% of the domain), 5% segments no
    longer than 5% of the domain, the domain scaled with the point count."""
    out = []
    for n, dist, seed in specs:
        L = 1000.0 * (n / 10000) ** 0.5
        P, S = gpu.generate_pslg(n, max(1, n // 20), L=L, maxlen=0.05 * L, dist=dist,
                                 min_angle=5.0, seed=seed)
        T, N, SS = gpu.build_cdt(P, S)
        out.append((P, T, N, SS))
    return out


FAMILY = [(3000, 1, 41), (1200, 0, 42), (6000, 1, 43), (800, 1, 44), (4500, 0, 45),
          (2500, 1, 46)]
FTHETA = [21.3, 29.5, 24.0, 27.7, 20.9, 25.5]
FMODE = [1, 0, 1, 0, 1, 1]


@pytest.mark.parametrize("graph", [0, 1])
def test_batch_family_equals_oracle_batch(gpu, oracle, graph):
    """Uniform and Gaussian-clustered instances of the batched family with
    per-instance theta and mode, under the chosen form's claim rule (three
    claim iterations, shared rings): bit-identical to the oracle's batched
    refinement, and every instance clean on the GPU checker (including
    segment coverage and every angle at least its own theta)."""
    print(f"batch family vs oracle graph={graph}", flush=True)
    ins = _family(gpu, FAMILY)
    off = [1] * len(ins)
    kw = dict(cand_min=512, cand_factor=3, claim_iters=3, shared_ring=1)
    b = gpu.Batch(ins, theta=FTHETA, insert=off, mode=FMODE, graph=graph, **kw)
    try:
        st = b.run()
        res = b.results()
    finally:
        b.close()
    ores, ost = oracle.oracle_refine_batch(ins, FTHETA, off, FMODE, **kw)
    assert st["status"] == ost["status"] == 0
    assert st["inserted"] == ost["inserted"] and st["seg_splits"] == ost["seg_splits"]
    for i, (P, T, N, S) in enumerate(ins):
        _same(res[i], ores[i])
        c = gpu.check_mesh(res[i].points, res[i].triangles, res[i].neighbors, res[i].segments,
                           points_in=P, segments_in=S, theta=FTHETA[i])
        assert gpu.violations(c) == 0, (i, c)
        assert c["min_angle"] >= FTHETA[i] - 1e-9 or c["exempt"] > 0, (i, c)


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
