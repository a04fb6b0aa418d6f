"""The headline workload at full size: 10^6 uniform points in [0, 1e6]^2 with
50,000 non-crossing segments no longer than 2% of the width (seed 2), refined
to about 10^7 triangles.  The CPU restatement is too slow to replay it, so
the meshes are held to size-independent properties: no violation in the GPU
checker or in the CPU checker (exact Delaunay across non-subsegment edges,
symmetry, orientation, boundary, subsegment conformity, input points
unmoved, every angle at least theta), run-to-run identical output, and a
Steiner count of the chosen form at most 5% above the plain path's."""
import os

import numpy as np
import pytest

from gpu_cdt_refine_amd import violations as gpu_violations

pytestmark = pytest.mark.gpu

CHOSEN = dict(pred_mode=1, claim_iters=3, shared_ring=1, cand_factor=3, priority=2,
              learned_shift=1, graph=0)
PLAIN = dict(pred_mode=0, claim_iters=1, priority=1, graph=0)


@pytest.fixture(scope="module")
def headline(gpu):
    P, S0 = gpu.generate_pslg(1_000_000, 50_000, L=1e6, maxlen=2e4, dist=0, min_angle=5.0,
                              seed=2)
    T, N, S = gpu.build_cdt(P, S0)
    return P, T, N, S


def _run(gpu, P, T, N, S, theta, insert, mode, opts):
    s = gpu.Session(P, T, N, S, theta=theta, insert=insert, mode=mode, **opts)
    try:
        if opts.get("priority") == 2:
            s.set_priority_model(np.load(os.path.join(os.path.dirname(gpu.__file__),
                                                      "priority_mlp.npy")))
        st = s.run()
        c = s.check()
        r = s.download()
    finally:
        s.close()
    return st, c, r


def _clean(oracle, st, c, r, P, S, theta):
    assert st["status"] == 0
    assert gpu_violations(c) == 0, c
    assert c["min_angle"] >= theta - 1e-6
    tot, viol, mina, _ = oracle.check_mesh(r.points, r.triangles, r.neighbors, r.segments,
                                           P, S, theta)
    assert tot == 0, viol
    assert mina >= theta - 1e-6
    assert np.array_equal(r.points[:len(P)].view(np.uint64), P.view(np.uint64))


@pytest.mark.parametrize("theta,insert,mode", [(25.0, 1, 1), (30.0, 0, 0)],
                         ids=["25-offcenter-ruppert", "30-circumcenter-chew"])
def test_headline_workload(gpu, oracle, headline, theta, insert, mode):
    P, T, N, S = headline
    print(f"1M workload theta={theta} insert={insert} mode={mode}", flush=True)
    stp, cp, rp = _run(gpu, P, T, N, S, theta, insert, mode, PLAIN)
    _clean(oracle, stp, cp, rp, P, S, theta)
    del rp
    st1, c1, r1 = _run(gpu, P, T, N, S, theta, insert, mode, CHOSEN)
    _clean(oracle, st1, c1, r1, P, S, theta)
    st2, _, r2 = _run(gpu, P, T, N, S, theta, insert, mode, CHOSEN)
    assert np.array_equal(r1.points.view(np.uint64), r2.points.view(np.uint64))
    assert np.array_equal(r1.triangles, r2.triangles)
    assert np.array_equal(r1.neighbors, r2.neighbors)
    assert np.array_equal(r1.segments, r2.segments)
    steiner_plain = stp["n_points"] - len(P)
    delta = (st1["n_points"] - stp["n_points"]) / steiner_plain
    print(f"  plain {stp['n_points']} points, {stp['rounds']} rounds; chosen "
          f"{st1['n_points']} points, {st1['rounds']} rounds ({100 * delta:+.1f}%)", flush=True)
    assert delta < 0.05
    assert st1["rounds"] < stp["rounds"]


@pytest.mark.parametrize("theta,insert,mode", [(25.0, 0, 1), (25.0, 1, 0), (25.0, 0, 0),
                                               (30.0, 1, 1), (30.0, 0, 1), (30.0, 1, 0)],
                         ids=["25-circumcenter-ruppert", "25-offcenter-chew",
                              "25-circumcenter-chew", "30-offcenter-ruppert",
                              "30-circumcenter-ruppert", "30-offcenter-chew"])
def test_headline_grid_cells(gpu, headline, theta, insert, mode):
    """The other six cells of the theta x mode x insert grid on the 10^6-point
    input: plain and chosen form converge, the GPU checker (exact Delaunay,
    conformity, coverage, every angle >= theta outside the MPW exemption,
    input points unmoved) finds nothing, the chosen form is reproducible and
    its Steiner count stays within 5% of the plain path's."""
    P, T, N, S = headline
    print(f"1M grid cell theta={theta} insert={insert} mode={mode}", flush=True)
    stp, cp, _ = _run(gpu, P, T, N, S, theta, insert, mode, PLAIN)
    st1, c1, r1 = _run(gpu, P, T, N, S, theta, insert, mode, CHOSEN)
    st2, _, r2 = _run(gpu, P, T, N, S, theta, insert, mode, CHOSEN)
    for st, c in ((stp, cp), (st1, c1)):
        assert st["status"] == 0, st
        assert gpu_violations(c) == 0, c
        assert c["min_angle"] >= theta - 1e-6 or c["exempt"] > 0, c
    assert np.array_equal(r1.points.view(np.uint64), r2.points.view(np.uint64))
    assert np.array_equal(r1.triangles, r2.triangles)
    delta = (st1["n_points"] - stp["n_points"]) / (stp["n_points"] - len(P))
    print(f"  plain {stp['n_points']} points; chosen {st1['n_points']} ({100 * delta:+.1f}%)",
          flush=True)
    assert delta < 0.05


def test_learned_config_full_size(gpu, oracle):
    """The learned-priority configuration at its full size: 2,000,000
    Gaussian-clustered points (sigma 5% of the domain), 100,000 segments,
    theta = 28, off-center, Ruppert; the chosen form with the trained network
    converges, passes the GPU and CPU checkers, is identical run to run, and
    its Steiner count is within 5% of the plain path's."""
    print("learned config: 2M clustered, 100k segments, theta 28", flush=True)
    P, S0 = gpu.generate_pslg(2_000_000, 100_000, L=1e6, maxlen=2e4, dist=1, min_angle=5.0,
                              seed=1002)
    T, N, S = gpu.build_cdt(P, S0)
    stp, cp, _ = _run(gpu, P, T, N, S, 28.0, 1, 1, PLAIN)
    assert stp["status"] == 0 and gpu_violations(cp) == 0, (stp, cp)
    st1, c1, r1 = _run(gpu, P, T, N, S, 28.0, 1, 1, CHOSEN)
    _clean(oracle, st1, c1, r1, P, S, 28.0)
    st2, _, r2 = _run(gpu, P, T, N, S, 28.0, 1, 1, CHOSEN)
    assert np.array_equal(r1.points.view(np.uint64), r2.points.view(np.uint64))
    assert np.array_equal(r1.triangles, r2.triangles)
    assert np.array_equal(r1.neighbors, r2.neighbors)
    delta = (st1["n_points"] - stp["n_points"]) / (stp["n_points"] - len(P))
    print(f"  plain {stp['n_points']} points, {stp['rounds']} rounds; chosen "
          f"{st1['n_points']} points, {st1['rounds']} rounds ({100 * delta:+.1f}%)", flush=True)
    assert abs(delta) < 0.05
    assert st1["rounds"] < stp["rounds"]
