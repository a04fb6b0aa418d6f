"""Device predicates (FP64 filter + exact, and FP32 -> FP64 -> exact) against
exact rational arithmetic, and the device's FP64 division / square root
against the host's correctly rounded ones (the refined coordinates are
reproduced bit for bit by the CPU oracle only if both round identically)."""
import numpy as np
import pytest

from test_predicates_cpu import exact_incircle, exact_orient, near_degenerate_queries

pytestmark = pytest.mark.gpu


def test_device_predicates_exact(gpu):
    rng = np.random.default_rng(7)
    q = np.concatenate([rng.uniform(-100, 100, size=(500, 8)), near_degenerate_queries(rng, 1500)])
    exp_o = np.array([exact_orient(r[0:2], r[2:4], r[4:6]) for r in q])
    exp_i = np.array([exact_incircle(r[0:2], r[2:4], r[4:6], r[6:8]) for r in q])
    for mode in (0, 1):
        o, ic, stages = gpu.device_predicates(q, pred_mode=mode)
        assert np.array_equal(o, exp_o), mode
        assert np.array_equal(ic, exp_i), mode
        assert stages[2] > 0  # the exact stage was exercised
        if mode == 1:
            assert stages[0] > 0  # the FP32 stage decided some signs


def test_fp32_filter_gives_identical_mesh(gpu, oracle):
    pts, segs = oracle.random_pslg(4000, 200, L=1000.0, maxlen=200.0, seed=17)
    P, T, N, S = oracle.triangle_cdt(pts, segs)
    a = gpu.refine(P, T, N, S, theta=25.0, pred_mode=0)
    b = gpu.refine(P, T, N, S, theta=25.0, pred_mode=1)
    assert np.array_equal(a.points.view(np.uint64), b.points.view(np.uint64))
    assert np.array_equal(a.triangles, b.triangles)
    assert b.stats["pred_fp32"] > 0
