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


def test_device_fp32_stage_exact_at_tiny_scales(gpu):
    from test_predicates_cpu import tiny_scale_queries
    rng = np.random.default_rng(99)
    q = tiny_scale_queries(rng, 600)
    o, ic, stages = gpu.device_predicates(q, pred_mode=1)
    exp_o = np.array([exact_orient(r[0:2], r[2:4], r[4:6]) for r in q])
    exp_i = np.array([exact_incircle(r[0:2], r[2:4], r[4:6], r[6:8]) for r in q])
    assert np.array_equal(o, exp_o)
    assert np.array_equal(ic, exp_i)


def test_device_constructions_equal_host(gpu):
    """The adaptive orient2d value and the Steiner point on the device equal
    the same source on the host bit for bit (which the CPU tests pin against
    the reference's Triangle)."""
    from test_predicates_cpu import near_collinear_triangles
    rng = np.random.default_rng(5)
    tris = near_collinear_triangles(rng, 3000)
    for theta, off in ((25.0, 1), (25.0, 0), (30.0, 1)):
        hc, hs, hok = gpu.host_constructions(tris, theta, off)
        dc, ds, dok = gpu.device_constructions(tris, theta, off)
        assert np.array_equal(hc.view(np.uint64), dc.view(np.uint64))
        assert np.array_equal(hok, dok)
        m = hok != 0
        assert np.array_equal(hs[m].view(np.uint64), ds[m].view(np.uint64))
