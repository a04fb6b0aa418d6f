"""Learned-priority MLP on the GPU: the MFMA operand layout, the forward pass
and the gradient against the CPU oracle (oracle/mlp_oracle.py, float64 with
the kernels' BF16 operand roundings), training convergence, and a learned-
priority refinement that still satisfies every exact mesh invariant."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_mfma_layout(gpu):
    from oracle.mlp_oracle import bf16
    rng = np.random.default_rng(0)
    A = rng.normal(size=(16, 16)).astype(np.float32)
    B = rng.normal(size=(16, 16)).astype(np.float32)
    D = gpu.mfma_probe(A, B)
    ref = bf16(A) @ bf16(B)
    assert np.allclose(D, ref, rtol=1e-5, atol=1e-5)


def test_mfma32_layout(gpu):
    """gfx950's 16x16x32 BF16 MFMA with the operand layout the inference
    kernel uses: equal to the product of the BF16-rounded operands."""
    from oracle.mlp_oracle import bf16
    rng = np.random.default_rng(2)
    A = rng.normal(size=(16, 32)).astype(np.float32)
    B = rng.normal(size=(32, 16)).astype(np.float32)
    D = gpu.mfma_probe32(A, B)
    ref = bf16(A).astype(np.float64) @ bf16(B).astype(np.float64)
    assert np.allclose(D, ref, rtol=1e-5, atol=1e-5)


def test_mlp_forward_and_grad_vs_oracle(gpu):
    from oracle import mlp_oracle as mo
    rng = np.random.default_rng(1)
    n = 1000
    X = rng.normal(size=(n, 16)).astype(np.float32)
    t = rng.normal(size=n).astype(np.float32)
    p = gpu.mlp_init(3)
    p[16 * 64:16 * 64 + 64] = 0.1  # nonzero biases
    y = gpu.mlp_forward(X, p)
    yr, _ = mo.forward(X, p)
    # FP32 accumulation of BF16 products: tolerance 1e-3 of the output scale
    assert np.max(np.abs(y - yr)) < 1e-3 * (1 + np.max(np.abs(yr)))
    g, loss = gpu.mlp_grad(X, t, p)
    gr, lr_ = mo.grad(X, t, p)
    assert abs(loss - lr_) < 1e-3 * (1 + lr_)
    # per-block relative error of the gradient, tolerance 2e-3
    for a, b in ((0, 1024), (1024, 1088), (1088, 5184), (5184, 5248), (5248, 5312), (5312, 5313)):
        err = np.linalg.norm(g[a:b] - gr[a:b]) / (np.linalg.norm(gr[a:b]) + 1e-12)
        assert err < 2e-3, (a, b, err)


def test_mlp_training_reduces_loss(gpu):
    rng = np.random.default_rng(2)
    n = 20000
    X = rng.uniform(-1, 1, size=(n, 16)).astype(np.float32)
    t = (np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2]).astype(np.float32)
    p0 = gpu.mlp_init(4)
    p, losses = gpu.mlp_train(X, t, p0, epochs=20, batch=1024, lr=3e-3, seed=1)
    assert losses[-1] < 0.25 * losses[0], losses


def test_learned_priority_refinement_invariants(gpu, oracle):
    """With a (random-init) model ordering candidates, every insertion is
    still decided by exact predicates: the mesh passes all invariants and the
    run is deterministic."""
    pts, segs = oracle.random_pslg(4000, 200, L=1000.0, maxlen=200.0, seed=51)
    P, T, N, S = oracle.triangle_cdt(pts, segs)
    res = []
    for _ in range(2):
        s = gpu.Session(P, T, N, S, theta=28.0, priority=2, claim_iters=2)
        s.set_priority_model(gpu.mlp_init(7))
        st = s.run()
        r = s.download()
        s.close()
        assert st["status"] == 0
        tot, viol, mina, ex = oracle.check_mesh(r.points, r.triangles, r.neighbors,
                                                r.segments, P, S, 28.0)
        assert tot == 0, viol
        res.append(r)
    assert np.array_equal(res[0].points.view(np.uint64), res[1].points.view(np.uint64))
    assert np.array_equal(res[0].triangles, res[1].triangles)


def test_training_logs(gpu, oracle):
    pts, segs = oracle.random_pslg(3000, 150, L=1000.0, maxlen=200.0, seed=52)
    P, T, N, S = oracle.triangle_cdt(pts, segs)
    s = gpu.Session(P, T, N, S, theta=25.0, log_training=1)
    s.run()
    X, t = s.logs()
    s.close()
    assert len(X) > 1000 and X.shape[1] == 16
    assert np.all(np.isfinite(X)) and np.all(t >= 0)
    assert X[:, 0].max() <= 25.0 / 60.0 + 1e-3  # bad triangles: min angle below theta


def test_trained_priority_gain(gpu):
    """What the trained network gains over the method without it: on a
    clustered PSLG (seed disjoint from the training seeds), the learned claim
    priority needs fewer refinement rounds than the shortest-edge rule, with
    the Steiner count within 5%.  The same run with an untrained (random)
    network is the control: it must not beat the trained one."""
    import os
    model = os.path.join(os.path.dirname(gpu.__file__), "priority_mlp.npy")
    P, S = gpu.generate_pslg(300000, 15000, L=1e6 * (0.3 ** 0.5), maxlen=2e4, dist=1,
                             min_angle=5.0, seed=4242)
    T, N, SS = gpu.build_cdt(P, S)
    out = {}
    for name, prio, params in (("plain", 1, None), ("learned", 2, np.load(model)),
                               ("untrained", 2, gpu.mlp_init(9))):
        s = gpu.Session(P, T, N, SS, theta=28.0, priority=prio, claim_iters=3, pred_mode=1)
        if params is not None:
            s.set_priority_model(params)
        out[name] = s.run()
        s.close()
    print({k: (v["rounds"], v["n_points"]) for k, v in out.items()}, flush=True)
    plain, learned, untrained = out["plain"], out["learned"], out["untrained"]
    assert learned["status"] == 0
    assert learned["rounds"] <= 0.92 * plain["rounds"]
    assert learned["rounds"] <= untrained["rounds"]
    steiner_plain = plain["n_points"] - len(P)
    assert abs(learned["n_points"] - plain["n_points"]) <= 0.05 * steiner_plain
