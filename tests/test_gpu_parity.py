"""GPU refinement vs the CPU oracle: bit-identical meshes, exact invariants,
run-to-run determinism.  Inputs are seeded PSLGs triangulated by the
reference's Triangle ("pzQnc", mesh.cpp:397)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # (n points, n segments, seed, theta, offcenter, mode)
    (2000, 100, 1, 25.0, 1, 1),
    (2000, 100, 1, 25.0, 1, 0),
    (2000, 100, 2, 30.0, 1, 1),
    (2000, 100, 3, 25.0, 0, 1),
    (2000, 100, 4, 30.0, 0, 0),
    (2000, 100, 5, 20.7, 1, 1),
    (20000, 1000, 6, 25.0, 1, 1),
    (20000, 1000, 7, 30.0, 1, 0),
]


def _input(oracle, n, ns, seed):
    pts, segs = oracle.random_pslg(n, ns, L=1000.0, maxlen=200.0, seed=seed)
    P, T, N, S = oracle.triangle_cdt(pts, segs)
    return P, T, N, S


@pytest.mark.parametrize("n,ns,seed,theta,off,mode", CASES)
def test_gpu_equals_oracle(gpu, oracle, n, ns, seed, theta, off, mode):
    print(f"case n={n} seed={seed} theta={theta} off={off} mode={mode}", flush=True)
    P, T, N, S = _input(oracle, n, ns, seed)
    r = gpu.refine(P, T, N, S, theta=theta, insert=off, mode=mode)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, theta=theta, offcenter=off, mode=mode)
    assert r.stats["status"] == 0 and st["status"] == 0
    assert r.stats["rounds"] == st["rounds"]
    assert np.array_equal(r.points.view(np.uint64), Po.view(np.uint64))
    assert np.array_equal(r.triangles, To)
    assert np.array_equal(r.neighbors, No)
    assert np.array_equal(r.segments, So)
    tot, viol, mina, ex = oracle.check_mesh(r.points, r.triangles, r.neighbors, r.segments,
                                            P, S, theta)
    assert tot == 0, viol
    assert mina >= theta - 1e-6


FAMILIES = [
    # (dist, n points, theta, offcenter, mode): the reference generator's
    # disk, circle (ring), grid and ellipse families (mesh.cpp:131-193),
    # segments no longer than 5% of the domain
    (3, 4000, 25.0, 1, 1),
    (4, 4000, 28.0, 1, 0),
    (5, 4000, 25.0, 0, 1),
    (6, 800, 28.0, 1, 1),
]


@pytest.mark.parametrize("dist,n,theta,off,mode", FAMILIES)
def test_gpu_equals_oracle_families(gpu, oracle, dist, n, theta, off, mode):
    print(f"family dist={dist} n={n} theta={theta} off={off} mode={mode}", flush=True)
    L = 1000.0 * (n / 10000) ** 0.5
    P0, S0 = gpu.generate_pslg(n, n // 20, L=L, maxlen=0.05 * L, dist=dist, min_angle=5.0,
                               seed=40 + dist)
    P, T, N, S = oracle.triangle_cdt(P0, S0)
    r = gpu.refine(P, T, N, S, theta=theta, insert=off, mode=mode)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, theta=theta, offcenter=off, mode=mode)
    assert r.stats["status"] == 0 and st["status"] == 0
    assert r.stats["rounds"] == st["rounds"]
    assert np.array_equal(r.points.view(np.uint64), Po.view(np.uint64))
    assert np.array_equal(r.triangles, To)
    assert np.array_equal(r.neighbors, No)
    assert np.array_equal(r.segments, So)
    tot, viol, mina, ex = oracle.check_mesh(r.points, r.triangles, r.neighbors, r.segments,
                                            P, S, theta)
    assert tot == 0, viol
    assert mina >= theta - 1e-6


def test_gpu_deterministic(gpu, oracle):
    P, T, N, S = _input(oracle, 5000, 250, 11)
    a = gpu.refine(P, T, N, S, theta=28.0)
    b = gpu.refine(P, T, N, S, theta=28.0)
    assert np.array_equal(a.points.view(np.uint64), b.points.view(np.uint64))
    assert np.array_equal(a.triangles, b.triangles)
    assert np.array_equal(a.neighbors, b.neighbors)
    assert np.array_equal(a.segments, b.segments)


@pytest.mark.parametrize("iters,seg_iters", [(2, 0), (3, 0), (3, 2)])
def test_gpu_equals_oracle_claim_iters(gpu, oracle, iters, seg_iters):
    """Claim iterations per pass; seg_iters > 0 gives passes that split only
    subsegments their own count (host loop and graph loop)."""
    P, T, N, S = _input(oracle, 5000, 250, 21)
    r = gpu.refine(P, T, N, S, theta=28.0, claim_iters=iters, seg_claim_iters=seg_iters)
    if seg_iters:
        rg = gpu.refine(P, T, N, S, theta=28.0, claim_iters=iters, seg_claim_iters=seg_iters,
                        graph=1)
        assert np.array_equal(rg.points.view(np.uint64), r.points.view(np.uint64))
        assert np.array_equal(rg.triangles, r.triangles)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, theta=28.0, claim_iters=iters,
                                                  seg_claim_iters=seg_iters)
    assert np.array_equal(r.points.view(np.uint64), Po.view(np.uint64))
    assert np.array_equal(r.triangles, To)
    assert np.array_equal(r.neighbors, No)


def circle_pslg(oracle, k=72, n_out=1500, seed=3):
    """k points on a circle of radius 100 around (500, 500), empty inside,
    plus random points outside: circumcenters of the skinny fan triangles
    inside sit near the centre, so their cavities hold most of the fan."""
    rng = np.random.default_rng(seed)
    a = 2 * np.pi * np.arange(k) / k
    circ = np.stack([500 + 100 * np.cos(a), 500 + 100 * np.sin(a)], 1)
    out = rng.uniform(0, 1000, size=(4 * n_out, 2))
    out = out[np.hypot(out[:, 0] - 500, out[:, 1] - 500) > 130][:n_out]
    pts = np.concatenate([circ, out])
    P, T, N, S = oracle.triangle_cdt(pts, np.zeros((0, 2), np.int32))
    return P, T, N, S


@pytest.mark.parametrize("cap0", [4, 64])
def test_gpu_cavity_capacity_paths(gpu, oracle, cap0):
    """Small K forces overflow and K growth; large cavities (most of a
    72-point circle's fan) exercise multi-level wave-parallel growth."""
    P, T, N, S = circle_pslg(oracle)
    kw = dict(cav_cap0=cap0, cav_cap_max=256)
    r = gpu.refine(P, T, N, S, theta=25.0, **kw)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, theta=25.0, **kw)
    if cap0 == 4:
        assert st["rej_overflow"] > 0
    assert st["final_cav_cap"] >= 64
    assert np.array_equal(r.points.view(np.uint64), Po.view(np.uint64))
    assert np.array_equal(r.triangles, To)
    assert np.array_equal(r.neighbors, No)
    assert np.array_equal(r.segments, So)
    tot, viol, mina, ex = oracle.check_mesh(r.points, r.triangles, r.neighbors, r.segments,
                                            P, S, 25.0)
    assert tot == 0, viol


def test_gpu_candidate_cap_by_priority(gpu, oracle):
    """A small workspace forces priority-ordered truncation every pass."""
    P, T, N, S = _input(oracle, 4000, 200, 41)
    kw = dict(cand_min=64, cand_factor=2, cand_budget=(6 * 32 + 16) * 500)
    r = gpu.refine(P, T, N, S, theta=25.0, **kw)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, theta=25.0, **kw)
    assert np.array_equal(r.points.view(np.uint64), Po.view(np.uint64))
    assert np.array_equal(r.triangles, To)


@pytest.mark.parametrize("pred_mode,graph", [(0, 0), (1, 0), (1, 1)])
def test_gpu_equals_oracle_precision_stress(gpu, oracle, pred_mode, graph):
    """Jittered lattice, cocircular groups and lattice-line segments: the
    inputs that drive the predicate cascade past FP64 and the near-vertex
    give-up rule (a bad triangle whose Steiner point lands within
    1e-6 x its shortest edge of a cavity vertex is given up).  The mesh is
    bit-identical to the oracle's under every predicate mode and both loop
    drivers."""
    import gpu_cdt_refine_amd as g
    print(f"precision stress pred_mode={pred_mode} graph={graph}", flush=True)
    P, S = g.generate_pslg(20000, 400, L=1000.0 * 0.02 ** 0.5, maxlen=20.0, dist=2,
                           min_angle=5.0, seed=2)
    T, N, SS = g.build_cdt(P, S)
    r = gpu.refine(P, T, N, SS, theta=25.0, pred_mode=pred_mode, graph=graph)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, SS, theta=25.0)
    assert st["giveups"] > 0
    for k in ("rounds", "inserted", "seg_splits", "giveups", "status", "giveup_tris_left",
              "giveup_segs_left"):
        assert r.stats[k] == st[k], k
    # converged except for the given-up triangles, which are reported
    left = st["giveup_tris_left"] + st["giveup_segs_left"]
    assert st["status"] == (3 if left else 0)
    tot, viol, mina, ex = oracle.check_mesh(r.points, r.triangles, r.neighbors, r.segments, P,
                                            SS, 25.0)
    assert viol["quality"] <= st["giveup_tris_left"], (viol, st)
    assert tot == viol["quality"], viol
    assert np.array_equal(r.points.view(np.uint64), Po.view(np.uint64))
    assert np.array_equal(r.triangles, To)
    assert np.array_equal(r.neighbors, No)
    assert np.array_equal(r.segments, So)
    assert r.stats["pred_exact"] + r.stats["pred_dd"] > 0


def _near_degenerate_family_instance(gpu):
    """A Gaussian-clustered instance of the batched family (69,161 points,
    3,458 segments) in which two segment chains pass within ~1e-7 of each
    other near coordinate 2000: the refinement must resolve that feature, and
    four segment splits there meet a nearly collinear neighbour vertex on the
    wrong side of FP64 rounding (their star would be inverted)."""
    n = 69161
    L = 1000.0 * (n / 10000) ** 0.5
    P, S = gpu.generate_pslg(n, n // 20, L=L, maxlen=0.05 * L, dist=1, min_angle=5.0, seed=7000 + 4009)
    T, N, SS = gpu.build_cdt(P, S)
    return P, T, N, SS


def test_gpu_equals_oracle_given_up_subsegments(gpu, oracle):
    """Given-up subsegments: bad triangles that encroach (or walk into) only
    given-up subsegments are given up too, so the run ends with status 3 and
    counts them; every quality violation left is one of them.  Bit-identical
    to the oracle."""
    print("given-up subsegments: family instance 4009", flush=True)
    P, T, N, S = _near_degenerate_family_instance(gpu)
    kw = dict(claim_iters=3, shared_ring=1, priority=1)
    theta = 21.33356473295036
    r = gpu.refine(P, T, N, S, theta=theta, **kw)
    Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, theta=theta, **kw)
    for k in ("status", "rounds", "inserted", "seg_splits", "giveups", "giveup_tris_left",
              "giveup_segs_left"):
        assert r.stats[k] == st[k], k
    assert st["status"] == 3 and st["giveup_segs_left"] > 0, st
    assert np.array_equal(r.points.view(np.uint64), Po.view(np.uint64))
    assert np.array_equal(r.triangles, To)
    assert np.array_equal(r.neighbors, No)
    c = gpu.check_mesh(r.points, r.triangles, r.neighbors, r.segments, points_in=P,
                       segments_in=S, theta=theta)
    assert c["quality"] <= st["giveup_tris_left"], (c, st)
    assert gpu.violations(c) == c["quality"], c


@pytest.mark.parametrize("graph", [0, 1])
def test_gpu_pstatus_matches_oracle(gpu, oracle, graph):
    """Debug point status (PStatus bits 1 Steiner, 2 segment split, 3 new:
    inserted in the run's last round, refine.h:36-106) of a run stopped by
    the round cap, and of a converged one, against the oracle's point flags."""
    P, T, N, S = _input(oracle, 5000, 250, 31)
    for rounds in (6, 100000):
        print(f"pstatus graph={graph} rounds={rounds}", flush=True)
        kw = dict(theta=27.0, mode=0, max_rounds=rounds)
        r = gpu.refine(P, T, N, S, graph=graph, **kw)
        Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, **kw)
        assert r.stats["status"] == st["status"] == (1 if rounds == 6 else 0)
        assert np.array_equal(r.points.view(np.uint64), Po.view(np.uint64))
        want = (np.where(pf & 1, 0, 2) | np.where(pf & 2, 4, 0) | np.where(pf & 16, 8, 0))
        assert np.array_equal(r.pstatus, want.astype(np.uint8))
        new = (r.pstatus & 8) != 0
        assert not new[:len(P)].any()
        if rounds == 6:
            assert new.sum() > 0
