"""Batched refinement in the CPU oracle: the union of independent CDTs, each
with its own theta / insertion policy / mode.  While no cross-instance
coupling can act -- every bad triangle fits in each pass (cand_min) and the
cavity capacity cannot change (cav_cap0 == cav_cap_max) -- every instance must
come out exactly as refined alone: instances share no triangle, candidate
order inside an instance is preserved by the union's index maps, and an
instance with nothing to do waits unchanged."""
import numpy as np
import pytest

FIXED_K = dict(cav_cap0=256, cav_cap_max=256, cand_min=1 << 29)


def _instances(oracle, specs):
    out = []
    for n, ns, seed in specs:
        pts, segs = oracle.random_pslg(n, ns, L=1000.0, maxlen=250.0, seed=seed)
        out.append(oracle.triangle_cdt(pts, segs))
    return out


def test_batch_of_one_is_the_single_mesh_path(oracle):
    (P, T, N, S), = _instances(oracle, [(1500, 60, 3)])
    res, st = oracle.oracle_refine_batch([(P, T, N, S)], [27.0], [1], [1])
    Po, To, No, So, pf, st1 = oracle.oracle_refine(P, T, N, S, theta=27.0)
    Pb, Tb, Nb, Sb, pfb = res[0]
    assert np.array_equal(Pb.view(np.uint64), Po.view(np.uint64))
    assert np.array_equal(Tb, To) and np.array_equal(Nb, No) and np.array_equal(Sb, So)
    assert np.array_equal(pfb, pf)
    for k in ("rounds", "inserted", "seg_splits", "giveups"):
        assert st[k] == st1[k]


def test_batch_instances_equal_solo_refinements(oracle):
    specs = [(800, 40, 11), (2500, 120, 12), (300, 20, 13), (1800, 90, 14), (1200, 80, 15)]
    ins = _instances(oracle, specs)
    theta = [20.7, 25.0, 30.0, 28.0, 22.5]
    off = [1, 1, 0, 1, 0]
    mode = [1, 0, 1, 1, 0]
    res, st = oracle.oracle_refine_batch(ins, theta, off, mode, **FIXED_K)
    assert st["status"] == 0
    for i, (P, T, N, S) in enumerate(ins):
        Po, To, No, So, pf, s1 = oracle.oracle_refine(P, T, N, S, theta=theta[i], offcenter=off[i],
                                                      mode=mode[i], **FIXED_K)
        Pb, Tb, Nb, Sb, pfb = res[i]
        assert np.array_equal(Pb.view(np.uint64), Po.view(np.uint64)), i
        assert np.array_equal(Tb, To), i
        assert np.array_equal(Nb, No), i
        assert np.array_equal(Sb, So), i
        tot, viol, mina, ex = oracle.check_mesh(Pb, Tb, Nb, Sb, P, S, theta[i])
        assert tot == 0, (i, viol)


def test_batch_parameters_are_per_instance(oracle):
    """The same CDT twice, at 20 and 30 degrees: the instances differ, and
    each meets its own bound."""
    (P, T, N, S), = _instances(oracle, [(1000, 50, 21)])
    res, st = oracle.oracle_refine_batch([(P, T, N, S), (P, T, N, S)], [20.0, 30.0], [1, 1], [1, 1])
    assert len(res[1][0]) > len(res[0][0])
    for i, th in enumerate([20.0, 30.0]):
        Pb, Tb, Nb, Sb, pfb = res[i]
        tot, viol, mina, ex = oracle.check_mesh(Pb, Tb, Nb, Sb, P, S, th)
        assert tot == 0, viol
        assert mina >= th - 1e-6 or ex > 0
