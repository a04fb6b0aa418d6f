"""CDT files: the reference's text layout (saveCDT / readCDT,
GPU_Refine/mesh.cpp:402-532), the binary layout and Triangle's
.node/.ele/.neigh/.poly files.  Host-only code; runs without a GPU."""
import os

import numpy as np
import pytest


@pytest.fixture(scope="module")
def g():
    import gpu_cdt_refine_amd as g
    g.lib()
    return g


def _mesh(g, n=500, seed=4):
    P, S = g.generate_pslg(n, n // 20, L=1000.0, maxlen=200.0, dist=0, min_angle=5.0, seed=seed)
    T, N, SS = g.build_cdt(P, S)
    rng = np.random.default_rng(seed)
    # coordinates that need all 17 digits
    P = P * (1.0 + rng.uniform(-1e-9, 1e-9, size=P.shape))
    return P, T, N, SS


def _reference_readcdt(path):
    """The reference reader's line logic (mesh.cpp:430-532), for checking that
    it accepts the files written here."""
    with open(path) as f:
        lines = f.read().split("\n")
    np_, nt, ns = int(lines[0]), int(lines[1]), int(lines[2])
    assert all(len(line) < 100 for line in lines)  # fgets(buf, 100, fp)
    o = 3
    P = np.array([[float(v) for v in lines[o + i].split()[:2]] for i in range(np_)]).reshape(-1, 2)
    o += np_
    T = np.array([[int(v) for v in lines[o + i].split()[:3]] for i in range(nt)], np.int32).reshape(-1, 3)
    o += nt
    N = np.array([[int(v) for v in lines[o + i].split()[:3]] for i in range(nt)], np.int32).reshape(-1, 3)
    o += nt
    S = np.array([[int(v) for v in lines[o + i].split()[:2]] for i in range(ns)], np.int32).reshape(-1, 2)
    return P, T, N, S


@pytest.mark.parametrize("binary", [False, True])
def test_round_trip_is_exact(g, tmp_path, binary):
    P, T, N, S = _mesh(g)
    path = tmp_path / ("m.bin" if binary else "m.cdt")
    g.save_cdt(path, P, T, N, S, binary=binary)
    P2, T2, N2, S2 = g.load_cdt(path)
    assert np.array_equal(P2.view(np.uint64), P.view(np.uint64))
    assert np.array_equal(T2, T) and np.array_equal(N2, N) and np.array_equal(S2, S)


def test_reference_reader_takes_our_text(g, tmp_path):
    P, T, N, S = _mesh(g, seed=5)
    path = tmp_path / "m.cdt"
    g.save_cdt(path, P, T, N, S)
    P2, T2, N2, S2 = _reference_readcdt(path)
    assert np.array_equal(P2.view(np.uint64), P.view(np.uint64))
    assert np.array_equal(T2, T) and np.array_equal(N2, N) and np.array_equal(S2, S)


def test_reads_reference_saved_files(g, tmp_path):
    """saveCDT's own output: "%lf" coordinates (six decimals), Windows line
    endings allowed."""
    P, T, N, S = _mesh(g, seed=6)
    path = tmp_path / "ref.cdt"
    rows = [f"{len(P)}", f"{len(T)}", f"{len(S)}"]
    rows += ["%f %f" % (x, y) for x, y in P]
    rows += ["%d %d %d" % tuple(t) for t in T]
    rows += ["%d %d %d" % tuple(nb) for nb in N]
    rows += ["%d %d" % tuple(s) for s in S]
    path.write_bytes(("\r\n".join(rows) + "\r\n").encode())
    P2, T2, N2, S2 = g.load_cdt(path)
    want = np.array([[float("%f" % x), float("%f" % y)] for x, y in P])
    assert np.array_equal(P2, want)
    assert np.array_equal(T2, T) and np.array_equal(N2, N) and np.array_equal(S2, S)


def test_malformed_files_raise(g, tmp_path):
    P, T, N, S = _mesh(g, n=100)
    path = tmp_path / "m.cdt"
    g.save_cdt(path, P, T, N, S)
    text = path.read_text().split("\n")
    (tmp_path / "short.cdt").write_text("\n".join(text[:len(text) // 2]))
    with pytest.raises(OSError, match="ends inside"):
        g.load_cdt(tmp_path / "short.cdt")
    (tmp_path / "bad.cdt").write_text("abc\n1\n1\n")
    with pytest.raises(OSError, match="bad point count"):
        g.load_cdt(tmp_path / "bad.cdt")
    with pytest.raises(OSError, match="cannot open"):
        g.load_cdt(tmp_path / "missing.cdt")
    b = tmp_path / "m.bin"
    g.save_cdt(b, P, T, N, S, binary=True)
    data = b.read_bytes()
    (tmp_path / "cut.bin").write_bytes(data[:len(data) - 10])
    with pytest.raises(OSError, match="too short"):
        g.load_cdt(tmp_path / "cut.bin")


def test_triangle_files(g, tmp_path):
    P, T, N, S = _mesh(g, seed=7)
    base = tmp_path / "mesh"
    g.save_triangle_files(base, P, T, N, S)

    def rows(ext):
        lines = [ln.split() for ln in open(f"{base}.{ext}").read().splitlines() if ln.strip()]
        return lines[0], lines[1:]

    h, r = rows("node")
    assert h == [str(len(P)), "2", "0", "0"]
    Pn = np.array([[float(a), float(b)] for _, a, b in r])
    assert [int(x[0]) for x in r] == list(range(len(P)))
    assert np.array_equal(Pn.view(np.uint64), P.view(np.uint64))
    h, r = rows("ele")
    assert h == [str(len(T)), "3", "0"]
    assert np.array_equal(np.array([[int(v) for v in x[1:]] for x in r]), T)
    h, r = rows("neigh")
    assert h == [str(len(T)), "3"]
    assert np.array_equal(np.array([[int(v) for v in x[1:]] for x in r]), N)
    h, r = rows("poly")
    assert h == ["0", "2", "0", "0"]
    assert r[0] == [str(len(S)), "0"]
    assert np.array_equal(np.array([[int(v) for v in x[1:]] for x in r[1:1 + len(S)]]), S)
    assert r[1 + len(S)] == ["0"]
