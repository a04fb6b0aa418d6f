"""The non-allocating checker entry point (gcdt_check_mesh_device_noalloc)
against the allocating one (gcdt_check_mesh): the same counts, exemptions and
smallest angle on two calls in a row through the C ABI, the second on a
smaller mesh with a planted fault (the shared static scratch keeps the first
call's contents, which must not leak into the second), and a mesh beyond the
scratch limit rejected before any device work.  The test's own device
buffers come from the HIP runtime the library links (test plumbing; the
entry point under test allocates nothing)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libgcdt.so uses."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        if a.nbytes:
            assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def get(self, ptr, a):
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(a.ctypes.data, C.c_void_p(ptr), a.nbytes, 2) == 0
        return a

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _noalloc(gpu, hip, P, T, N, S, Pin, Sin, theta):
    d = [hip.put(x) for x in (P, T, N, S, Pin, Sin)]
    out = hip.put(np.full(13, -7, np.int64))
    gpu.check_mesh_device_noalloc(d[0], len(P), d[1], d[2], len(T), d[3], len(S), d[4],
                                  len(Pin), d[5], len(Sin), theta, out, None)
    return hip.get(out, np.zeros(13, np.int64))


def _expect(gpu, P, T, N, S, Pin, Sin, theta):
    c = gpu.check_mesh(P, T, N, S, points_in=Pin, segments_in=Sin, theta=theta)
    v = [c[k] for k in gpu.CHECK_NAMES]
    return v, c


def _compare(gpu, got, P, T, N, S, Pin, Sin, theta):
    v, c = _expect(gpu, P, T, N, S, Pin, Sin, theta)
    assert list(got[:10]) == v, (list(got[:10]), c)
    assert got[10] == c["exact_signs"]
    assert got[11] == c["exempt"]
    assert np.int64(got[12]).view(np.float64) == c["min_angle"], (got[12], c)
    return c


@pytest.fixture
def hip(gpu):
    h = Hip()
    yield h
    h.free()


def test_noalloc_check_matches_allocating_entry(gpu, oracle, hip):
    pts, segs = oracle.random_pslg(6000, 300, L=1000.0, maxlen=200.0, seed=211)
    P, T, N, S = oracle.triangle_cdt(pts, segs)
    r = gpu.refine(P, T, N, S, theta=25.0, claim_iters=3, shared_ring=1)
    print(f"noalloc call 1: {len(r.triangles)} triangles", flush=True)
    got = _noalloc(gpu, hip, r.points, r.triangles, r.neighbors, r.segments, P, S, 25.0)
    c = _compare(gpu, got, r.points, r.triangles, r.neighbors, r.segments, P, S, 25.0)
    assert gpu.violations(c) == 0 and c["min_angle"] >= 25.0 - 1e-9

    # a smaller mesh with a flipped interior edge (Delaunay / quality faults)
    pts2, segs2 = oracle.random_pslg(1500, 80, L=1000.0, maxlen=200.0, seed=212)
    P2, T2, N2, S2 = oracle.triangle_cdt(pts2, segs2)
    r2 = gpu.refine(P2, T2, N2, S2, theta=28.0)
    To = r2.triangles.copy()
    segset = {tuple(sorted(x)) for x in r2.segments.tolist()}
    for t in range(len(To)):
        j = next((j for j in range(3) if r2.neighbors[t, j] > t and
                  tuple(sorted((To[t, (j + 1) % 3], To[t, (j + 2) % 3]))) not in segset), None)
        if j is not None:
            u = r2.neighbors[t, j]
            k = [kk for kk in range(3) if r2.neighbors[u, kk] == t][0]
            a, b, cc, dd = To[t, (j + 1) % 3], To[t, (j + 2) % 3], To[t, j], To[u, k]
            To[t] = [cc, a, dd]
            To[u] = [dd, b, cc]
            break
    print(f"noalloc call 2: {len(To)} triangles, planted flip", flush=True)
    got2 = _noalloc(gpu, hip, r2.points, To, r2.neighbors, r2.segments, P2, S2, 28.0)
    c2 = _compare(gpu, got2, r2.points, To, r2.neighbors, r2.segments, P2, S2, 28.0)
    assert gpu.violations(c2) > 0


def test_noalloc_check_rejects_oversized_mesh(gpu, hip):
    """The size check runs on the host before anything is enqueued (no
    pointer is dereferenced), so tiny buffers serve."""
    p = hip.put(np.zeros(16, np.int64))
    with pytest.raises(RuntimeError, match="scratch"):
        gpu.check_mesh_device_noalloc(p, 2_000_000, p, p, 4_000_000, p, 10, p, 0, p, 4, 25.0, p)
