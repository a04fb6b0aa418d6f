"""The drop-in GPU_Refine_Quality (reference signature, refine.h:385-386).

- The reference-style C++ caller (examples/dropin_main.cpp) linked against
  libgcdt.so through include/gpu_cdt_refine.h gets the chosen form: the same
  mesh as a device-resident session with the chosen options and the trained
  network, and a mesh that passes every invariant.
- In-process calls through the C++ symbol: with the plain options the mesh
  equals the oracle's bit for bit; calls in a row on inputs of different
  sizes (the cached refiner's buffers reused, grown) equal fresh runs; the
  debug status arrays equal the C ABI's."""
import os
import struct
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "examples", "dropin_main")
MODEL = os.path.join(ROOT, "gpu_cdt_refine_amd", "priority_mlp.npy")


def _same(a, P, T, N, S):
    assert np.array_equal(a.points.view(np.uint64), P.view(np.uint64))
    assert np.array_equal(a.triangles, T) and np.array_equal(a.neighbors, N)
    assert np.array_equal(a.segments, S)


def _chosen_session(gpu, P, T, N, S, theta):
    s = gpu.Session(P, T, N, S, theta=theta, **gpu.CHOSEN_FORM)
    try:
        s.set_priority_model(np.load(MODEL))
        st = s.run()
        r = s.download()
    finally:
        s.close()
    return r, st


def test_dropin_binary(gpu, oracle, tmp_path):
    if not os.path.exists(BIN):
        pytest.skip("examples/dropin_main not built")
    pts, segs = oracle.random_pslg(3000, 150, L=1000.0, maxlen=200.0, seed=23)
    P, T, N, S = oracle.triangle_cdt(pts, segs)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(struct.pack("iii", len(P), len(T), len(S)))
        f.write(P.astype(np.float64).tobytes())
        f.write(T.astype(np.int32).tobytes())
        f.write(N.astype(np.int32).tobytes())
        f.write(S.astype(np.int32).tobytes())
    r = subprocess.run([BIN, str(fin), str(fout), "25", "1"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(fout, "rb").read()
    np_, nt, ns = struct.unpack("iii", raw[:12])
    off = 12
    Po = np.frombuffer(raw, np.float64, 2 * np_, off).reshape(-1, 2); off += 16 * np_
    To = np.frombuffer(raw, np.int32, 3 * nt, off).reshape(-1, 3); off += 12 * nt
    No = np.frombuffer(raw, np.int32, 3 * nt, off).reshape(-1, 3); off += 12 * nt
    So = np.frombuffer(raw, np.int32, 2 * ns, off).reshape(-1, 2)
    c, _ = _chosen_session(gpu, P, T, N, S, 25.0)
    _same(c, Po, To, No, So)
    tot, viol, mina, ex = oracle.check_mesh(Po, To, No, So, P, S, 25.0)
    assert tot == 0, viol
    steiner = int(r.stdout.split("steiner")[1])
    assert steiner == np_ - len(P)


def test_dropin_plain_equals_oracle_and_reuse(gpu, oracle):
    """Plain options: bit-identical to the oracle.  Four calls in a row on
    inputs of 2000, 6000, 1000 and 6000 points (buffers reused, then grown,
    then reused): each equals its oracle run."""
    gpu.dropin_set_options(plain=True)
    try:
        for n, seed, theta, mode in ((2000, 31, 25.0, 1), (6000, 32, 28.0, 0),
                                     (1000, 33, 22.0, 1), (6000, 32, 28.0, 0)):
            print(f"dropin plain n={n} seed={seed}", flush=True)
            pts, segs = oracle.random_pslg(n, n // 20, L=1000.0, maxlen=200.0, seed=seed)
            P, T, N, S = oracle.triangle_cdt(pts, segs)
            r, dt = gpu.dropin_refine(P, T, N, S, theta=theta, mode=mode, debug=True)
            Po, To, No, So, pf, st = oracle.oracle_refine(P, T, N, S, theta=theta, mode=mode)
            _same(r, Po, To, No, So)
            assert r.stats["status"] == st["status"] == 0
            h = gpu.refine(P, T, N, S, theta=theta, mode=mode)
            assert np.array_equal(r.pstatus, h.pstatus) and np.array_equal(r.tstatus, h.tstatus)
            assert dt > 0
    finally:
        gpu.dropin_reset()


def test_dropin_chosen_form_default(gpu, oracle):
    """The default drop-in (chosen form, compiled-in network) equals a
    session with the chosen options and the network file, twice in a row,
    and passes the invariants."""
    gpu.dropin_reset()
    pts, segs = oracle.random_pslg(20000, 1000, L=1000.0, maxlen=200.0, seed=34)
    P, T, N, S = oracle.triangle_cdt(pts, segs)
    c, st = _chosen_session(gpu, P, T, N, S, 25.0)
    for _ in range(2):
        r, dt = gpu.dropin_refine(P, T, N, S, theta=25.0)
        _same(r, c.points, c.triangles, c.neighbors, c.segments)
        assert r.stats["rounds"] == st["rounds"]
    tot, viol, mina, ex = oracle.check_mesh(r.points, r.triangles, r.neighbors, r.segments, P,
                                            S, 25.0)
    assert tot == 0, viol
    gpu.dropin_reset()
