"""The reference-style C++ caller (examples/dropin_main.cpp) linked against
libgcdt.so through include/gpu_cdt_refine.h: GPU_Refine_Quality with the
reference's signature returns the same mesh as the C ABI and the oracle."""
import os
import struct
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "examples", "dropin_main")


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
    Pq, Tq, Nq, Sq, pf, st = oracle.oracle_refine(P, T, N, S, theta=25.0)
    assert np.array_equal(Po.view(np.uint64), Pq.view(np.uint64))
    assert np.array_equal(To, Tq) and np.array_equal(No, Nq) and np.array_equal(So, Sq)
    steiner = int(r.stdout.split("steiner")[1])
    assert steiner == np_ - len(P)
