"""Writes tests/golden/*.json.  Run in the build container, where the
reference tree is present:

    python tests/golden/make_golden.py

triangle_q_counts.json: output sizes of the reference's own Triangle -q
(GPU_Refine/triangle.cpp, compiled into oracle/_ref; switches of
CPU_Triangle_Quality, main.cpp:389-407) on the project's seeded config-1
inputs (10,000 uniform points in [0,1000]^2, 500 non-crossing segments).
oracle_digest.json: SHA-256 of the oracle's refined mesh for one small case,
a regression pin for the exact contract the GPU reproduces."""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import gpu_cdt_refine_amd as g  # noqa: E402
from oracle import pyoracle  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def config1(seed):
    P, S = g.generate_pslg(10000, 500, L=1000.0, maxlen=0.0, dist=0, min_angle=5.0, seed=seed)
    T, N, SS = g.build_cdt(P, S)
    return P, T, N, SS


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    counts = []
    for seed in (1, 2):
        P, T, N, SS = config1(seed)
        for theta, mode in ((25, 1), (25, 0), (30, 1)):
            Pt, Tt, Nt, St = pyoracle.triangle_refine(P, T, SS, theta, mode)
            counts.append({"seed": seed, "theta": theta, "mode": mode,
                           "cdt_points": int(len(P)), "cdt_triangles": int(len(T)),
                           "cdt_segments": int(len(SS)), "cdt_digest": digest(P, T, N, SS),
                           "points": int(len(Pt)), "triangles": int(len(Tt)),
                           "segments": int(len(St))})
    with open(os.path.join(HERE, "triangle_q_counts.json"), "w") as f:
        json.dump(counts, f, indent=1)
    P, S = g.generate_pslg(1500, 75, L=1000.0, maxlen=200.0, dist=0, min_angle=5.0, seed=77)
    T, N, SS = g.build_cdt(P, S)
    Po, To, No, So, pf, st = pyoracle.oracle_refine(P, T, N, SS, theta=25.0)
    with open(os.path.join(HERE, "oracle_digest.json"), "w") as f:
        json.dump({"input": "generate_pslg(1500, 75, L=1000, maxlen=200, dist=0, "
                            "min_angle=5, seed=77) -> build_cdt; theta 25, off-center, Ruppert",
                   "points": int(len(Po)), "triangles": int(len(To)),
                   "segments": int(len(So)), "digest": digest(Po, To, No, So)}, f, indent=1)


if __name__ == "__main__":
    main()
