"""Train the learned claim priority on the GPU and write its weights.

    python -m gpu_cdt_refine_amd.train_priority [--out FILE] [--points N]

Training data are the conflict logs the refinement itself produces
(options.log_training = 1): for every valid bad-pass candidate, its 16
features (smallest angle, circumradius / shortest edge, cavity size, ring
size, distance to the nearest subsegment, pass index) and its conflict
degree -- how many other candidates claim a triangle of its cavity or ring;
the target is log2(1 + degree).  Inputs are uniform PSLGs of the benchmark's
family (segments capped at 2% of the width, theta 25 and 30 degrees,
off-center, Ruppert) with seeds TRAIN_SEEDS, disjoint from every seed the
benchmarks evaluate on.  Training runs the library's BF16-MFMA forward /
backward / Adam kernels (gcdt_mlp_train).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import gpu_cdt_refine_amd as g  # noqa: E402

TRAIN_SEEDS = (101, 102, 103)
TRAIN_THETAS = (25.0, 30.0)
DEFAULT_OUT = os.path.join(HERE, "priority_mlp.npy")


def collect(points, segments_per_point=0.05, claim_iters=2, max_samples=3_000_000, seeds=TRAIN_SEEDS,
            thetas=TRAIN_THETAS, log=print):
    Xs, ts = [], []
    for seed in seeds:
        L = 1e6 * (points / 1e6) ** 0.5
        P, S = g.generate_pslg(points, int(points * segments_per_point), L=L, maxlen=0.02e6,
                               dist=0, min_angle=5.0, seed=seed)
        T, N, SS = g.build_cdt(P, S)
        for theta in thetas:
            s = g.Session(P, T, N, SS, theta=theta, log_training=1, claim_iters=claim_iters)
            try:
                st = s.run()
                X, t = s.logs()
            finally:
                s.close()
            log(f"# seed {seed} theta {theta}: {len(X)} samples, {st['rounds']} rounds, "
                f"{st['n_triangles']} triangles")
            Xs.append(X)
            ts.append(t)
    X = np.concatenate(Xs)
    t = np.concatenate(ts)
    if len(X) > max_samples:
        idx = np.random.default_rng(0).choice(len(X), max_samples, replace=False)
        X, t = X[idx], t[idx]
    return X, t


def train(X, t, epochs=12, batch=4096, lr=2e-3, seed=0):
    p0 = g.mlp_init(seed)
    p, losses = g.mlp_train(X, t, p0, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return p, losses


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--points", type=int, default=300_000)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--claim-iters", type=int, default=2)
    args = ap.parse_args()
    t0 = time.time()
    X, t = collect(args.points, claim_iters=args.claim_iters)
    t1 = time.time()
    # hold out 10% to report generalisation
    rng = np.random.default_rng(1)
    perm = rng.permutation(len(X))
    nv = len(X) // 10
    va, tr = perm[:nv], perm[nv:]
    p, losses = train(X[tr], t[tr], epochs=args.epochs)
    t2 = time.time()
    yv = g.mlp_forward(X[va], p)
    val_mse = float(np.mean((yv - t[va]) ** 2))
    base_mse = float(np.mean((t[va] - t[tr].mean()) ** 2))
    np.save(args.out, p.astype(np.float32))
    meta = {"train_seeds": list(TRAIN_SEEDS), "thetas": list(TRAIN_THETAS),
            "points": args.points, "samples": int(len(tr)), "epochs": args.epochs,
            "losses": [float(x) for x in losses], "val_mse": val_mse,
            "val_mse_constant_predictor": base_mse, "collect_s": t1 - t0, "train_s": t2 - t1}
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    if os.path.abspath(args.out) == os.path.abspath(DEFAULT_OUT):
        # the drop-in's compiled-in copy (rebuild libgcdt.so to pick it up)
        from gpu_cdt_refine_amd import embed_model
        embed_model.main()
    print(json.dumps(meta))


if __name__ == "__main__":
    main()
