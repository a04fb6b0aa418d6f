"""Benchmark: CDT quality refinement on one MI355X.

Workload (the headline single-GPU configuration): 1,000,000 uniform FP64
points in [0, 1e6]^2 plus 50,000 random non-crossing segments no longer than
2% of the domain width, convex hull enclosed by segments, constrained
Delaunay triangulated on the host (untimed input preparation), then refined
at theta = 25 deg with off-centers in Ruppert mode.

A "step" is one full refinement of that CDT (GPU_Refine_Quality's work),
starting from the CDT already resident in HBM and ending with the compacted
output mesh in HBM.  value = output triangles / second of wall time.

Prints ONE JSON line (rank 0).  Also reports:
  roofline      the dominant kernel's algorithmic bytes per launch / its
                average HIP-event duration, against the 8 TB/s HBM peak
  whole_job     the whole refinement's algorithmic bytes (record-size model)
                over its wall time
  cpu_baseline  the reference's Triangle "-q" path (compiled from its own
                source into oracle/_ref) on the same CDT, one host core
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Record sizes (bytes) of the refinement's data model.
B_POINT, B_TRI, B_SUBSEG = 16, 36, 12
B_OUT_TRI, B_OUT_SEG = 24, 8
B_PER_INSERT = 424  # 3 parent vertices + 4-triangle cavity read, 6-triangle star + point written

# Algorithmic bytes per work item for each kernel:
KERNEL_BYTES = {
    # read parent record+vertices (36+48), 4 cavity records (144) + their new
    # vertices (32), write cavity+ring list (10*4) and candidate record (40)
    "bad_cavity": 36 + 48 + 144 + 32 + 40 + 40,
    "seg_cavity": 12 + 32 + 72 + 144 + 32 + 40 + 40,
    "apply": B_PER_INSERT,
    "claim": 10 * 4 + 10 * 8,   # list read + 64-bit claim RMW
    "win": 10 * 4 + 10 * 8 + 4,
    "mark_bad": 1 + 12 + 48 + 1,
    "compact": 2 * 1,
    "export": 1 + 4 + 24 + 24 + 12,
    "init": 24 + 24 + 48 + 8,
}


def workload_config(args):
    return {
        "workload": "cdt-refine-1M-uniform-50k-segs",
        "points": args.points,
        "segments": args.segments,
        "domain": args.domain,
        "max_segment_length": args.domain * 0.02,
        "theta_deg": args.theta,
        "insert": "offcenter",
        "mode": "ruppert",
        "seed": args.seed,
    }


def make_input(args):
    import gpu_cdt_refine_amd as g
    t0 = time.time()
    P, S = g.generate_pslg(args.points, args.segments, L=args.domain,
                           maxlen=0.02 * args.domain, dist=0, min_angle=5.0,
                           seed=args.seed)
    T, N, SS = g.build_cdt(P, S)
    return P, T, N, SS, time.time() - t0


def cpu_baseline(args, n_sample):
    """Reference Triangle -q, one host core, on a bounded sample of the same
    workload: an n_sample-point instance of the same PSLG family (same point
    density, segment density and length cap, same seed), refined with the
    reference's CPU_Triangle_Quality switches (main.cpp:389-407)."""
    try:
        from oracle import pyoracle
        import gpu_cdt_refine_amd as g
        if not os.path.exists(pyoracle.TRIREF_SO):
            return None
        L = args.domain * (n_sample / args.points) ** 0.5
        P, S = g.generate_pslg(n_sample, int(args.segments * n_sample / args.points), L=L,
                               maxlen=0.02 * args.domain, dist=0, min_angle=5.0, seed=args.seed)
        T, N, SS = g.build_cdt(P, S)
        t0 = time.time()
        Po, To, No, So = pyoracle.triangle_refine(P, T, SS, args.theta, mode=1)
        dt = time.time() - t0
        return {"value": len(To) / dt, "unit": "triangles/s", "cores": 1,
                "kind": "reference",
                "sample": f"Triangle 1.6 'pzDQnrq{args.theta:g}' (reference GPU_Refine/triangle.cpp) "
                          f"on a {n_sample}-point instance of the same PSLG family "
                          f"({len(SS)} segments, same density); {dt:.1f} s on 1 core",
                "seconds": dt, "triangles_out": int(len(To)), "points_out": int(len(Po)),
                "points_in": int(len(P))}
    except Exception as e:  # baseline is informative only
        return {"error": str(e)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--points", type=int, default=1_000_000)
    ap.add_argument("--segments", type=int, default=50_000)
    ap.add_argument("--domain", type=float, default=1e6)
    ap.add_argument("--theta", type=float, default=25.0)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--pred-mode", type=int, default=0)
    ap.add_argument("--priority", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-points", type=int, default=200_000)
    ap.add_argument("--claim-iters", type=int, default=1)
    ap.add_argument("--no-profile-step", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is a design note only)"}))
        sys.exit(2)

    import gpu_cdt_refine_amd as g
    P, T, N, SS, prep_s = make_input(args)
    print(f"# input: {len(P)} points, {len(T)} triangles, {len(SS)} segments "
          f"(prep {prep_s:.1f}s)", file=sys.stderr, flush=True)
    sess = g.Session(P, T, N, SS, theta=args.theta, insert=g.Offcenter, mode=g.RUPPERT,
                     pred_mode=args.pred_mode, priority=args.priority,
                     claim_iters=args.claim_iters)
    for i in range(args.warmup):
        st = sess.run()
        print(f"# warmup {i}: {st['device_ms']:.1f} ms device, {st['n_triangles']} tris, "
              f"rounds {st['rounds']}", file=sys.stderr, flush=True)
    walls, devs = [], []
    for i in range(args.steps):
        t0 = time.perf_counter()
        st = sess.run()
        walls.append(time.perf_counter() - t0)
        devs.append(st["device_ms"])
        print(f"# step {i}: wall {walls[-1]*1e3:.1f} ms, device {st['device_ms']:.1f} ms",
              file=sys.stderr, flush=True)
    ms = 1e3 * float(np.mean(walls))
    tris = int(st["n_triangles"])
    value = tris / (ms / 1e3)

    # algorithmic bytes of the whole refinement (record-size model)
    n_ins = int(st["inserted"] + st["seg_splits"])
    b_alg = (B_POINT * len(P) + B_TRI * len(T) + B_SUBSEG * len(SS) + B_PER_INSERT * n_ins +
             B_POINT * st["n_points"] + B_OUT_TRI * tris + B_OUT_SEG * st["n_subsegs"])
    whole = {"bytes": b_alg, "achieved_GBs": b_alg / (ms / 1e3) / 1e9,
             "frac": (b_alg / (ms / 1e3) / 1e9) / HBM_PEAK_GBS}

    roof = None
    kt = None
    if not args.no_profile_step:
        prof = g.Session(P, T, N, SS, theta=args.theta, insert=g.Offcenter, mode=g.RUPPERT,
                         pred_mode=args.pred_mode, priority=args.priority,
                         claim_iters=args.claim_iters, profile=1)
        prof.run()
        kt = prof.kernel_times()
        prof.close()
        dom = max((k for k in kt if k in KERNEL_BYTES), key=lambda k: kt[k]["total_ms"])
        d = kt[dom]
        per_launch_ms = d["total_ms"] / max(1, d["launches"])
        bytes_per_launch = KERNEL_BYTES[dom] * d["items"] / max(1, d["launches"])
        achieved = bytes_per_launch / (per_launch_ms / 1e3) / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "avg_launch_ms": per_launch_ms, "launches": d["launches"],
                "bytes_per_item": KERNEL_BYTES[dom]}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, args.cpu_sample_points)

    out = {
        "metric": "output triangles/sec (refine, theta=25 deg, off-center, Ruppert)",
        "value": value,
        "unit": "triangles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "device_ms_per_step": float(np.mean(devs)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform PSLG, own CDT builder)",
        "config": workload_config(args),
        "output": {"points": int(st["n_points"]), "triangles": tris,
                   "subsegments": int(st["n_subsegs"]), "rounds": st["rounds"],
                   "seg_passes": st["seg_passes"], "bad_passes": st["bad_passes"],
                   "status": st["status"], "giveups": st["giveups"],
                   "cand_evals": st["cand_evals"], "lost": st["lost"],
                   "big_cavities": st["big_cavities"], "inserted": st["inserted"],
                   "seg_splits": st["seg_splits"]},
        "roofline": roof,
        "whole_job": whole,
        "cpu_baseline": cpu,
        "kernel_times_ms": {k: round(v["total_ms"], 3) for k, v in kt.items()} if kt else None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
