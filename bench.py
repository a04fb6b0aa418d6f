"""Benchmark: CDT quality refinement on one MI355X.

Default workload (the headline single-GPU configuration): 1,000,000 uniform
FP64 points in [0, 1e6]^2 plus 50,000 random non-crossing segments no longer
than 2% of the domain width, convex hull enclosed by segments, constrained
Delaunay triangulated on the host (untimed input preparation), then refined
at theta = 25 deg with off-centers in Ruppert mode.

A "step" is one call of the drop-in GPU_Refine_Quality (reference signature,
refine.h:385-386) on that CDT: host triangulateio in, upload, refinement,
compaction, download, host triangulateio out -- the boundary the reference
times (main.cpp:1860-1868).  Two forms are timed in the same run:
  plain   FP64-filtered exact predicates, one claim iteration, Triangle's
          shortest-edge queue order, host-driven passes (the reference's
          structure)
  chosen  the drop-in's default: FP32 -> FP64 -> double-double -> exact
          predicate cascade, claim iterations with shared rings (winners'
          rings may overlap, their cavities may not meet another winner's
          region), the learned claim priority compiled into the library,
          fixed launch shapes
value = output triangles / median wall time of the chosen form's calls.
The same forms are also timed device-resident (a session whose CDT stays
in HBM; the chosen form with autotuned shapes), and the plain form with the
reference's lowest-index claim rule (priority 0, refine.cu:4159-4172).

Prints ONE JSON line.  Also reports:
  roofline      the dominant kernel's algorithmic bytes per launch / its
                average HIP-event duration (events on the refinement stream),
                against the 8 TB/s HBM peak
  whole_job     the whole refinement's algorithmic bytes (record-size model)
                over its wall time
  cpu_baseline  the reference's Triangle "-q" path (compiled from its own
                source into oracle/_ref) on a bounded sample, one host core
  median_step_ms, host
                median step times of each timed form; the host's CPU count,
                model and clock
Other workloads (--workload learned | batch | precision | grid) print their own
line; grid runs the 1M workload over theta {25, 30} x mode x insert.
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

# Cavity kernels: algorithmic bytes are counted from the units the kernels
# report -- every triangle record examined (walk step, seed, conflict test)
# needs its 3 vertex ids, 3 neighbours, 3 subsegments and 3 vertex
# coordinates (36 + 48 B); every candidate reads its id and writes status,
# point, cavity / ring sizes and priority (36 B); every workspace int
# written (cavity + ring list, ordered star boundary) is 4 B; under the
# learned priority a bad candidate also writes 8 float features (32 B).
B_EXAM, B_CAND, B_WS, B_FEAT = 84, 36, 4, 32

# Algorithmic bytes per work item of the other timed kernel groups:
KERNEL_BYTES = {
    "bad_cavity": None,
    "seg_cavity": None,
    "apply": B_PER_INSERT,
    "claim": 10 * 4 + 10 * 8,   # list read + 64-bit claim RMW
    "win": 10 * 4 + 10 * 8 + 4,
    "mark_bad": 1 + 16 + 3 * 16 + 1,
    "compact": 2 * 1,
    "export": 1 + 4 + 16 + 16 + 24,
    "init": 24 + 24 + 48 + 8,
}

MODEL = os.path.join(ROOT, "gpu_cdt_refine_amd", "priority_mlp.npy")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_input(n, nseg, L, maxlen, dist, seed):
    import gpu_cdt_refine_amd as g
    P, S = g.generate_pslg(n, nseg, L=L, maxlen=maxlen, dist=dist, min_angle=5.0, seed=seed)
    T, N, SS = g.build_cdt(P, S)
    return P, T, N, SS


def plain_options():
    return dict(pred_mode=0, claim_iters=1, priority=1, graph=0)


def chosen_options(args):
    # FP32 filter, claim iterations, shared rings, a bad pass evaluating up to
    # 3x the previous pass's winners, learned priority inside
    # 2-half-octave buckets (learned_shift 1 keeps the Steiner count within a
    # few percent of the plain path), segment phases that hand their last
    # few (at most 4) encroached subsegments to the bad pass (host-driven
    # loop; the graph-driven loop is timed separately, without it)
    # launch shapes start from the drop-in's (the autotuner may move them)
    o = dict(pred_mode=1, claim_iters=args.claim_iters, seg_claim_iters=2, mix_below=4,
             priority=1, graph=0,
             shared_ring=1,
             cand_factor=args.cand_factor, grid_cavity=4096, grid_claim=2048, grid_apply=2048,
             cavity_lanes=16, cav_lds_cap=32)
    if args.learned and os.path.exists(MODEL):
        o["priority"] = 2
        o["learned_shift"] = args.learned_shift
    return o


def timed_runs(sess, steps, warmup, tag):
    for i in range(warmup):
        st = sess.run()
        log(f"# {tag} warmup {i}: {st['device_ms']:.1f} ms device, {st['n_triangles']} tris, "
            f"rounds {st['rounds']}")
    walls, devs = [], []
    for i in range(steps):
        t0 = time.perf_counter()
        st = sess.run()
        walls.append(time.perf_counter() - t0)
        devs.append(st["device_ms"])
        log(f"# {tag} step {i}: wall {walls[-1] * 1e3:.1f} ms, device {st['device_ms']:.1f} ms")
    MEDIANS[tag] = {"wall_ms": float(np.median(walls)) * 1e3, "device_ms": float(np.median(devs))}
    return st, float(np.mean(walls)) * 1e3, float(np.mean(devs))


MEDIANS = {}  # per timed form: median step time (wall and HIP-event device time)


def host_info():
    """The host the run used: logical CPUs, CPU model and clock."""
    info = {"nproc": os.cpu_count(), "cpu_model": None, "cpu_mhz": None}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k == "model name" and info["cpu_model"] is None:
                    info["cpu_model"] = v.strip()
                elif k == "cpu MHz" and info["cpu_mhz"] is None:
                    info["cpu_mhz"] = float(v)
    except (OSError, ValueError):
        pass
    return info


def summary(st):
    keys = ["n_points", "n_triangles", "n_subsegs", "rounds", "seg_passes", "bad_passes", "status",
            "giveups", "cand_evals", "lost", "big_cavities", "inserted", "seg_splits",
            "pred_fp32", "pred_fp64", "pred_dd", "pred_exact"]
    return {k: st[k] for k in keys}


def whole_job_bytes(npts, ntri, nseg, st):
    n_ins = int(st["inserted"] + st["seg_splits"])
    return (B_POINT * npts + B_TRI * ntri + B_SUBSEG * nseg + B_PER_INSERT * n_ins +
            B_POINT * st["n_points"] + B_OUT_TRI * st["n_triangles"] + B_OUT_SEG * st["n_subsegs"])


def profile_kernels(P, T, N, SS, theta, opts):
    """One refinement with a HIP-event pair around every launch (host-driven
    passes: the same kernels and work as the graph loop, same mesh)."""
    import gpu_cdt_refine_amd as g
    o = dict(opts)
    o["graph"] = 0
    prof = g.Session(P, T, N, SS, theta=theta, insert=g.Offcenter, mode=g.RUPPERT, profile=1, **o)
    if o.get("priority") == 2:
        prof.set_priority_model(np.load(MODEL))
    st = prof.run()
    kt = prof.kernel_times()
    prof.close()
    return kt, st


TRAFFIC = os.path.join(ROOT, "profiles", "traffic_bad_cavity_refine_1M.json")


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, calibrated), when one exists for this
    workload; None otherwise."""
    try:
        t = json.load(open(TRAFFIC))
    except (OSError, ValueError):
        return None
    if t.get("kernel") != kernel or t.get("workload") != workload:
        return None
    return t["traffic_bytes_per_launch"]


def kernel_bytes(name, d, st, learned):
    """Algorithmic bytes of all launches of kernel group `name`, and a
    description of the model."""
    if name in ("bad_cavity", "seg_cavity"):
        kind = "bad" if name == "bad_cavity" else "seg"
        e, w, c = st[f"tri_examined_{kind}"], st[f"ws_ints_{kind}"], d["items"]
        feat = B_FEAT * c if (learned and kind == "bad") else 0
        return (B_EXAM * e + B_CAND * c + B_WS * w + feat,
                {"bytes_per_examined_triangle": B_EXAM, "bytes_per_candidate": B_CAND
                 + (B_FEAT if feat else 0), "bytes_per_workspace_int": B_WS,
                 "examined_triangles_per_launch": e / max(1, d["launches"]),
                 "candidates_per_launch": c / max(1, d["launches"]),
                 "workspace_ints_per_launch": w / max(1, d["launches"])})
    return KERNEL_BYTES[name] * d["items"], {"bytes_per_item": KERNEL_BYTES[name],
                                             "items_per_launch": d["items"] / max(1, d["launches"])}


def roofline(kt, st, workload, learned):
    dom = max((k for k in kt if k in KERNEL_BYTES), key=lambda k: kt[k]["total_ms"])
    d = kt[dom]
    per_launch_ms = d["total_ms"] / max(1, d["launches"])
    total, model = kernel_bytes(dom, d, st, learned)
    bytes_per_launch = total / max(1, d["launches"])
    achieved = bytes_per_launch / (per_launch_ms / 1e3) / 1e9
    out = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
           "traffic": measured_traffic(dom, workload),
           "algorithmic_bytes_per_launch": bytes_per_launch,
           "avg_launch_ms": per_launch_ms, "launches": d["launches"]}
    out.update(model)
    return out


def cpu_baseline(args, n_sample):
    """Reference Triangle -q, one host core, on a bounded sample of the same
    workload: an n_sample-point instance of the same PSLG family (same point
    density, segment density and length cap, same seed), refined with the
    reference's CPU_Triangle_Quality switches (main.cpp:389-407)."""
    try:
        from oracle import pyoracle
        if not os.path.exists(pyoracle.TRIREF_SO):
            return None
        L = args.domain * (n_sample / args.points) ** 0.5
        P, T, N, SS = make_input(n_sample, int(args.segments * n_sample / args.points), L,
                                 0.02 * args.domain, 0, args.seed)
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


def calibration_copy():
    """A plain 256 MiB device copy (256 MiB read + 256 MiB written): the
    reference dispatch for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counts.
    Runs before the library touches the device."""
    import torch
    a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    torch.empty_like(a).copy_(a)
    torch.cuda.synchronize()


def refine_workload_name(args):
    return (f"cdt-refine-{args.points // 1000000}M-uniform-{args.segments // 1000}k-segs"
            if args.points % 1000000 == 0 else
            f"cdt-refine-{args.points}-uniform-{args.segments}-segs")


def dropin_runs(g, P, T, N, SS, steps, warmup, tag, theta):
    """W untimed and K timed calls of the drop-in GPU_Refine_Quality (host
    arrays in, host arrays out); each call's own wall time, as the
    reference times it (steady clock around the call, main.cpp:1860-1868)."""
    for i in range(warmup):
        r, dt = g.dropin_refine(P, T, N, SS, theta=theta)
        log(f"# {tag} warmup {i}: {dt * 1e3:.1f} ms wall, {r.stats['device_ms']:.1f} ms device, "
            f"{len(r.triangles)} tris, rounds {r.stats['rounds']}")
    walls, devs = [], []
    for i in range(steps):
        r, dt = g.dropin_refine(P, T, N, SS, theta=theta)
        walls.append(dt * 1e3)
        devs.append(r.stats["device_ms"])
        log(f"# {tag} step {i}: wall {dt * 1e3:.1f} ms, device {r.stats['device_ms']:.1f} ms")
    MEDIANS[tag] = {"wall_ms": float(np.median(walls)), "device_ms": float(np.median(devs))}
    return r, walls, devs


def bench_refine(args):
    if args.calibration_copy:
        calibration_copy()
    import gpu_cdt_refine_amd as g
    t0 = time.time()
    P, T, N, SS = make_input(args.points, args.segments, args.domain, 0.02 * args.domain, 0,
                             args.seed)
    log(f"# input: {len(P)} points, {len(T)} triangles, {len(SS)} segments "
        f"(prep {time.time() - t0:.1f}s)")
    kw = dict(theta=args.theta, insert=g.Offcenter, mode=g.RUPPERT)
    # ---- headline: the drop-in, host arrays in and out (the reference's
    # timing boundary: steady clock around GPU_Refine_Quality, main.cpp:1860-1868)
    rp = None
    if not args.no_plain:
        g.dropin_set_options(plain=True, **plain_options())
        rp, wp, dp = dropin_runs(g, P, T, N, SS, args.steps, args.warmup, "dropin plain",
                                 args.theta)
    g.dropin_reset()  # the chosen form with the compiled-in network
    rc, wc, dc = dropin_runs(g, P, T, N, SS, args.steps, args.warmup, "dropin chosen", args.theta)
    g.dropin_reset()
    st = rc.stats
    tris = int(st["n_triangles"])
    ms = float(np.median(wc))
    value = tris / (ms / 1e3)
    # ---- device-resident sessions (extra fields): the CDT already in HBM
    resident = {}
    st_p = None
    if not args.no_plain:
        for name, o in (("plain", plain_options()),
                        ("plain_lowest_index", dict(plain_options(), priority=0))):
            sp = g.Session(P, T, N, SS, **kw, **o)
            s_, m_, d_ = timed_runs(sp, args.steps, args.warmup, name)
            sp.close()
            resident[name] = {"ms_per_step": MEDIANS[name]["wall_ms"],
                              "device_ms_per_step": MEDIANS[name]["device_ms"],
                              "value": s_["n_triangles"] / (MEDIANS[name]["wall_ms"] / 1e3),
                              "rounds": s_["rounds"], "n_points": s_["n_points"]}
            if name == "plain":
                st_p = s_
    co = chosen_options(args)
    sc = g.Session(P, T, N, SS, **kw, **co)
    if co["priority"] == 2:
        sc.set_priority_model(np.load(MODEL))
    tune = None
    if not args.no_autotune:
        # each setting timed three times (median): boxes vary by ~10%, a
        # single run per setting picks noise
        best, trials = sc.autotune(reps=3)
        tune = {"best": best, "trials": trials}
        log(f"# autotune: {best}")
        # the graph-driven loop at its own best structure, for the record
        gbest = None
        for gr in (1, 2, 4, 8):
            for gs in (1, 2, 3):
                o = dict(co, **{k: v for k, v in best.items() if k.startswith("grid_") or
                                k in ("cavity_lanes", "cav_lds_cap")})
                o.update(graph=1, graph_rounds=gr, graph_seg_passes=gs, mix_below=0)
                sg = g.Session(P, T, N, SS, **kw, **o)
                if co["priority"] == 2:
                    sg.set_priority_model(np.load(MODEL))
                ts = sorted(sg.run()["device_ms"] for _ in range(3))
                sg.close()
                if gbest is None or ts[1] < gbest["device_ms"]:
                    gbest = {"device_ms": ts[1], "graph_rounds": gr, "graph_seg_passes": gs}
        tune["graph_loop_best"] = gbest
        log(f"# graph loop best: {gbest}")
    st_s, ms_s, dev_s = timed_runs(sc, args.steps, args.warmup, "chosen")
    resident["chosen"] = {"ms_per_step": MEDIANS["chosen"]["wall_ms"],
                          "device_ms_per_step": MEDIANS["chosen"]["device_ms"],
                          "value": st_s["n_triangles"] / (MEDIANS["chosen"]["wall_ms"] / 1e3),
                          "rounds": st_s["rounds"], "n_points": st_s["n_points"],
                          "options": co, "autotune": tune}
    if "plain" in resident:
        resident["speedup_vs_plain"] = (resident["plain"]["ms_per_step"] /
                                        resident["chosen"]["ms_per_step"])
    # exact invariants of the chosen form's mesh, checked in HBM (untimed):
    # range, symmetry, orientation, boundary, Delaunay, subsegments, segment
    # coverage, input points, quality with the MPW exemption, multiplicity
    t1 = time.perf_counter()
    gcheck = sc.check()
    gcheck["ms"] = (time.perf_counter() - t1) * 1e3
    sc.close()
    b_alg = whole_job_bytes(len(P), len(T), len(SS), st)
    whole = {"bytes": b_alg, "achieved_GBs": b_alg / (ms / 1e3) / 1e9,
             "frac": (b_alg / (ms / 1e3) / 1e9) / HBM_PEAK_GBS}
    roof, kt = None, None
    if not args.no_profile_step:
        # per-launch events need the host-driven passes; the drop-in's chosen
        # form and launch shapes
        po = dict(g.CHOSEN_FORM)
        kt, kst = profile_kernels(P, T, N, SS, args.theta, po)
        roof = roofline(kt, kst, refine_workload_name(args), po["priority"] == 2)
    cpu = None if args.no_cpu_baseline else cpu_baseline(args, args.cpu_sample_points)
    out = {
        "metric": "output triangles/sec (refine, theta=25 deg, off-center, Ruppert): chosen form",
        "value": value,
        "unit": "triangles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "mean_ms_per_step": float(np.mean(wc)),
        "device_ms_per_step": float(np.median(dc)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform PSLG, own CDT builder identical to Triangle pzQnc)",
        "config": {"workload": refine_workload_name(args), "points": args.points,
                   "segments": args.segments, "domain": args.domain,
                   "max_segment_length": args.domain * 0.02, "theta_deg": args.theta,
                   "insert": "offcenter", "mode": "ruppert", "seed": args.seed,
                   "boundary": "drop-in GPU_Refine_Quality: host triangulateio in and out, "
                               "median call wall time after warm-up (main.cpp:1860-1868)",
                   "chosen_options": g.CHOSEN_FORM},
        "plain": None if rp is None else {
            "value": rp.stats["n_triangles"] / (np.median(wp) / 1e3),
            "ms_per_step": float(np.median(wp)), "device_ms_per_step": float(np.median(dp)),
            "options": plain_options(), "output": summary(rp.stats)},
        "speedup_vs_plain": None if rp is None else float(np.median(wp)) / ms,
        "steiner_delta_vs_plain": None if rp is None else
            (st["n_points"] - rp.stats["n_points"]) / max(1, rp.stats["n_points"] - len(P)),
        "dropin_overhead": ms / float(np.median(dc)),
        "output": summary(st),
        "device_resident": resident,
        "roofline": roof,
        "whole_job": whole,
        "gpu_check": gcheck,
        "cpu_baseline": cpu,
        "kernel_times_ms": {k: round(v["total_ms"], 3) for k, v in kt.items()} if kt else None,
        "median_step_ms": dict(MEDIANS),
        "host": host_info(),
    }
    print(json.dumps(out))


def bench_grid(args):
    """The 1M-point workload over theta {25, 30} x mode {Ruppert, Chew} x
    insert {circumcenter, off-center}: chosen form vs the plain path per cell,
    each chosen mesh checked in HBM; for the off-center cells Triangle -q (the
    reference's CPU path, which always uses off-centers) on a bounded sample."""
    import gpu_cdt_refine_amd as g
    P, T, N, SS = make_input(args.points, args.segments, args.domain, 0.02 * args.domain, 0,
                             args.seed)
    log(f"# input: {len(P)} points, {len(T)} triangles, {len(SS)} segments")
    n_cpu = args.cpu_sample_points // 2
    Pc = Tc = Sc = None
    if not args.no_cpu_baseline:
        L = args.domain * (n_cpu / args.points) ** 0.5
        Pc, Tc, _, Sc = make_input(n_cpu, int(args.segments * n_cpu / args.points), L,
                                   0.02 * args.domain, 0, args.seed)
    co = chosen_options(args)
    cells = []
    for theta in (25.0, 30.0):
        for mode, mname in ((g.RUPPERT, "ruppert"), (g.CHEW, "chew")):
            for insert, iname in ((g.Circumcenter, "circumcenter"), (g.Offcenter, "offcenter")):
                kw = dict(theta=theta, insert=insert, mode=mode)
                tag = f"{theta:g}/{mname}/{iname}"
                cell = {"theta_deg": theta, "mode": mname, "insert": iname}
                for name, o in (("plain", plain_options()), ("chosen", co)):
                    s = g.Session(P, T, N, SS, **kw, **o)
                    if o.get("priority") == 2:
                        s.set_priority_model(np.load(MODEL))
                    st, ms, dev = timed_runs(s, args.steps, args.warmup, f"{tag} {name}")
                    r = {"ms_per_step": ms, "value": st["n_triangles"] / (ms / 1e3),
                         "steiner": int(st["n_points"] - len(P)), "triangles": int(st["n_triangles"]),
                         "rounds": int(st["rounds"]), "status": int(st["status"])}
                    if name == "chosen":
                        c = s.check()
                        r["check_violations"] = int(g.violations(c))
                        r["min_angle"] = c["min_angle"]
                    s.close()
                    cell[name] = r
                cell["speedup"] = cell["plain"]["ms_per_step"] / cell["chosen"]["ms_per_step"]
                cell["steiner_delta"] = (cell["chosen"]["steiner"] - cell["plain"]["steiner"]) / \
                    max(1, cell["plain"]["steiner"])
                if Pc is not None and iname == "offcenter":
                    from oracle import pyoracle
                    if os.path.exists(pyoracle.TRIREF_SO):
                        t0 = time.time()
                        _, To, _, _ = pyoracle.triangle_refine(Pc, Tc, Sc, theta,
                                                               mode=1 if mode == g.RUPPERT else 0)
                        dt = time.time() - t0
                        cell["cpu_triangle"] = {"value": len(To) / dt, "seconds": dt,
                                                "points_in": int(len(Pc)), "cores": 1}
                log(f"# {tag}: chosen {cell['chosen']['ms_per_step']:.1f} ms, plain "
                    f"{cell['plain']['ms_per_step']:.1f} ms, {cell['speedup']:.2f}x, Steiner "
                    f"{100 * cell['steiner_delta']:+.1f}%")
                cells.append(cell)
    prim = [c for c in cells if c["theta_deg"] == 25.0 and c["mode"] == "ruppert"
            and c["insert"] == "offcenter"][0]
    out = {"metric": "output triangles/sec over the theta x mode x insert grid (1M uniform)",
           "value": prim["chosen"]["value"], "unit": "triangles/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
           "dtype": "f64", "data": "synthetic (seeded uniform PSLG)",
           "config": {"workload": refine_workload_name(args) + "-grid", "points": args.points,
                      "segments": args.segments, "seed": args.seed, "chosen_options": co,
                      "cpu_sample_points": None if Pc is None else int(len(Pc))},
           "cells": cells}
    print(json.dumps(out))


def bench_learned(args):
    """Clustered 2M-point PSLG, 100k segments, theta 28: plain priority vs the
    learned priority, both with the chosen form's other settings."""
    import gpu_cdt_refine_amd as g
    n = args.points if args.points != 1_000_000 else 2_000_000
    nseg = n // 20
    P, T, N, SS = make_input(n, nseg, args.domain, 0.02 * args.domain, 1, args.seed + 1000)
    log(f"# input: {len(P)} points, {len(T)} triangles, {len(SS)} segments")
    res = {}
    runs = [("plain_path", None), ("plain_priority", 1), ("learned_priority", 2)]
    for name, prio in runs:
        if prio is None:
            o = plain_options()
        else:
            o = chosen_options(args)
            o["priority"] = prio
        s = g.Session(P, T, N, SS, theta=args.learned_theta, **o)
        if prio == 2:
            s.set_priority_model(np.load(MODEL))
        st, ms, dev = timed_runs(s, args.steps, args.warmup, name)
        s.close()
        res[name] = {"ms_per_step": ms, "value": st["n_triangles"] / (ms / 1e3), "output": summary(st)}
    p, a, b = res["plain_path"], res["plain_priority"], res["learned_priority"]
    th = f"{args.learned_theta:g}"
    out = {"metric": f"refine time, learned vs plain claim priority (2M clustered, theta={th})",
           "value": b["value"], "unit": "triangles/s", "n_gpus": 1, "higher_is_better": True,
           "config": {"workload": f"cdt-refine-{n // 1000}k-clustered-{nseg // 1000}k-segs-theta{th}",
                      "points": n, "segments": nseg, "seed": args.seed + 1000,
                      "theta_deg": args.learned_theta},
           "speedup": a["ms_per_step"] / b["ms_per_step"],
           "steiner_delta": (b["output"]["n_points"] - a["output"]["n_points"]) /
                            max(1, a["output"]["n_points"] - len(P)),
           "speedup_vs_plain_path": p["ms_per_step"] / b["ms_per_step"],
           "steiner_delta_vs_plain_path": (b["output"]["n_points"] - p["output"]["n_points"]) /
                                          max(1, p["output"]["n_points"] - len(P)),
           "rounds": [p["output"]["rounds"], a["output"]["rounds"], b["output"]["rounds"]],
           "runs": res}
    print(json.dumps(out))


def batch_instances(k, seed, max_points):
    """Config-4 family: n log-uniform in [200, max_points], uniform or
    Gaussian-clustered 50/50, 5% segments, theta U[20.7, 30], random mode."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(200), np.log(max_points), size=k)).astype(int)
    dist = rng.integers(0, 2, size=k)
    theta = rng.uniform(20.7, 30.0, size=k)
    mode = rng.integers(0, 2, size=k)
    return ns, dist, theta, mode


def _triangle_q_count(a):
    from oracle import pyoracle
    P, T, SS, theta, mode = a
    Po, To, No, So = pyoracle.triangle_refine(P, T, SS, theta, mode=mode)
    return len(To)


def _prep_instance(a):
    n, d, s = a
    L = 1000.0 * (n / 10000) ** 0.5
    return make_input(int(n), max(1, int(0.05 * n)), L, 0.05 * L, int(d), int(s))


def bench_batch(args):
    import gpu_cdt_refine_amd as g
    from concurrent.futures import ProcessPoolExecutor
    ns, dist, theta, mode = batch_instances(args.instances, args.seed, args.batch_max_points)
    t0 = time.time()
    jobs = [(int(n), int(d), 7000 + i) for i, (n, d) in enumerate(zip(ns, dist))]
    with ProcessPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        ins = list(ex.map(_prep_instance, jobs, chunksize=8))
    log(f"# {len(ins)} instances, {sum(len(x[0]) for x in ins)} points (prep {time.time() - t0:.1f}s)")
    co = chosen_options(args)
    b = g.Batch(ins, theta=theta, insert=1, mode=mode, **co)
    if co["priority"] == 2:
        b.set_priority_model(np.load(MODEL))
    best = None
    if not args.no_autotune:
        best, trials = b.autotune(reps=1)
        log(f"# autotune: {best}")
    n_waves = b.waves()
    for i in range(args.warmup):
        b.run()
    runs = []
    for i in range(args.steps):
        st = b.run()
        runs.append(st["run_ms"])
        log(f"# batch step {i}: {st['run_ms']:.1f} ms, {st['n_triangles']} tris")
    ms = float(np.mean(runs))
    res = b.results()
    b.close()
    out = {"metric": "batched refinement: output triangles/sec over all instances",
           "value": st["n_triangles"] / (ms / 1e3), "unit": "triangles/s", "n_gpus": 1,
           "ms_per_step": ms, "higher_is_better": True,
           "config": {"workload": f"batch-{args.instances}-pslgs", "instances": args.instances,
                      "input_points": int(sum(len(x[0]) for x in ins)),
                      "max_points": args.batch_max_points, "waves": n_waves, "autotune": best},
           "output": summary(st)}
    # exact invariants on a few instances (the oracle's checker)
    from oracle import pyoracle
    checked = []
    for i in range(0, len(ins), max(1, len(ins) // 4)):
        P, T, N, SS = ins[i]
        r = res[i]
        tot, viol, mina, ex = pyoracle.check_mesh(r.points, r.triangles, r.neighbors, r.segments,
                                                  P, SS, float(theta[i]))
        checked.append({"instance": i, "violations": int(tot), "min_angle": mina})
    out["checked_instances"] = checked
    # every instance on the GPU checker (index range, symmetry, orientation,
    # boundary, exact Delaunay, subsegment edges, segment coverage, moved
    # points, quality with the MPW exemption at the instance's theta,
    # multiplicity)
    t1 = time.time()
    tot = {}
    min_margin = None
    for i, (P, T, N, SS) in enumerate(ins):
        r = res[i]
        c = g.check_mesh(r.points, r.triangles, r.neighbors, r.segments, points_in=P,
                         segments_in=SS, theta=float(theta[i]))
        margin = c.pop("min_angle") - float(theta[i])
        min_margin = margin if min_margin is None else min(min_margin, margin)
        for k, v in c.items():
            tot[k] = tot.get(k, 0) + v
    tot["min_angle_minus_theta"] = min_margin
    tot["instances"] = len(ins)
    tot["seconds"] = time.time() - t1
    out["gpu_check_all_instances"] = tot
    if not args.no_cpu_baseline:
        # the reference's Triangle -q on a sample of the instances, one process
        # per core (Triangle is sequential and keeps global state)
        sample = list(range(0, len(ins), max(1, len(ins) // 48)))
        jobs = [(ins[i][0], ins[i][1], ins[i][3], float(theta[i]), int(mode[i])) for i in sample]
        cores = min(16, os.cpu_count() or 1)
        t1 = time.time()
        with ProcessPoolExecutor(max_workers=cores) as ex:
            tris = sum(ex.map(_triangle_q_count, jobs))
        dt = time.time() - t1
        gpu_tris_sample = sum(len(res[i].triangles) for i in sample)
        out["cpu_baseline"] = {"value": tris / dt, "unit": "triangles/s", "cores": cores,
                               "kind": "reference",
                               "sample": f"Triangle 1.6 -q on {len(sample)} of the {len(ins)} "
                                         f"instances ({tris} output triangles; the GPU made "
                                         f"{gpu_tris_sample} on them), {cores} processes, {dt:.1f} s"}
    print(json.dumps(out))


def bad_triangles(P, T, theta):
    """Mask of the triangles whose smallest angle is below theta (atan2)."""
    a, b, c = P[T[:, 0]], P[T[:, 1]], P[T[:, 2]]

    def ang(p, q, r):
        u, v = q - p, r - p
        return np.degrees(np.abs(np.arctan2(u[:, 0] * v[:, 1] - u[:, 1] * v[:, 0],
                                            (u * v).sum(1))))
    return np.minimum(np.minimum(ang(a, b, c), ang(b, c, a)), ang(c, a, b)) < theta


def bench_precision(args):
    """Precision stress: lattice + cocircular + uniform points, lattice-line
    segments; stage hit rates of the predicate cascade and Steiner-point error
    of the FP64 constructions against exact rational arithmetic."""
    import gpu_cdt_refine_amd as g
    from fractions import Fraction
    from oracle import pyoracle
    n = args.points
    P, T, N, SS = make_input(n, n // 50, 1000.0, 20.0, 2, args.seed)
    log(f"# input: {len(P)} points, {len(T)} triangles, {len(SS)} segments")
    co = chosen_options(args)
    s = g.Session(P, T, N, SS, theta=25.0, **co)
    if co["priority"] == 2:
        s.set_priority_model(np.load(MODEL))
    st, ms, dev = timed_runs(s, args.steps, args.warmup, "precision")
    r = s.download()
    s.close()
    tot_pred = st["pred_fp32"] + st["pred_fp64"] + st["pred_dd"] + st["pred_exact"]
    # Steiner error on the triangles the refinement splits: the bad triangles
    # of the input CDT and of the meshes after 2, 8 and 32 rounds (sessions
    # capped with max_rounds), 300 of each, both insertion modes; the FP64
    # construction (bit-identical to the kernels') against exact rationals
    rng = np.random.default_rng(0)
    errs = {1: [], 0: []}
    for cap in (0, 2, 8, 32):
        if cap == 0:
            Pm, Tm = P, T
        else:
            sc = g.Session(P, T, N, SS, theta=25.0, max_rounds=cap, **co)
            if co["priority"] == 2:
                sc.set_priority_model(np.load(MODEL))
            sc.run()
            rm = sc.download()
            sc.close()
            Pm, Tm = rm.points, rm.triangles
        idx = np.nonzero(bad_triangles(Pm, Tm, 25.0))[0]
        idx = rng.choice(idx, min(300, len(idx)), replace=False)
        for t in idx:
            a, o, d = Pm[Tm[t, 0]], Pm[Tm[t, 1]], Pm[Tm[t, 2]]
            for off in (1, 0):
                c = pyoracle.steiner_point(a, o, d, 25.0, off)
                if c is None:
                    continue
                ex = pyoracle.steiner_point_exact(a, o, d, 25.0, off)
                err = max(abs(Fraction(c[0]) - ex[0]), abs(Fraction(c[1]) - ex[1]))
                scale = max(abs(float(ex[0])), abs(float(ex[1])), 1e-300)
                errs[off].append((float(err), float(err) / (np.spacing(scale))))
    steiner_error = {}
    for off, name in ((1, "offcenter"), (0, "circumcenter")):
        e = np.array(errs[off])
        steiner_error[name] = {"samples": int(len(e)), "max_abs": float(e[:, 0].max()),
                               "median_ulps": float(np.median(e[:, 1])),
                               "p99_ulps": float(np.percentile(e[:, 1], 99)),
                               "max_ulps": float(e[:, 1].max())}
    out = {"metric": "precision stress: predicate stage hit rates and Steiner error",
           "value": st["n_triangles"] / (ms / 1e3), "unit": "triangles/s", "n_gpus": 1,
           "ms_per_step": ms, "higher_is_better": True,
           "config": {"workload": f"precision-stress-{n}", "points": n, "segments": len(SS)},
           "stage_rates": {k: st[k] / max(1, tot_pred) for k in
                           ("pred_fp32", "pred_fp64", "pred_dd", "pred_exact")},
           "stage_counts": {k: st[k] for k in ("pred_fp32", "pred_fp64", "pred_dd", "pred_exact")},
           "steiner_error": steiner_error,
           "giveups": {k: st[k] for k in ("giveups", "giveup_tris_left", "giveup_segs_left")},
           "output": summary(st)}
    print(json.dumps(out))


def bench_cdt(args):
    """Construction of the input CDT (points + segments -> triangles with hull
    segments, Triangle's "pzQnc" contract) on the GPU (gcdt_build_cdt_device:
    host arrays in and out, median wall time), next to this library's host
    builder and the reference's Triangle (oracle/_ref, one core), on the
    refine workload's 10^6 uniform points / 50,000 segments and the learned
    workload's 2*10^6 clustered points / 100,000 segments.  The GPU result is
    compared with the host builder's (the same triangle and segment sets)."""
    import gpu_cdt_refine_amd as g
    from oracle import pyoracle

    def canon(T):
        T = np.asarray(T).reshape(-1, 3)
        r = np.argmin(T, axis=1)
        i = np.arange(len(T))
        return set(map(tuple, np.stack([T[i, r], T[i, (r + 1) % 3], T[i, (r + 2) % 3]],
                                       1).tolist()))

    cases = []
    for n, ns, dist, seed, name in ((args.points, args.segments, 0, args.seed, "uniform"),
                                    (2 * args.points, 2 * args.segments, 1, args.seed + 1000,
                                     "clustered")):
        t0 = time.perf_counter()
        P, S0 = g.generate_pslg(n, ns, L=args.domain, maxlen=0.02 * args.domain, dist=dist,
                                min_angle=5.0, seed=seed)
        t_gen = time.perf_counter() - t0
        t0 = time.perf_counter()
        Th, Nh, Sh = g.build_cdt(P, S0)
        t_host = time.perf_counter() - t0
        walls, devs = [], []
        for i in range(args.warmup + args.steps):
            st = {}
            t0 = time.perf_counter()
            T, N, S = g.build_cdt(P, S0, device=True, stats=st)
            if i >= args.warmup:
                walls.append((time.perf_counter() - t0) * 1e3)
                devs.append(st["device_ms"])
            log(f"# cdt {name} {n}: {(time.perf_counter() - t0) * 1e3:.1f} ms wall, {st}")
        same = (canon(T) == canon(Th) and
                {tuple(sorted(x)) for x in S.tolist()} == {tuple(sorted(x)) for x in Sh.tolist()})
        case = {"workload": f"{name}-{n}-points-{ns}-segments", "points": n, "segments": ns,
                "triangles": int(len(T)), "gpu_ms": float(np.median(walls)),
                "gpu_device_ms": float(np.median(devs)), "rounds": st["rounds"],
                "segment_passes": st["segment_passes"], "host_builder_ms": t_host * 1e3,
                "generator_ms": t_gen * 1e3, "same_as_host_builder": bool(same)}
        if not args.no_cpu_baseline and os.path.exists(pyoracle.TRIREF_SO):
            t0 = time.perf_counter()
            Pt, Tt, Nt, St = pyoracle.triangle_cdt(P, S0)
            case["triangle_pzQnc_ms"] = (time.perf_counter() - t0) * 1e3
        cases.append(case)
    big = cases[-1]
    out = {"metric": "input CDT construction on the GPU: points/sec (host arrays in and out)",
           "value": big["points"] / (big["gpu_ms"] / 1e3), "unit": "points/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": big["gpu_ms"],
           "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "cdt-construction"}, "cases": cases,
           "cpu_baseline": None if "triangle_pzQnc_ms" not in big else {
               "value": big["points"] / (big["triangle_pzQnc_ms"] / 1e3), "unit": "points/s",
               "cores": 1, "kind": "reference",
               "sample": f"Triangle 1.6 'pzQnc' (reference GPU_Refine/triangle.cpp) on the same "
                         f"{big['points']}-point input, one run"},
           "host": host_info()}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="refine",
                    choices=["refine", "learned", "batch", "precision", "grid", "cdt"])
    ap.add_argument("--points", type=int, default=1_000_000)
    ap.add_argument("--segments", type=int, default=50_000)
    ap.add_argument("--domain", type=float, default=1e6)
    ap.add_argument("--theta", type=float, default=25.0)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--claim-iters", type=int, default=3)
    ap.add_argument("--learned", type=int, default=1)
    ap.add_argument("--learned-theta", type=float, default=28.0)
    ap.add_argument("--learned-shift", type=int, default=1)
    ap.add_argument("--cand-factor", type=int, default=3)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile-step", action="store_true")
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="start with a 256 MiB device copy (counter calibration)")
    ap.add_argument("--cpu-sample-points", type=int, default=200_000)
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--batch-max-points", type=int, default=200_000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is a design note only)"}))
        sys.exit(2)
    {"refine": bench_refine, "learned": bench_learned, "batch": bench_batch,
     "precision": bench_precision, "grid": bench_grid, "cdt": bench_cdt}[args.workload](args)


if __name__ == "__main__":
    main()
