#!/usr/bin/env python3
"""Benchmark of the segment-tree clipping path on one MI355X.

Main workload (the value of the JSON line): plain pairwise clip of two
harmonic star rings, base 200,000 vertices (seed 2001), overlay 1,000,000
vertices (seed 2002, centre (0.15,-0.1), rotated 17 degrees), FP64.  One step
= the whole hot path on the device with both rings already resident in HBM:
edge records, elementary intervals, node intervals, end lists, cover lists,
stab + candidate dedupe, intersections / augmented rings / initial labels,
labelling, traversal and clean-up of the output rings.  Value: input edges
clipped per second (N_b + N_c per step).

Batched workloads (reported under "batched", each next to the plain per-pair
GPU loop on a sample of the same pairs):
  one_to_many   base 100,000-vertex harmonic star against 4,096 star overlays
                (log-uniform 8..20,000 vertices): one tree, every overlay
                clipped in one batch
  small_pairs   16,384 pairs of 4..64-vertex rings: clip + area/IoU forward
                (the persistent small-pair kernel; also replayed as a
                captured hipGraph, and on the staged path for comparison),
                and forward + backward of the area / IoU
  large_pairs   256 pairs of 10k..200k-vertex harmonic stars: forward +
                backward with the clip recomputed in backward and with the
                forward's state kept, peak HBM of each, the stab's HBM rate

roofline: the stab kernel (the dominant kernel), HBM bound.  Algorithmic
bytes per launch = 36 (L + E) + 12 K: every cover entry (L) and leaf end entry
(E; upper-level end lists are ranges of the leaf entries, not stored) read
once with its 32-byte coordinate record, plus a 12-byte triple per emitted
candidate (K), divided by the launch time from HIP events on the library's
stream, against 8 TB/s.  traffic: the kernel's measured bytes per launch
from the rocprofv3 FETCH_SIZE / WRITE_SIZE counter runs summarised in
profiles/clip_harmonic_200k_1m_hbm_traffic.md (default workload only).
Those counters count L2-to-fabric requests, so Infinity Cache hits are
included: an upper bound on DRAM bytes.

cpu_baseline: the CPU oracle (restatement of the reference's multicore
path, OpenMP loops where the reference has them, results independent of the
thread count) on the same 200,000 x 1,000,000 pair, timed on this host with
OMP_NUM_THREADS (or every available core) threads.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "segment-tree-based-clipping_amd", "python"))

METRIC = "edges clipped/s and pairs/s on 1 GPU, batched vs plain; fwd+bwd area pairs/s"
HBM_PEAK_GBPS = 8000.0


def stab_bytes(cover_entries, end_entries, candidates):
    """Algorithmic HBM bytes of one stab launch: lists read once with a 32-byte
    coordinate record per entry (4-byte entry + 32-byte gather), 12 bytes per
    emitted candidate triple."""
    return 36 * (cover_entries + end_entries) + 12 * candidates


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "clip_harmonic_200k_1m_hbm_traffic.md")
# settings chosen by `bench.py --tune` on the plain 200k x 1M clip (segclip.tune)
TUNING_FILE = os.path.join(ROOT, "segment-tree-based-clipping_amd", "tuning.json")
USE_TUNING = True


def tuning_for(name):
    """Settings that `bench.py --tune` chose for workload 'name' (None: defaults)."""
    if not USE_TUNING or not os.path.exists(TUNING_FILE):
        return None
    from segclip import tune
    return tune.load(TUNING_FILE, name) or None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary, or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith(kernel + " traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_threads():
    """Host threads for the CPU baseline: OMP_NUM_THREADS when set, else the cores this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(nb, nc):
    """The oracle (C restatement of the reference's multicore path, its OpenMP
    loops where the reference has them) on the same construction, timed on
    this host with every available core: tree build and the whole clip
    including labelling and traversal; input generation excluded."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import segclip
    import oracle_py as O
    b = segclip.synth.harmonic(nb, 2001)
    c = segclip.synth.harmonic(nc, 2002, 0.15, -0.1, 1.0, 17.0 * math.pi / 180.0)
    threads = cpu_threads()
    O.set_threads(threads)
    t0 = time.perf_counter()
    tr = O.Tree(b, [c])
    t1 = time.perf_counter()
    r = tr.clip(1, canonical=True)
    dt = time.perf_counter() - t0
    return {"value": (nb + nc) / dt, "unit": "edges/s", "cores": threads, "kind": "port",
            "sample": "oracle (C restatement of the reference multicore path, OpenMP where the reference is "
                      "parallel), harmonic star pair %d x %d (the workload itself), tree %.1f s + clip incl. "
                      "labelling %.1f s = %.1f s on %d threads, %d rings" % (nb, nc, t1 - t0, dt - (t1 - t0), dt,
                                                                            threads, len(r.rings))}


def timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def plain_loop_rate(bases, ovs, sample):
    """Pairs/s of the plain per-pair GPU path: every pair its own context, tree and clip."""
    import segclip
    idx = list(range(0, len(ovs), max(1, len(ovs) // sample)))[:sample]
    cl = segclip.Clipper()
    one = bases if isinstance(bases, list) else None
    cl.set_polygons(one[idx[0]] if one else bases, [ovs[idx[0]]]).clip(1)  # warm-up
    t0 = time.perf_counter()
    for i in idx:
        cl.set_polygons(one[i] if one else bases, [ovs[i]])
        cl.clip(1)
    dt = time.perf_counter() - t0
    return len(idx) / dt, len(idx)


def bench_one_to_many(steps):
    import segclip
    base, ovs = segclip.synth.one_to_many(100000, 4096, 3000)
    cl = segclip.Clipper().set_polygons(base, ovs)
    cl.set_timing(True)
    tuning = tuning_for("one_to_many")
    cl.apply_tuning(tuning or {})

    def step():
        cl.reset()
        cl.clip_batch(1, len(ovs))

    dt = timed(step, steps, 1)
    cnt = cl.counts()
    log("[bench] one_to_many stage ms %s counts %s" % (json.dumps(cl.stage_times()), json.dumps(cnt)))
    plain, n = plain_loop_rate(base, ovs, 48)
    edges = len(base) * len(ovs) + sum(len(o) for o in ovs)
    return {"workload": "one-to-many: harmonic base 100,000 vertices x 4,096 star overlays (8..20,000 vertices)",
            "pairs": len(ovs), "ms_per_batch": 1e3 * dt, "pairs_per_s": len(ovs) / dt,
            "edges_per_s": edges / dt, "plain_pairs_per_s": plain, "plain_sample": n,
            "speedup_vs_plain": (len(ovs) / dt) / plain, "rings": cnt["rings"], "out_vertices": cnt["out_vertices"],
            "tuning": tuning}


def bench_small_pairs(steps):
    import torch
    import segclip
    from segclip.autograd import clip_area_iou
    bases, ovs = segclip.synth.small_pairs(16384, 4000)
    cl = segclip.Clipper().set_pairs(bases, ovs)
    tuning = tuning_for("small_pairs")  # the forward batch (the autograd path keeps the defaults)
    cl.apply_tuning(tuning or {})
    J = len(bases)
    out = torch.empty((4, J), dtype=torch.float64, device="cuda")

    def fwd():
        cl.reset()
        cl.clip_batch()
        cl.batch_area_device(out.data_ptr())

    dt = timed(fwd, steps, 1)
    fused = cl.get_tuning("small_fallback") == 0
    import numpy as np
    bx = torch.tensor(np.concatenate(bases), dtype=torch.float64, device="cuda", requires_grad=True)
    ox = torch.tensor(np.concatenate(ovs), dtype=torch.float64, device="cuda", requires_grad=True)
    bsz, csz = [len(b) for b in bases], [len(o) for o in ovs]
    # the same forward replayed as one captured hipGraph on device-resident rings
    gcl = segclip.Clipper().set_pairs_device(bx.data_ptr(), bsz, ox.data_ptr(), csz)
    gout = torch.empty((4, J), dtype=torch.float64, device="cuda")
    dtg = timed(lambda: gcl.pairs_area_graph(bx.data_ptr(), ox.data_ptr(), gout.data_ptr()), steps, 1)
    # the staged path (one launch per stage for all pairs) on the same batch
    scl = segclip.Clipper().set_pairs(bases, ovs).set_tuning("small_fused", 0)
    scl.apply_tuning(tuning or {})

    def fwd_staged():
        scl.reset()
        scl.clip_batch()
        scl.batch_area_device(out.data_ptr())

    dts = timed(fwd_staged, steps, 1)

    def fwd_bwd():
        bx.grad = None
        ox.grad = None
        a, iou = clip_area_iou(bx, bsz, ox, csz)
        (a.sum() + iou.sum()).backward()

    dtb = timed(fwd_bwd, steps, 1)
    plain, n = plain_loop_rate(bases, ovs, 256)
    return {"workload": "16,384 pairs, 4..64 vertices (convex / star)", "pairs": J,
            "ms_per_batch": 1e3 * dt, "pairs_per_s": J / dt, "persistent_kernel": fused,
            "ms_per_batch_graph": 1e3 * dtg, "pairs_per_s_graph": J / dtg,
            "ms_per_batch_staged": 1e3 * dts, "pairs_per_s_staged": J / dts,
            "fwd_bwd_ms": 1e3 * dtb, "fwd_bwd_pairs_per_s": J / dtb,
            "plain_pairs_per_s": plain, "plain_sample": n, "speedup_vs_plain": (J / dt) / plain, "tuning": tuning}


def bench_large_pairs(steps):
    import numpy as np
    import torch
    import segclip
    from segclip.autograd import clip_area_iou
    t0 = time.perf_counter()
    bases, ovs = segclip.synth.large_pairs(256, 5000)
    log("[bench] large pairs generated in %.1f s" % (time.perf_counter() - t0))
    J = len(bases)
    bx = torch.tensor(np.concatenate(bases), dtype=torch.float64, device="cuda", requires_grad=True)
    ox = torch.tensor(np.concatenate(ovs), dtype=torch.float64, device="cuda", requires_grad=True)
    bsz, csz = [len(b) for b in bases], [len(o) for o in ovs]

    def fwd_bwd():
        bx.grad = None
        ox.grad = None
        a, iou = clip_area_iou(bx, bsz, ox, csz, recompute=True)
        (a.sum() + iou.sum()).backward()

    fwd_bwd()
    segclip.memory(reset_peak=True)
    dt = timed(fwd_bwd, steps, 0)
    cur, peak = segclip.memory()

    def fwd_bwd_stored():
        bx.grad = None
        ox.grad = None
        a, iou = clip_area_iou(bx, bsz, ox, csz, recompute=False)
        (a.sum() + iou.sum()).backward()

    fwd_bwd_stored()
    segclip.memory(reset_peak=True)
    dts = timed(fwd_bwd_stored, steps, 0)
    _, peak_stored = segclip.memory()
    # forward only, batched, for the same pairs
    cl = segclip.Clipper().set_pairs(bases, ovs)
    tuning = tuning_for("large_pairs")  # the forward batch (the autograd path keeps the defaults)
    cl.apply_tuning(tuning or {})

    def fwd():
        cl.reset()
        cl.clip_batch()

    dtf = timed(fwd, steps, 1)
    cl.set_timing(True)
    fwd()
    st = cl.stage_times()
    cnt = cl.counts()
    sb = stab_bytes(cnt["cover_entries"], cnt["leaf_end_entries"], cnt["candidates"])
    cl.release()
    plain, n = plain_loop_rate(bases, ovs, 8)
    edges = sum(bsz) + sum(csz)
    return {"workload": "256 pairs of harmonic stars, 10k..200k vertices", "pairs": J,
            "fwd_bwd_ms": 1e3 * dt, "fwd_bwd_pairs_per_s": J / dt, "fwd_bwd_edges_per_s": edges / dt,
            "peak_hbm_bytes_fwd_bwd": peak, "fwd_bwd_stored_ms": 1e3 * dts,
            "peak_hbm_bytes_fwd_bwd_stored": peak_stored,
            "stab_GBps": round(sb / (st["stab_kernel"] * 1e-3) / 1e9, 1) if st["stab_kernel"] > 0 else None,
            "stab_roofline_frac": round(sb / (st["stab_kernel"] * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
            if st["stab_kernel"] > 0 else None,
            "stage_ms": {k: round(v, 3) for k, v in st.items()},
            "fwd_ms": 1e3 * dtf, "fwd_pairs_per_s": J / dtf,
            "plain_fwd_pairs_per_s": plain, "plain_sample": n, "speedup_fwd_vs_plain": (J / dtf) / plain,
            "tuning": tuning}


def run_tuner(nb, nc):
    """Time every tuning candidate (segclip.tune.SETTINGS, one setting after the
    other) on the plain clip and the three batched
    forward workloads; write the fastest per workload to TUNING_FILE."""
    import torch
    import segclip
    from segclip import tune
    entries = {}

    def one(name, text, cl, step, reps):
        log("[tune] %s" % name)
        chosen, ms = tune.autotune_all(cl, lambda c: step(), warmup=1, reps=reps)
        entries[name] = {"settings": chosen, "workload": text, "timings_ms": ms}
        log("[tune] %s: %s" % (name, json.dumps(entries[name])))
        cl.close()

    base = segclip.synth.harmonic(nb, 2001)
    ov = segclip.synth.harmonic(nc, 2002, 0.15, -0.1, 1.0, 17.0 * math.pi / 180.0)
    cl = segclip.Clipper().set_polygons(base, [ov])
    one("plain", "plain pairwise clip, harmonic star rings %d x %d (seeds 2001/2002)" % (nb, nc), cl,
        lambda: (cl.reset(), cl.clip(1)), 5)
    base, ovs = segclip.synth.one_to_many(100000, 4096, 3000)
    cl = segclip.Clipper().set_polygons(base, ovs)
    one("one_to_many", "harmonic base 100,000 vertices x 4,096 star overlays", cl,
        lambda: (cl.reset(), cl.clip_batch(1, len(ovs))), 3)
    bases, ovs = segclip.synth.small_pairs(16384, 4000)
    cl = segclip.Clipper().set_pairs(bases, ovs)
    one("small_pairs", "16,384 pairs, 4..64 vertices, forward batch", cl, lambda: (cl.reset(), cl.clip_batch()), 5)
    bases, ovs = segclip.synth.large_pairs(256, 5000)
    cl = segclip.Clipper().set_pairs(bases, ovs)
    one("large_pairs", "256 pairs of harmonic stars, 10k..200k vertices, forward batch", cl,
        lambda: (cl.reset(), cl.clip_batch()), 3)
    return tune.save(TUNING_FILE, entries, torch.cuda.get_device_name(0))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nb", type=int, default=200000)
    ap.add_argument("--nc", type=int, default=1000000)
    ap.add_argument("--cpu-sample", default="200000x1000000", help="NBxNC of the oracle sample, or 'none'")
    ap.add_argument("--batched", default="one_to_many,small_pairs,large_pairs",
                    help="comma list of batched workloads to add, or 'none'")
    ap.add_argument("--only", default=None, metavar="WORKLOAD",
                    help="run only one batched workload (one_to_many, small_pairs, large_pairs) and print its figures")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (3 launches), to scale HBM counters")
    ap.add_argument("--tune", action="store_true",
                    help="time every candidate of each tuning setting on the plain clip and the batched workloads and "
                         "write the fastest per workload to "
                         "segment-tree-based-clipping_amd/tuning.json")
    ap.add_argument("--no-tuning", action="store_true", help="ignore tuning.json (library defaults)")
    args = ap.parse_args()
    global USE_TUNING
    USE_TUNING = not args.no_tuning
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.arange(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 8 * n, "bytes_written": 8 * n, "launches": 3}))
        return
    if args.only:
        fn = {"one_to_many": bench_one_to_many, "small_pairs": bench_small_pairs,
              "large_pairs": bench_large_pairs}[args.only]
        print(json.dumps({args.only: fn(args.steps)}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU (the box has one); --gpus %d is not supported"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.tune:
        print(json.dumps(run_tuner(args.nb, args.nc)), flush=True)
        return
    import segclip
    nb, nc = args.nb, args.nc
    log("[bench] generating harmonic rings %d x %d" % (nb, nc))
    base = segclip.synth.harmonic(nb, 2001)
    ov = segclip.synth.harmonic(nc, 2002, 0.15, -0.1, 1.0, 17.0 * math.pi / 180.0)
    cl = segclip.Clipper(device=0)
    cl.set_polygons(base, [ov])  # host -> HBM once, outside the timed region
    cl.set_timing(True)

    def step():
        cl.reset()
        cl.clip(1)

    tuning = tuning_for("plain")
    cl.apply_tuning(tuning or {})
    if tuning:
        log("[bench] tuning %s" % json.dumps(tuning))

    for i in range(args.warmup):
        step()
        log("[bench] warmup %d done" % i)
    stab_ms, stages = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        st = cl.stage_times()
        stages.append(st)
        stab_ms.append(st["stab_kernel"])
    t1 = time.perf_counter()
    ms = 1000.0 * (t1 - t0) / args.steps
    cnt = cl.counts()
    log("[bench] %d steps, %.3f ms/step, counts %s" % (args.steps, ms, json.dumps(cnt)))
    avg = {k: sum(s[k] for s in stages) / len(stages) for k in stages[0]}
    log("[bench] stage ms: %s" % json.dumps({k: round(v, 4) for k, v in avg.items()}))
    # E summed over every level (the reference's end-list layout), outside the timed region
    materialized_end = None
    try:
        cl.materialize_end_lists()
        materialized_end = cl.counts()["end_entries"]
        log("[bench] materialised end-list entries (all levels): %d" % materialized_end)
    except Exception as e:
        log("[bench] end-list materialisation skipped: %s" % e)
    cl.release()
    del cl

    stab = sum(stab_ms) / len(stab_ms)
    algo = stab_bytes(cnt["cover_entries"], cnt["leaf_end_entries"], cnt["candidates"])
    achieved = algo / (stab * 1e-3) / 1e9 if stab > 0 else 0.0
    traffic = measured_traffic("stab_stage") if (nb, nc) == (200000, 1000000) else None
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
            "kernel": "stab stage: k_stab + k_stab_rows + k_near_classify", "kernel_ms": round(stab, 4),
            "algorithmic_bytes": algo, "near_pairs": cnt.get("near_pairs"),
            "work": "36 (L + E) + 12 K bytes with L=%d cover entries, E=%d leaf end entries, K=%d candidates; "
                    "W=%d pair tests" % (cnt["cover_entries"], cnt["leaf_end_entries"], cnt["candidates"],
                                          cnt["tests"]),
            "traffic_note": "FETCH_SIZE + WRITE_SIZE of the three kernels per stage, scaled by a 1 GiB copy: "
                            "L2-to-fabric bytes, Infinity Cache hits included"}
    if materialized_end is not None:
        roof["materialized_end_entries"] = materialized_end

    batched = {}
    if args.batched != "none":
        for name in args.batched.split(","):
            fn = {"one_to_many": bench_one_to_many, "small_pairs": bench_small_pairs,
                  "large_pairs": bench_large_pairs}[name]
            log("[bench] batched workload %s" % name)
            try:
                batched[name] = fn(3)
                log("[bench] %s: %s" % (name, json.dumps(batched[name])))
            except Exception as e:  # reported, never hides the main line
                batched[name] = {"error": "%s: %s" % (type(e).__name__, e)}
                log("[bench] %s failed: %s" % (name, e))

    cpu = None
    if args.cpu_sample != "none":
        snb, snc = (int(v) for v in args.cpu_sample.split("x"))
        log("[bench] cpu baseline sample %d x %d" % (snb, snc))
        cpu = cpu_baseline(snb, snc)

    edges = nb + nc
    line = {
        "metric": METRIC, "value": edges / (ms * 1e-3), "unit": "edges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "plain pairwise clip, harmonic star rings %d x %d (seeds 2001/2002), "
                               "inputs resident in HBM" % (nb, nc),
                   "pairs_per_s": 1000.0 / ms, "intersections": cnt["overlay_pairs"], "rings": cnt["rings"],
                   "out_vertices": cnt["out_vertices"], "stage_ms": {k: round(v, 4) for k, v in avg.items()},
                   "tuning": tuning},
        "roofline": roof, "cpu_baseline": cpu, "batched": batched,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
