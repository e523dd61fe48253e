#!/usr/bin/env python3
"""Benchmark of the segment-tree clipping path on one MI355X.

Workload (default): plain pairwise clip of two harmonic star rings, base
200,000 vertices (seed 2001), overlay 1,000,000 vertices (seed 2002, centre
(0.15,-0.1), rotated 17 degrees), FP64.  One step = the whole hot path on the
device with both rings already resident in HBM: edge records, elementary
intervals, node intervals, end lists, cover lists, stab + candidate dedupe,
intersections / augmented rings / initial labels, labelling, traversal and
clean-up of the output rings.  Reported value: input edges clipped per second
(N_b + N_c per step).

roofline: the stab kernel (the dominant kernel), FP64 work = W pair tests x 8
FP64 comparisons (the LSMF bounding-box test every pair runs) per launch,
divided by its launch time from HIP events on the library's stream, against
the 78.6 TFLOP/s FP64 peak.

cpu_baseline: the CPU oracle (single-thread restatement of the reference's
multicore path) on a bounded sample of the same construction, timed on this
host.
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
FP64_PEAK_TFLOPS = 78.6
FLOPS_PER_TEST = 8.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(nb, nc):
    """Oracle timing on a bounded sample (same construction, smaller rings)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import segclip
    import oracle_py as O
    b = segclip.synth.harmonic(nb, 2001)
    c = segclip.synth.harmonic(nc, 2002, 0.15, -0.1, 1.0, 17.0 * math.pi / 180.0)
    t0 = time.perf_counter()
    tr = O.Tree(b, [c])
    r = tr.clip(1, canonical=True)
    dt = time.perf_counter() - t0
    return {"value": (nb + nc) / dt, "unit": "edges/s", "cores": 1, "kind": "port",
            "sample": "oracle (single-thread C restatement of the reference multicore path), harmonic star "
                      "pair %d x %d (same construction as the workload), full clip incl. labelling, %.1f s, "
                      "%d rings" % (nb, nc, dt, len(r.rings))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nb", type=int, default=200000)
    ap.add_argument("--nc", type=int, default=1000000)
    ap.add_argument("--cpu-sample", default="100000x500000", help="NBxNC of the oracle sample, or 'none'")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU (the box has one); --gpus %d is not supported"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

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

    stab = sum(stab_ms) / len(stab_ms)
    flops = FLOPS_PER_TEST * cnt["tests"]
    achieved = flops / (stab * 1e-3) / 1e12 if stab > 0 else 0.0
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": None,
            "kernel": "k_stab", "kernel_ms": round(stab, 4),
            "work": "W=%d pair tests x %d FP64 ops (LSMF box test), FP64 VALU" % (cnt["tests"], FLOPS_PER_TEST)}

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
                   "out_vertices": cnt["out_vertices"], "stage_ms": {k: round(v, 4) for k, v in avg.items()}},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
