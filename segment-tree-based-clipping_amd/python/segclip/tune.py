"""Autotuner for a context's tuning settings.

A setting changes how the library splits its work, never its results: the
stab's "stab_tiles" (64-column tiles per work unit, one wave each) changes the
number and length of the waves, "stab_units_per_wave" how many consecutive
units one wave runs, and the candidate list is sorted afterwards,
so every value yields the same pairs, intersections and rings.  The tuner
therefore measures speed only; the tests check that results agree across
values.

Usage: time each candidate value on a workload (a function that runs one clip
on a Clipper), keep the fastest, and persist it as JSON:

    best, ms = tune.autotune(clipper, step)
    tune.save(path, {"plain": {"settings": {"stab_tiles": best}, "workload": text,
                               "timings_ms": {"stab_tiles": ms}}})
    clipper.apply_tuning(tune.load(path, "plain"))

The reference has no counterpart: its stab runs one tree node per
parallel-loop iteration (oacc_segment_tree.cpp:6250-6526) with nothing to
choose.
"""

import json
import math
import os
import time

# candidate values per setting, and the library's defaults
SETTINGS = {"stab_tiles": (1, 2, 4, 8), "stab_units_per_wave": (1, 2, 4, 8, 16)}
DEFAULTS = {"stab_tiles": 2, "stab_units_per_wave": 2}
LIMITS = {"stab_tiles": (1, 16), "stab_units_per_wave": (1, 32)}


def autotune(clipper, step, name="stab_tiles", candidates=None, warmup=1, reps=3):
    """Time step(clipper) under each candidate value of one setting.

    Returns (best value, {value: best-of-reps ms}) and leaves the clipper set
    to the best value.  step must run one whole workload and may return before
    the device is done; the device is synchronised around every timing.
    """
    import torch

    if name not in SETTINGS:
        raise ValueError("unknown tuning setting %r" % name)
    candidates = tuple(candidates or SETTINGS[name])
    lo, hi = LIMITS[name]
    for v in candidates:
        if not lo <= int(v) <= hi:
            raise ValueError("%s=%d is outside %d..%d" % (name, v, lo, hi))
    ms = {}
    for v in candidates:
        clipper.set_tuning(name, v)
        for _ in range(warmup):
            step(clipper)
        torch.cuda.synchronize()
        best = math.inf
        for _ in range(reps):
            t0 = time.perf_counter()
            step(clipper)
            torch.cuda.synchronize()
            best = min(best, 1000.0 * (time.perf_counter() - t0))
        ms[int(v)] = best
    choice = min(ms, key=ms.get)
    clipper.set_tuning(name, choice)
    return choice, ms


def autotune_all(clipper, step, warmup=1, reps=3):
    """Time every combination of the settings' candidate values (a grid: the
    settings interact) and keep the fastest.  Returns ({setting: value},
    {"grid": {"v1,v2,...": ms}}), values in SETTINGS order."""
    import itertools
    import torch

    names = list(SETTINGS)
    grid = {}
    for combo in itertools.product(*(SETTINGS[n] for n in names)):
        for n, v in zip(names, combo):
            clipper.set_tuning(n, v)
        for _ in range(warmup):
            step(clipper)
        torch.cuda.synchronize()
        best = math.inf
        for _ in range(reps):
            t0 = time.perf_counter()
            step(clipper)
            torch.cuda.synchronize()
            best = min(best, 1000.0 * (time.perf_counter() - t0))
        grid[",".join(str(v) for v in combo)] = best
    key = min(grid, key=grid.get)
    chosen = {n: int(v) for n, v in zip(names, key.split(","))}
    for n, v in chosen.items():
        clipper.set_tuning(n, v)
    return chosen, {"grid": grid}


def save(path, entries, device=None):
    """Persist chosen settings per named workload as JSON.

    entries maps a workload name to {"settings": {...}, "workload": text,
    "timings_ms": {setting: {value: ms}}}.
    """
    out = {}
    for name, e in entries.items():
        check(e["settings"])
        out[name] = {"settings": {k: int(v) for k, v in e["settings"].items()},
                     "workload": e.get("workload", ""),
                     "timings_ms": {k: {str(a): round(b, 4) for a, b in t.items()}
                                    for k, t in e.get("timings_ms", {}).items()}}
    rec = {"device": device, "workloads": out}
    with open(path, "w") as f:
        json.dump(rec, f, indent=1, sort_keys=True)
        f.write("\n")
    return rec


def load(path, name):
    """Settings of workload 'name' in a file written by save(), validated;
    {} (the library's defaults) when the file has no entry for it."""
    with open(path) as f:
        rec = json.load(f)
    e = rec.get("workloads", {}).get(name)
    settings = {k: int(v) for k, v in (e or {}).get("settings", {}).items()}
    check(settings)
    return settings


TUNING_FILE = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))),
                           "tuning.json")


def workload_for(paired, max_ring):
    """The tuned workload whose shape a batch has: many small pairs, many large
    pairs, or one base ring against many overlays."""
    if not paired:
        return "one_to_many"
    return "small_pairs" if max_ring <= 64 else "large_pairs"


def settings_for(paired, max_ring, path=None):
    """Settings persisted for the workload of this shape ({} when none)."""
    path = path or TUNING_FILE
    if not os.path.exists(path):
        return {}
    return load(path, workload_for(paired, max_ring))


def check(settings):
    for k, v in settings.items():
        if k not in LIMITS:
            raise ValueError("unknown tuning setting %r" % k)
        lo, hi = LIMITS[k]
        if not lo <= int(v) <= hi:
            raise ValueError("%s=%d is outside %d..%d" % (k, v, lo, hi))
