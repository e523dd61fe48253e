"""Seeded synthetic rings for the clipping workloads (host code in libsegclip).

Workloads (all FP64, rings counter-clockwise, star-shaped, no repeated
closing vertex):
  star1k      base 1,000 vertices, r~U[0.5,1], centre (0,0), seed 1001;
              overlay 1,000, centre (0.2,0.1), seed 1002
  harmonic    base 200,000 harmonic star (32 harmonics, k^-1.2, jitter
              +-0.02), seed 2001; overlay 1,000,000, centre (0.15,-0.1),
              rotated 17 degrees, seed 2002
  one_to_many base 100,000 harmonic (seed 3000) and 4,096 random stars with
              log-uniform sizes [8,20000] and radii [0.005,0.3]
  small_pairs 16,384 pairs with 4..64 vertices, convex / star alternating
  large_pairs 256 pairs of harmonic stars with 10k..200k vertices
"""
import math

import numpy as np

from ._lib import as_ptr, lib

import ctypes as C


def star(n, seed, cx=0.0, cy=0.0, rmin=0.5, rmax=1.0, rot=0.0):
    a = np.zeros((n, 2), np.float64)
    rc = lib().segclip_synth_star(as_ptr(a, C.c_double), int(n), int(seed), cx, cy, rmin, rmax, rot)
    if rc != 0:
        raise ValueError("synth_star failed (%d)" % rc)
    return a


def harmonic(n, seed, cx=0.0, cy=0.0, scale=1.0, rot=0.0):
    a = np.zeros((n, 2), np.float64)
    rc = lib().segclip_synth_harmonic(as_ptr(a, C.c_double), int(n), int(seed), cx, cy, scale, rot)
    if rc != 0:
        raise ValueError("synth_harmonic failed (%d)" % rc)
    return a


def convex(n, seed, cx=0.0, cy=0.0, scale=1.0, rot=0.0):
    a = np.zeros((n, 2), np.float64)
    rc = lib().segclip_synth_convex(as_ptr(a, C.c_double), int(n), int(seed), cx, cy, scale, rot)
    if rc != 0:
        raise ValueError("synth_convex failed (%d)" % rc)
    return a


def uniform(n, seed):
    a = np.zeros(n, np.float64)
    lib().segclip_synth_uniform(as_ptr(a, C.c_double), int(n), int(seed))
    return a


def count_vertical(ring):
    ring = np.ascontiguousarray(ring, dtype=np.float64)
    return lib().segclip_count_vertical(as_ptr(ring, C.c_double), len(ring))


def star1k():
    return star(1000, 1001, 0.0, 0.0, 0.5, 1.0), [star(1000, 1002, 0.2, 0.1, 0.5, 1.0)]


def harmonic_pair(nb=200000, nc=1000000):
    b = harmonic(nb, 2001, 0.0, 0.0, 1.0, 0.0)
    c = harmonic(nc, 2002, 0.15, -0.1, 1.0, 17.0 * math.pi / 180.0)
    return b, [c]


def one_to_many(nb=100000, count=4096, seed=3000):
    base = harmonic(nb, seed, 0.0, 0.0, 1.0, 0.0)
    lo, hi = base.min(axis=0), base.max(axis=0)
    u = uniform(6 * count, seed + 77)
    ovs = []
    for i in range(count):
        n = int(round(math.exp(math.log(8) + u[6 * i] * (math.log(20000) - math.log(8)))))
        R = math.exp(math.log(0.005) + u[6 * i + 1] * (math.log(0.3) - math.log(0.005)))
        cx = lo[0] + u[6 * i + 2] * (hi[0] - lo[0])
        cy = lo[1] + u[6 * i + 3] * (hi[1] - lo[1])
        ovs.append(star(max(n, 3), seed + 1 + i, cx, cy, 0.5 * R, R, 0.0))
    return base, ovs
