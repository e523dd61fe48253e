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


def _normal_pairs(u):
    """Box-Muller on pairs of uniforms in (0,1]: returns len(u)//2 standard normals."""
    u1 = np.maximum(u[0::2], 1e-300)
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def small_pairs(count=16384, seed=4000):
    """Pairs with 4..64 vertices per ring: even pairs convex (sorted angles,
    radius 1), odd pairs star (radii U[0.4,1.0]); the overlay is rotated by
    U[0,2pi) and displaced by N(0, 0.3^2) per coordinate."""
    u = uniform(8 * count, seed)
    bases, ovs = [], []
    for i in range(count):
        nb = 4 + min(60, int(u[8 * i] * 61))
        nc = 4 + min(60, int(u[8 * i + 1] * 61))
        rot = 2.0 * math.pi * u[8 * i + 2]
        d = _normal_pairs(u[8 * i + 3:8 * i + 7])
        dx, dy = 0.3 * d[0], 0.3 * d[1]
        sb, sc = seed + 1 + 2 * i, seed + 2 + 2 * i
        if i % 2 == 0:
            bases.append(convex(nb, sb, 0.0, 0.0, 1.0, 0.0))
            ovs.append(convex(nc, sc, dx, dy, 1.0, rot))
        else:
            bases.append(star(nb, sb, 0.0, 0.0, 0.4, 1.0, 0.0))
            ovs.append(star(nc, sc, dx, dy, 0.4, 1.0, rot))
    return bases, ovs


def large_pairs(count=256, seed=5000):
    """Pairs of harmonic stars with log-uniform 10k..200k vertices; the overlay
    is offset by U[-0.3,0.3]^2 and rotated by U[0,2pi)."""
    u = uniform(5 * count, seed)
    bases, ovs = [], []
    lo, hi = math.log(10000), math.log(200000)
    for i in range(count):
        nb = int(round(math.exp(lo + u[5 * i] * (hi - lo))))
        nc = int(round(math.exp(lo + u[5 * i + 1] * (hi - lo))))
        dx, dy = -0.3 + 0.6 * u[5 * i + 2], -0.3 + 0.6 * u[5 * i + 3]
        rot = 2.0 * math.pi * u[5 * i + 4]
        bases.append(harmonic(nb, seed + 1 + 2 * i, 0.0, 0.0, 1.0, 0.0))
        ovs.append(harmonic(nc, seed + 2 + 2 * i, dx, dy, 1.0, rot))
    return bases, ovs
