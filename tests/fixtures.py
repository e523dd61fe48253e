"""Input rings shared by the tests: seeded workloads and crafted degenerate cases."""
import math

import numpy as np

import segclip.synth as S


def square(x0, y0, x1, y1):
    return np.array([[x0, y0], [x1, y0], [x1, y1], [x0, y1]], np.float64)


def workloads_small():
    """(name, base, overlays) for sizes the oracle handles in about a second."""
    out = []
    b, o = S.star1k()
    out.append(("star1k", b, o))
    b = S.harmonic(20000, 2001)
    c = S.harmonic(100000, 2002, 0.15, -0.1, 1.0, 17.0 * math.pi / 180.0)
    out.append(("harmonic20k_100k", b, [c]))
    b = S.harmonic(3000, 77)
    c = S.star(2000, 78, 0.1, 0.05, 0.4, 0.9)
    out.append(("harmonic3k_star2k", b, [c]))
    return out


def degenerate_cases():
    """Crafted inputs with the special cases whose answers are known in closed form."""
    cases = []
    # overlapping axis-aligned squares (vertical edges, X intersections): area = overlap product
    cases.append(("squares", square(0.0, 0.0, 2.0, 2.0), [square(1.0, 0.5, 3.0, 1.5)], 1.0 * 1.0))
    # containment: overlay strictly inside the base -> output = overlay ring
    b = S.star(500, 11, 0.0, 0.0, 2.0, 3.0)
    c = S.star(300, 12, 0.0, 0.0, 0.5, 1.0)
    cases.append(("contained", b, [c], None))
    # disjoint: no output
    cases.append(("disjoint", square(0.0, 0.0, 1.0, 1.0), [square(5.0, 5.0, 6.0, 6.0)], 0.0))
    # rotated squares crossing (no vertical edges)
    th = math.radians(30.0)
    R = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
    sq = square(-1.0, -1.0, 1.0, 1.0) @ R.T
    sq2 = square(-1.0, -1.0, 1.0, 1.0) @ np.array([[math.cos(-0.4), -math.sin(-0.4)],
                                                     [math.sin(-0.4), math.cos(-0.4)]]).T + np.array([0.3, 0.2])
    cases.append(("rotated_squares", sq, [sq2], None))
    # identical rings: every vertex a V-overlap (ON_ON labels), output = the base ring
    r = S.star(200, 21, 0.0, 0.0, 0.5, 1.0)
    cases.append(("identical", r, [r.copy()], None))
    # shared collinear edge pieces: squares sliding along x (overlap chains), area 2
    cases.append(("shared_edges", square(0.0, 0.0, 2.0, 2.0), [square(1.0, 0.0, 3.0, 2.0)], 2.0))
    # overlay touching the base boundary from inside at one vertex (bouncing)
    tri = np.array([[1.0, 1.0], [3.0, 1.0], [2.0, 0.0]])
    cases.append(("touch_inside", square(0.0, 0.0, 4.0, 4.0), [tri], 1.0))
    # overlay vertex on a base edge, overlay crossing the base (T intersection)
    q = np.array([[2.0, 1.0], [3.0, -1.0], [4.0, 1.0], [3.0, 3.0]])
    cases.append(("vertex_on_edge", square(0.0, 0.0, 3.0, 2.0), [q], None))
    return cases


def status_agrees(gpu_status, oracle_status):
    """A job's status against the oracle's.  Where an intersection got no slot
    (bit 1) the reference's later steps are undefined, so only bits 1 and 16
    must agree; otherwise the statuses are equal."""
    if oracle_status & 1:
        return (gpu_status & 17) == (oracle_status & 17)
    return gpu_status == oracle_status


def rings_equal(g, o):
    """None when the two ring lists are bit-identical, else what differs."""
    if len(g) != len(o):
        return "ring count %d != %d" % (len(g), len(o))
    for i, (a, b) in enumerate(zip(g, o)):
        if a.shape != b.shape:
            return "ring %d: %d vs %d vertices" % (i, len(a), len(b))
        if not np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64)):
            return "ring %d: coordinates differ (max %g)" % (i, np.max(np.abs(a - b)))
    return None


def mixed_overlays(seed=40):
    """Overlays of a unit-scale base: crossing, inside, outside, enclosing, tiny."""
    out = []
    for i, (n, cx, cy, r0, r1) in enumerate([(50, 0.2, 0.1, 0.1, 0.4), (300, -0.5, 0.2, 0.1, 0.4),
                                               (1000, 0.0, 0.0, 0.6, 1.4), (12, 0.9, -0.3, 0.1, 0.4),
                                               (700, -0.2, -0.8, 0.1, 0.4), (40, 0.0, 0.0, 0.05, 0.1),
                                               (25, 6.0, 6.0, 0.5, 1.0), (200, 0.0, 0.0, 3.0, 4.0),
                                               (8, 0.3, 0.3, 0.005, 0.01), (5000, 0.1, -0.1, 0.3, 1.2)]):
        out.append(S.star(n, seed + i, cx, cy, r0, r1))
    return out


def rings_digest(rings):
    """64-bit digest of a job's output rings: the ring lengths (int32) then the
    coordinates (float64 bits), in ring order (blake2b, 8 bytes)."""
    import hashlib
    h = hashlib.blake2b(digest_size=8)
    h.update(np.array([len(r) for r in rings], np.int32).tobytes())
    for r in rings:
        h.update(np.ascontiguousarray(r, np.float64).tobytes())
    return int.from_bytes(h.digest(), "little")


def rings_area(rings):
    """Unsigned area of a job's output: sum of the rings' |shoelace| / 2."""
    a = 0.0
    for r in rings:
        x, y = r[:, 0], r[:, 1]
        a += abs(float(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)))
    return 0.5 * a
