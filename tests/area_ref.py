"""FP64 CPU reference of the differentiable area / IoU, over the oracle's clip.

area(): shoelace over the oracle's output rings.  grad(): the gradient of the
intersection area with respect to every input vertex, found independently of
the GPU's slot provenance: each output vertex is matched by exact coordinates
to an input vertex or to a type-1 intersection point of the oracle's pair
list, whose gradient is formed here from the intersection formula by the
chain rule.  fd_grad(): central finite differences of the oracle's area.
"""
import numpy as np

import oracle_py as O


def shoelace2(ring):
    x, y = ring[:, 0], ring[:, 1]
    return float(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y))


def clip_rings(b, o):
    return O.Tree(b, [o]).clip(1, canonical=True)


def area(b, o):
    """Intersection area: the output rings' unsigned areas (an intersection of
    two simple rings has no holes; with a clockwise input the oracle's output
    can hold rings of both orientations)."""
    return 0.5 * sum(abs(shoelace2(r)) for r in clip_rings(b, o).rings)


def area_iou(b, o):
    i = area(b, o)
    p, q = 0.5 * abs(shoelace2(b)), 0.5 * abs(shoelace2(o))
    u = p + q - i
    return i, (i / u if u != 0 else 0.0)


def _A(p, q, r):
    return (q[0] - p[0]) * (r[1] - p[1]) - (q[1] - p[1]) * (r[0] - p[0])


def _dA(p, q, r):
    """Partial derivatives of _A w.r.t. p, q, r."""
    return (np.array([q[1] - r[1], r[0] - q[0]]), np.array([r[1] - p[1], p[0] - r[0]]),
            np.array([p[1] - q[1], q[0] - p[0]]))


def grad(b, o):
    """d(intersection area)/d(b), d/d(o) as (nb,2), (nc,2) arrays."""
    res = clip_rings(b, o)
    gb, go = np.zeros_like(b), np.zeros_like(o)
    vb = {(float(x), float(y)): i for i, (x, y) in enumerate(b)}
    vo = {(float(x), float(y)): i for i, (x, y) in enumerate(o)}
    ip = {}
    for p in range(len(res.pair_type)):
        if res.pair_type[p] == 1:
            ip[(float(res.intersections[p, 0]), float(res.intersections[p, 1]))] = p
    nb, nc = len(b), len(o)
    unmatched = 0
    for ring in res.rings:
        n = len(ring)
        sg = -1.0 if shoelace2(ring) < 0 else 1.0  # d|S| = sgn(S) dS
        for i in range(n):
            nx, pv = ring[(i + 1) % n], ring[(i - 1) % n]
            g = 0.5 * sg * np.array([nx[1] - pv[1], pv[0] - nx[0]])
            key = (float(ring[i, 0]), float(ring[i, 1]))
            if key in vb:
                gb[vb[key]] += g
            elif key in vo:
                go[vo[key]] += g
            elif key in ip:
                p = ip[key]
                bi, ci = int(res.pair_b[p]) // 2, int(res.pair_c[p]) // 2
                b1, b2 = b[bi], b[(bi + 1) % nb]
                c1, c2 = o[ci], o[(ci + 1) % nc]
                ap1, ap2 = _A(b1, c1, c2), _A(b2, c1, c2)
                d = ap1 - ap2
                al = ap1 / d
                gal = g @ (b2 - b1)
                g1, g2 = -gal * ap2 / d ** 2, gal * ap1 / d ** 2
                p1, q1, r1 = _dA(b1, c1, c2)
                p2, q2, r2 = _dA(b2, c1, c2)
                gb[bi] += (1 - al) * g + g1 * p1
                gb[(bi + 1) % nb] += al * g + g2 * p2
                go[ci] += g1 * q1 + g2 * q2
                go[(ci + 1) % nc] += g1 * r1 + g2 * r2
            else:
                unmatched += 1
    assert unmatched == 0, "%d output vertices matched no input vertex or crossing" % unmatched
    return gb, go


def fd_grad(b, o, h=1e-7):
    gb, go = np.zeros_like(b), np.zeros_like(o)
    for arr, g in ((b, gb), (o, go)):
        for i in range(len(arr)):
            for k in range(2):
                keep = arr[i, k]
                arr[i, k] = keep + h
                ap = area(b, o)
                arr[i, k] = keep - h
                am = area(b, o)
                arr[i, k] = keep
                g[i, k] = (ap - am) / (2 * h)
    return gb, go


def min_vertex_edge_distance(b, o):
    """Smallest distance from a vertex of one ring to an edge of the other."""
    def one(P, Q):
        a = Q
        c = np.roll(Q, -1, axis=0)
        d = c - a
        L2 = np.maximum(np.sum(d * d, axis=1), 1e-300)
        best = np.inf
        for p in P:
            t = np.clip(np.sum((p - a) * d, axis=1) / L2, 0.0, 1.0)
            q = a + t[:, None] * d
            best = min(best, float(np.min(np.hypot(*(p - q).T))))
        return best
    return min(one(b, o), one(o, b))


def min_consecutive_area(ring):
    """Smallest |A(v[i-1], v[i], v[i+1])| of a ring: the clean-up removes a vertex
    below 1e-9, so the area jumps where a triple crosses that tolerance."""
    p, q, n = np.roll(ring, 1, 0), ring, np.roll(ring, -1, 0)
    a = (q[:, 0] - p[:, 0]) * (n[:, 1] - p[:, 1]) - (q[:, 1] - p[:, 1]) * (n[:, 0] - p[:, 0])
    return float(np.min(np.abs(a)))


def fd_safe(b, o):
    """Pairs where central differences with h = 1e-7 see a smooth area: every
    vertex farther than 1e-4 from the other ring's edges and no consecutive
    triple near the clean-up tolerance."""
    return (min_vertex_edge_distance(b, o) > 1e-4 and min_consecutive_area(b) > 1e-6
            and min_consecutive_area(o) > 1e-6)
