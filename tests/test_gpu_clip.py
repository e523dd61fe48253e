"""GPU parity of the whole clip (rows A8-A17): output rings bit-exact to the oracle.

Covers the parallel traversal (generic crossings) and the sequential device
path (intersection chains, ON_ON rings, bouncing touches); the oracle itself
is pinned to the reference's Foster clipper in test_oracle.py.
"""
import numpy as np
import pytest

import oracle_py as O
from fixtures import degenerate_cases, rings_equal as _rings_equal, workloads_small

pytestmark = pytest.mark.gpu


def _cases():
    out = [(n, b, o) for n, b, o in workloads_small()]
    out += [(n, b, o) for n, b, o, _ in degenerate_cases()]
    return out


CASES = _cases()


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_clip_rings_bit_exact(gpu, case):
    import segclip
    name, b, ovs = case
    print("[gpu-clip]", name, len(b), [len(o) for o in ovs], flush=True)
    tr = O.Tree(b, ovs)
    cl = segclip.Clipper()
    cl.set_polygons(b, ovs)
    for k in range(1, len(ovs) + 1):
        r = tr.clip(k, canonical=True)
        cl.clip(k)
        cnt = cl.counts()
        assert (cnt["undefined"] & 1) == (r.status & 1)
        err = _rings_equal(cl.rings(), r.rings)
        assert err is None, err
        assert cnt["rings"] == len(r.rings)


def test_one_to_many_matches_per_overlay_oracle(gpu):
    import segclip
    b = segclip.synth.harmonic(4000, 31)
    ovs = [segclip.synth.star(n, 40 + i, cx, cy, 0.1, 0.4) for i, (n, cx, cy) in
           enumerate([(50, 0.2, 0.1), (300, -0.5, 0.2), (1000, 0.0, 0.0), (12, 0.9, -0.3), (700, -0.2, -0.8)])]
    tr = O.Tree(b, ovs)
    cl = segclip.Clipper().set_polygons(b, ovs)
    for k in range(1, len(ovs) + 1):
        r = tr.clip(k, canonical=True)
        cl.clip(k)
        pb, pc, pt = cl.pairs(k)
        assert np.array_equal(pb, r.pair_b) and np.array_equal(pc, r.pair_c) and np.array_equal(pt, r.pair_type)
        err = _rings_equal(cl.rings(), r.rings)
        assert err is None, "overlay %d: %s" % (k, err)


def _sawtooth(T):
    """Counter-clockwise sawtooth: a spine at y=-2 and T teeth reaching y=1, valleys at y=-1."""
    pts = [(0.0, -2.0), (1.0, -2.0)]
    for i in range(T, 0, -1):
        pts.append(((i - 0.5) / T, 1.0))
        pts.append(((i - 1.0) / T, -1.0))
    return np.array(pts, dtype=np.float64)


@pytest.mark.parametrize("T", [32, 33, 40, 64, 65, 128, 129, 256, 257, 2048, 2500])
@pytest.mark.parametrize("swap", [False, True])
def test_many_intersections_on_one_edge(gpu, T, swap):
    """2T intersections on each horizontal edge of a rectangle, on the base side
    and (swapped) on the overlay side: per-edge groups sorted by one wave with
    1, 2, 4 or 8 keys per lane (64 up to 512 pairs, each side of every
    boundary), by one block (514 and 4096 pairs) and beyond the LDS limit
    (5000 pairs)."""
    import segclip
    rect = np.array([(-0.5, 0.0), (1.5, 0.0), (1.5, 0.5), (-0.5, 0.5)], dtype=np.float64)
    saw = _sawtooth(T)
    b, ov = (saw, rect) if swap else (rect, saw)
    print("[gpu-clip] sawtooth", T, "swap", swap, flush=True)
    tr = O.Tree(b, [ov])
    r = tr.clip(1, canonical=True)
    cl = segclip.Clipper().set_polygons(b, [ov])
    cl.clip(1)
    pb, pc, pt = cl.pairs(1)
    assert len(pb) == 4 * T
    assert np.array_equal(pb, r.pair_b) and np.array_equal(pc, r.pair_c) and np.array_equal(pt, r.pair_type)
    err = _rings_equal(cl.rings(), r.rings)
    assert err is None, err
    assert cl.counts()["rings"] == len(r.rings) == T
