"""GPU parity of the tree and candidate stages (rows A3-A15) against the oracle.

Bit-exact comparison: elementary points, node intervals, cover lists (as sets
per node: the order inside a node is unspecified, as in the reference's CPU
path), leaf end-list targets, materialised end lists in (id,type) order, the
unique pair set with types, intersection points, quantised alphas, augmented
rings, neighbour links and initial labels.
"""
import numpy as np
import pytest

import oracle_py as O
from fixtures import degenerate_cases, workloads_small

pytestmark = pytest.mark.gpu


def _bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def _comb(teeth):
    """A ring whose zigzag puts `teeth` vertices at x = 1 and as many at x = 2
    (runs of equal endpoint values longer than the library's short walk),
    crossed by a tilted rectangle."""
    pts = [(0.0, 0.0), (4.0, 0.0), (4.0, float(teeth) + 1.0)]
    for i in range(teeth, 0, -1):
        pts.append((2.0, float(i) + 0.5))
        pts.append((1.0, float(i)))
    base = np.array(pts, np.float64)
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    rect = np.array([[-0.5, -0.2], [3.5, -0.2], [3.5, 0.6], [-0.5, 0.6]]) @ R.T + np.array([0.2, 2.1])
    return base, rect


def _cases():
    out = [(n, b, o) for n, b, o in workloads_small()]
    out += [(n, b, o) for n, b, o, _ in degenerate_cases()]
    b, r = _comb(24)
    out.append(("comb24", b, [r]))
    return out


CASES = _cases()


@pytest.fixture(scope="module")
def built():
    import segclip
    res = {}
    for name, b, ovs in CASES:
        print("[gpu-tree] case", name, len(b), [len(o) for o in ovs], flush=True)
        cl = segclip.Clipper()
        cl.set_polygons(b, ovs).build_tree().materialize_end_lists().find_candidates()
        res[name] = (cl, O.Tree(b, ovs))
    return res


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_elementary_and_intervals(gpu, built, name):
    cl, tr = built[name]
    assert _bits_equal(cl.elementary(), tr.elementary())
    g, o = cl.node_intervals(), tr.node_intervals()
    assert _bits_equal(g[2:], o[2:])


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_cover_lists(gpu, built, name):
    cl, tr = built[name]
    goff, ge = cl.cover_lists()
    ooff, oe = tr.cover()
    assert np.array_equal(goff, ooff)
    for v in np.nonzero(np.diff(ooff))[0]:
        s, t = ooff[v], ooff[v + 1]
        assert np.array_equal(np.sort(ge[s:t]), np.sort(oe[s:t])), "node %d" % v


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_end_lists(gpu, built, name):
    cl, tr = built[name]
    ga, gb = cl.leaf_targets()
    oa, ob = tr.leaf_targets()
    assert np.array_equal(ga, oa) and np.array_equal(gb, ob)
    goff, ge = cl.end_lists()
    ooff, oe = tr.end_lists()
    assert np.array_equal(goff[1:], ooff[1:])
    assert np.array_equal(ge, oe)


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_pairs_and_intersections(gpu, built, name):
    cl, tr = built[name]
    for k in range(1, tr.K + 1):
        r = tr.clip(k, canonical=True)
        cl.intersections(k)
        b, c, t = cl.pairs(k)
        assert np.array_equal(b, r.pair_b) and np.array_equal(c, r.pair_c)
        assert np.array_equal(t, r.pair_type)
        xy, qa, qb = cl.intersection_points()
        assert _bits_equal(xy.reshape(-1), r.intersections.reshape(-1))
        assert np.array_equal(qa, r.alpha_b) and np.array_equal(qb, r.alpha_c)
        cnt = cl.counts()
        assert (cnt["undefined"] & 17) == (r.status & 17)
        if r.status & 1:
            continue
        for side, (axy, nbr, av, lab) in enumerate([(r.aug_b, r.nbr_b, r.alphav_b, r.label_b),
                                                     (r.aug_c, r.nbr_c, r.alphav_c, r.label_c)]):
            gxy, gn, ga, gl = cl.augmented(side)
            assert _bits_equal(gxy.reshape(-1), axy.reshape(-1)), "side %d coords" % side
            assert np.array_equal(gn, nbr), "side %d neighbours" % side
            assert np.array_equal(ga, av), "side %d alpha" % side
            assert np.array_equal(gl, lab), "side %d labels" % side
