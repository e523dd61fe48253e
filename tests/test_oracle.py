"""CPU checks of the oracle itself (no GPU).

What pins each oracle stage (the reference has no tests or golden files):
  * output rings vs the reference's own sequential Foster clipper
    (oracle/_ref/polyclip, built from lib/optimizedFostersAlgorithm): same
    component count, same vertex counts, same cyclic order, coordinates equal
    to its 6-significant-digit print;
  * pair set vs brute force over all base x overlay edges;
  * elementary points / node intervals / cover and end lists vs their closed
    forms computed here independently in numpy;
  * special cases with known answers (overlapping squares, containment,
    disjoint rings);
  * committed golden vectors (tests/golden) for the predicates.
"""
import json
import os
import tempfile

import numpy as np
import pytest

import oracle_py as O
from fixtures import degenerate_cases, workloads_small

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _small():
    return [w for w in workloads_small() if w[0] != "harmonic20k_100k"]


@pytest.mark.parametrize("case", _small() + [(n, b, o) for n, b, o, _ in degenerate_cases()], ids=lambda c: c[0])
def test_pairs_equal_brute_force(case):
    name, b, ovs = case
    tr = O.Tree(b, ovs)
    r = tr.clip(1, canonical=True)
    bb, bc, bt = tr.brute_pairs(1)
    assert set(zip(r.pair_b, r.pair_c, r.pair_type)) == set(zip(bb, bc, bt))


@pytest.mark.parametrize("case", _small(), ids=lambda c: c[0])
def test_literal_and_canonical_orders_agree(case):
    name, b, ovs = case
    tr = O.Tree(b, ovs)
    lit, can = tr.clip(1, canonical=False), tr.clip(1, canonical=True)
    assert lit.status == 0 and can.status == 0
    assert len(lit.rings) == len(can.rings)
    for x, y in zip(lit.rings, can.rings):
        assert np.array_equal(x, y)


def _fmt6(v):
    return float("%g" % v)


@pytest.mark.skipif(not os.path.exists(O.REF_POLYCLIP), reason="reference Foster clipper not built (no reference tree)")
@pytest.mark.parametrize("case", _small() + [(n, b, o) for n, b, o, _ in degenerate_cases()], ids=lambda c: c[0])
def test_rings_match_reference_foster(case):
    name, b, ovs = case
    r = O.Tree(b, ovs).clip(1, canonical=True)
    with tempfile.TemporaryDirectory() as d:
        ref = O.run_reference_foster(b, ovs[0], d)
    assert len(ref) == len(r.rings)
    for mine, theirs in zip(r.rings, ref):
        assert mine.shape == theirs.shape
        printed = np.vectorize(_fmt6)(mine)
        assert np.array_equal(printed, theirs)


@pytest.mark.skipif(not os.path.exists(O.REF_POLYCLIP17), reason="reference Foster clipper not built (no reference tree)")
@pytest.mark.parametrize("case", _small() + [(n, b, o) for n, b, o, _ in degenerate_cases()], ids=lambda c: c[0])
def test_rings_match_reference_foster_full_precision(case):
    """Rings against the reference's own clipper printed at 17 digits: same
    components and vertices in the same cyclic order, coordinates within 1e-12
    relative (the two programs compute intersection points by the same
    formula but not necessarily with the same rounding)."""
    name, b, ovs = case
    r = O.Tree(b, ovs).clip(1, canonical=True)
    with tempfile.TemporaryDirectory() as d:
        ref = O.run_reference_foster(b, ovs[0], d, full_precision=True)
    assert len(ref) == len(r.rings)
    for mine, theirs in zip(r.rings, ref):
        assert mine.shape == theirs.shape
        tol = 1e-12 * np.maximum(1.0, np.abs(theirs))
        assert np.all(np.abs(mine - theirs) <= tol), np.max(np.abs(mine - theirs))


@pytest.mark.parametrize("case", _small() + [(n, b, o) for n, b, o, _ in degenerate_cases()], ids=lambda c: c[0])
def test_intersection_points_closed_form(case):
    """A12's points against an independent solution: an X crossing (type 1)
    is the solution of the 2x2 system b1 + a (b2 - b1) = c1 + t (c2 - c1)
    (numpy.linalg.solve, within 1e-12 relative); every other type lies on an
    input vertex exactly: c1 for types 3, 4, 5, 7, 8 and b1 for 2 and 6
    (saveIntersectionsParallel, oacc_segment_tree.cpp:7257-7341)."""
    name, b, ovs = case
    b = np.asarray(b, dtype=np.float64).reshape(-1, 2)
    c = np.asarray(ovs[0], dtype=np.float64).reshape(-1, 2)
    r = O.Tree(b, ovs).clip(1, canonical=True)
    n, m = len(b), len(c)
    for eid, fid, t, p in zip(r.pair_b, r.pair_c, r.pair_type, r.intersections):
        e, f = eid // 2, fid // 2  # edge ids are 2 x the first vertex's offset (Edge::id)
        b1, b2 = b[e], b[(e + 1) % n]
        c1, c2 = c[f], c[(f + 1) % m]
        if t == 1:
            A = np.array([[b2[0] - b1[0], -(c2[0] - c1[0])], [b2[1] - b1[1], -(c2[1] - c1[1])]])
            a, _ = np.linalg.solve(A, c1 - b1)
            x = b1 + a * (b2 - b1)
            scale = max(1.0, float(np.max(np.abs(np.concatenate([b1, b2, c1, c2])))))
            assert np.all(np.abs(p - x) <= 1e-12 * scale), (name, e, f, p, x)
        elif t in (3, 4, 5, 7, 8):
            assert p[0] == c1[0] and p[1] == c1[1], (name, e, f, t)
        else:
            assert t in (2, 6), t
            assert p[0] == b1[0] and p[1] == b1[1], (name, e, f, t)


def test_elementary_closed_form():
    b, ovs = workloads_small()[0][1:]
    tr = O.Tree(b, ovs)
    e = tr.edges()
    pts = np.concatenate([e["start"], e["end"]])
    u = np.unique(pts)
    # last-of-run rule of the reference drops the largest value (no vertical edges here)
    kept = u[:-1]
    n = 1 << int(np.ceil(np.log2(len(kept))))
    want = np.concatenate([kept, np.full(n + 1 - len(kept), kept[-1])])
    assert np.array_equal(tr.elementary(), want)
    T = tr.tree_size
    niv = tr.node_intervals().reshape(-1, 2)
    D = int(np.log2(T))
    for v in range(1, T):
        d = int(np.floor(np.log2(v)))
        sh = D - 1 - d
        first = (v - (1 << d)) << sh
        assert niv[v, 0] == want[first] and niv[v, 1] == want[first + (1 << sh)]


def test_cover_and_end_lists_closed_form():
    b, ovs = workloads_small()[2][1:]
    tr = O.Tree(b, ovs)
    e = tr.edges()
    ea = tr.elementary()
    T = tr.tree_size
    half = T // 2
    niv = tr.node_intervals().reshape(-1, 2)
    off, ce = tr.cover()
    # cover: Int(v) inside [s,e] and Int(parent) not inside
    for idx in range(0, len(e), 7):
        s, t = e["start"][idx], e["end"][idx]
        inside = (niv[:, 0] >= s) & (niv[:, 1] <= t)
        inside[0] = False
        par = np.arange(T) // 2
        want = np.nonzero(inside & ~np.where(np.arange(T) > 1, inside[par], False))[0]
        got = [v for v in range(1, T) if idx in set(ce[off[v]:off[v + 1]])]
        assert list(want) == got
    # end lists: sorted set of edges whose after/before leaf lies under v
    after, before = tr.leaf_targets()
    eo, ee = tr.end_lists()
    key = e["id"].astype(np.int64) * 4096 + e["type"]
    D = int(np.log2(T))
    for v in list(range(1, 64)) + list(range(half, half + 64)):
        d = int(np.floor(np.log2(v)))
        sh = D - 1 - d
        lo = half + ((v - (1 << d)) << sh)
        hi = lo + (1 << sh)
        members = np.nonzero(((after >= lo) & (after < hi)) | ((before >= lo) & (before < hi)))[0]
        want = members[np.argsort(key[members], kind="stable")]
        assert np.array_equal(ee[eo[v]:eo[v + 1]], want)


@pytest.mark.parametrize("case", degenerate_cases(), ids=lambda c: c[0])
def test_special_case_areas(case):
    name, b, ovs, area = case
    r = O.Tree(b, ovs).clip(1, canonical=True)
    if area is not None:
        assert abs(abs(r.area()) - area) < 1e-12
    if name == "contained":
        assert len(r.rings) == 1 and len(r.rings[0]) == len(ovs[0])


def test_predicate_golden_vectors():
    with open(os.path.join(GOLDEN, "predicates.json")) as f:
        cases = json.load(f)
    for c in cases:
        t, a, b = O.classify(np.array(c["p"]))
        assert t == c["type"], c
        assert O.lsmf(np.array(c["p"])) == c["lsmf"]
        if "alpha" in c:
            assert a == c["alpha"] and b == c["beta"]


@pytest.mark.skipif(not os.path.exists(O.REF_POLYCLIP), reason="reference Foster clipper not built (no reference tree)")
@pytest.mark.parametrize("case", _small() + [(n, b, o) for n, b, o, _ in degenerate_cases()], ids=lambda c: c[0])
def test_union_rings_match_reference_foster(case):
    """The oracle's union (the UNION switch of foster/polyclip.h:39) equals the
    reference's sequential clipper run with -union: components, vertex counts,
    order and the 6-digit coordinates it prints."""
    name, b, ovs = case
    r = O.Tree(b, ovs).clip(1, canonical=True, union=True)
    with tempfile.TemporaryDirectory() as d:
        ref = O.run_reference_foster(b, ovs[0], d, union=True)
    assert len(ref) == len(r.rings)
    for mine, theirs in zip(r.rings, ref):
        assert mine.shape == theirs.shape
        assert np.array_equal(np.vectorize(_fmt6)(mine), theirs)
    # inclusion-exclusion: |P u Q| = |P| + |Q| - |P n Q|.  Two overlapping simple
    # rings have a connected union: one outer ring (the largest) and holes, all
    # emitted with the same orientation, so its area is 2 max|A| - sum|A|; with
    # an empty intersection the union is both rings.  The interior touch of
    # touch_inside gives no ring at all: its bouncing vertex counts as a crossing
    # for the union (polygon.h:471), so no special case adds a ring, and no
    # CROSSING vertex starts a traversal (the reference's own result).
    if name != "touch_inside":
        def area(x):
            return abs(0.5 * float(np.sum(x[:, 0] * np.roll(x[:, 1], -1) - np.roll(x[:, 0], -1) * x[:, 1])))
        a = [area(x) for x in r.rings]
        inter = sum(area(x) for x in O.Tree(b, ovs).clip(1, canonical=True).rings)
        want = area(b) + area(ovs[0]) - inter
        got = (2 * max(a) - sum(a)) if inter > 0 else sum(a)
        assert got == pytest.approx(want, rel=1e-9, abs=1e-12)
