"""GPU parity at the workloads' full sizes.

* Plain 200,000 x 1,000,000 harmonic clip: elementary points, node intervals,
  cover sets per node, leaf end targets, materialised end lists, the pair set
  with types, intersection points and quantised alphas, augmented rings,
  neighbour links, initial labels and output rings, all bit-exact against the
  oracle run live (its OpenMP loops give the single-thread results).
* One-to-many 100,000 x 4,096: the whole batch against per-job oracle
  summaries (tests/golden/one_to_many_100k_x4096.npz: status bits, ring and
  vertex counts, a digest of the rings, area), and the jobs that take the
  sequential route or are flagged undefined against the live oracle.
* 16,384 small pairs and 256 large pairs: the whole batch against the per-pair
  summaries; the 8 smallest large pairs bit-exact against the live oracle;
  recompute-in-backward equal to the stored path (areas bit for bit,
  gradients to summation order); a directional central finite-difference
  check of the gradient (h = 1e-7, rel 1e-6).

Status bits where the reference is undefined: bit 1 (an intersection never
gets a slot), 4 (a traversal does not close), 16 (a quantised alpha beyond the
int range).  On a job with bit 1 the oracle stops before labelling while the
GPU goes on, so only bits 1 and 16 are compared there and the rings are not.
"""
import os

import numpy as np
import pytest

import oracle_py as O
import segclip.synth as S
from fixtures import rings_area, rings_digest, rings_equal, status_agrees

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def _threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return 8


def test_harmonic_200k_1m_bit_exact(gpu):
    import segclip
    b, ovs = S.harmonic_pair()
    print("[gpu-full] harmonic 200k x 1M: oracle", flush=True)
    O.set_threads(_threads())
    tr = O.Tree(b, ovs)
    r = tr.clip(1, canonical=True)
    print("[gpu-full] harmonic 200k x 1M: gpu", flush=True)
    cl = segclip.Clipper().set_polygons(b, ovs)
    cl.build_tree().materialize_end_lists().find_candidates()
    assert _bits_equal(cl.elementary(), tr.elementary())
    assert _bits_equal(cl.node_intervals()[2:], tr.node_intervals()[2:])
    # cover lists: the same set per node (order inside a node is unspecified)
    goff, ge = cl.cover_lists()
    ooff, oe = tr.cover()
    assert np.array_equal(goff, ooff)
    node = np.repeat(np.arange(len(ooff) - 1), np.diff(ooff))
    assert np.array_equal(ge[np.lexsort((ge, node))], oe[np.lexsort((oe, node))])
    ga, gb = cl.leaf_targets()
    oa, ob = tr.leaf_targets()
    assert np.array_equal(ga, oa) and np.array_equal(gb, ob)
    goff, ge = cl.end_lists()
    ooff, oe = tr.end_lists()
    assert np.array_equal(goff[1:], ooff[1:]) and np.array_equal(ge, oe)
    assert cl.counts()["end_entries"] == len(oe)
    # pairs, intersections, augmented rings, labels
    cl.intersections(1)
    pb, pc, pt = cl.pairs(1)
    assert len(pb) == len(r.pair_b) > 10000
    assert np.array_equal(pb, r.pair_b) and np.array_equal(pc, r.pair_c) and np.array_equal(pt, r.pair_type)
    xy, qa, qb = cl.intersection_points()
    assert _bits_equal(xy.reshape(-1), r.intersections.reshape(-1))
    assert np.array_equal(qa, r.alpha_b) and np.array_equal(qb, r.alpha_c)
    for side, (axy, nbr, av, lab) in enumerate([(r.aug_b, r.nbr_b, r.alphav_b, r.label_b),
                                                 (r.aug_c, r.nbr_c, r.alphav_c, r.label_c)]):
        gxy, gn, gav, gl = cl.augmented(side)
        assert _bits_equal(gxy.reshape(-1), axy.reshape(-1)), side
        assert np.array_equal(gn, nbr) and np.array_equal(gav, av) and np.array_equal(gl, lab), side
    cl.label_traverse()
    assert cl.counts()["undefined"] == r.status == 0
    err = rings_equal(cl.rings(), r.rings)
    assert err is None, err


def _golden(name):
    z = np.load(os.path.join(GOLDEN, name))
    return {k: z[k] for k in z.files}


def _check_batch(cl, g, J, areas=None):
    rings = cl.batch_rings()
    st = cl.batch_status()
    assert len(rings) == len(st) == J
    bad = []
    for j in range(J):
        ok = status_agrees(int(st[j]), int(g["status"][j]))
        if ok and not (g["status"][j] & 1):
            ok = (len(rings[j]) == g["rings"][j] and sum(len(x) for x in rings[j]) == g["verts"][j]
                  and rings_digest(rings[j]) == int(g["digest"][j]))
        if ok and areas is not None and not (g["status"][j] & 1):
            ok = areas[j] == pytest.approx(float(g["area"][j]), rel=1e-12, abs=1e-14)
        if not ok:
            bad.append((j, int(st[j]), int(g["status"][j]), len(rings[j]), int(g["rings"][j])))
    assert not bad, "%d jobs differ from the oracle, first: %s" % (len(bad), bad[:5])


def test_one_to_many_full_batch(gpu):
    import segclip
    base, ovs = S.one_to_many(100000, 4096, 3000)
    g = _golden("one_to_many_100k_x4096.npz")
    print("[gpu-full] one-to-many 100k x 4096", flush=True)
    cl = segclip.Clipper().set_polygons(base, ovs)
    cl.clip_batch(1, len(ovs))
    _check_batch(cl, g, len(ovs))
    # every job the reference leaves undefined is flagged, and some are
    assert np.count_nonzero(g["status"]) >= 1


def test_one_to_many_flagged_jobs_vs_oracle(gpu):
    """The overlays that take the sequential route or carry status bits,
    clipped as one batch against the same base, equal the live oracle."""
    import segclip
    base, ovs = S.one_to_many(100000, 4096, 3000)
    g = _golden("one_to_many_100k_x4096.npz")
    cl = segclip.Clipper().set_polygons(base, ovs)
    cl.clip_batch(1, len(ovs))
    route, _ = cl.batch_routes()
    pick = sorted(set(np.nonzero(route == 2)[0].tolist()) | set(np.nonzero(g["status"])[0].tolist()))
    assert len(pick) >= 8
    sub = [ovs[i] for i in pick]
    print("[gpu-full] one-to-many flagged jobs", pick, flush=True)
    cl2 = segclip.Clipper().set_polygons(base, sub)
    cl2.clip_batch(1, len(sub))
    rings, st = cl2.batch_rings(), cl2.batch_status()
    O.set_threads(_threads())
    for q, i in enumerate(pick):
        r = O.Tree(base, [ovs[i]]).clip(1, canonical=True)
        assert status_agrees(int(st[q]), r.status), (i, int(st[q]), r.status)
        if r.status & 1:
            continue
        err = rings_equal(rings[q], r.rings)
        assert err is None, "overlay %d: %s" % (i, err)


def test_one_to_many_sequential_walks_agree(gpu):
    """The sequential route's rings walked one per lane (the default) equal
    the reference's single walk (seq_parallel_walks = 0) on every job of the
    one-to-many batch that takes that route, failures included."""
    import segclip
    base, ovs = S.one_to_many(100000, 4096, 3000)
    cl = segclip.Clipper().set_polygons(base, ovs)
    cl.clip_batch(1, len(ovs))
    route, _ = cl.batch_routes()
    pick = np.nonzero(route == 2)[0].tolist()
    assert len(pick) >= 8
    sub = [ovs[i] for i in pick]
    got = {}
    for par in (1, 0):
        print("[gpu-full] sequential route, seq_parallel_walks=%d, %d jobs" % (par, len(sub)), flush=True)
        c = segclip.Clipper().set_polygons(base, sub)
        c.set_tuning("seq_parallel_walks", par)
        assert c.get_tuning("seq_parallel_walks") == par
        c.clip_batch(1, len(sub))
        got[par] = (c.batch_rings(), c.batch_status())
    assert np.array_equal(got[0][1], got[1][1])
    assert np.count_nonzero(got[0][1] & 4) >= 1  # some jobs end in a failed walk
    for q in range(len(sub)):
        err = rings_equal(got[1][0][q], got[0][0][q])
        assert err is None, "job %d: %s" % (pick[q], err)
        for a_, b_ in zip(got[1][0][q], got[0][0][q]):
            assert np.array_equal(a_.view(np.uint64), b_.view(np.uint64)), pick[q]


def test_small_pairs_full_batch(gpu):
    import segclip
    bases, ovs = S.small_pairs(16384, 4000)
    g = _golden("small_pairs_16384.npz")
    print("[gpu-full] small pairs 16384", flush=True)
    cl = segclip.Clipper().set_pairs(bases, ovs).clip_batch()
    _check_batch(cl, g, len(bases), areas=cl.batch_area()[0])


@pytest.fixture(scope="module")
def large_pairs():
    print("[gpu-full] generating 256 large pairs", flush=True)
    return S.large_pairs(256, 5000)


def test_large_pairs_full_batch(gpu, large_pairs):
    import segclip
    bases, ovs = large_pairs
    g = _golden("large_pairs_256.npz")
    print("[gpu-full] large pairs 256", flush=True)
    cl = segclip.Clipper().set_pairs(bases, ovs).clip_batch()
    _check_batch(cl, g, len(bases), areas=cl.batch_area()[0])


def _smallest(large_pairs, n):
    bases, ovs = large_pairs
    order = np.argsort([len(b) + len(o) for b, o in zip(bases, ovs)], kind="stable")[:n]
    return [bases[i] for i in order], [ovs[i] for i in order]


def test_large_pairs_smallest_vs_oracle(gpu, large_pairs):
    import segclip
    bases, ovs = _smallest(large_pairs, 8)
    print("[gpu-full] 8 smallest large pairs vs oracle", [len(b) + len(o) for b, o in zip(bases, ovs)], flush=True)
    cl = segclip.Clipper().set_pairs(bases, ovs).clip_batch()
    rings, st = cl.batch_rings(), cl.batch_status()
    area = cl.batch_area()[0]
    O.set_threads(_threads())
    for j, (b, o) in enumerate(zip(bases, ovs)):
        r = O.Tree(b, [o]).clip(1, canonical=True)
        assert int(st[j]) == r.status == 0
        err = rings_equal(rings[j], r.rings)
        assert err is None, "pair %d: %s" % (j, err)
        assert area[j] == pytest.approx(rings_area(r.rings), rel=1e-12)


def test_large_pairs_recompute_and_fd(gpu, large_pairs):
    """recompute=True gives the stored path's areas and gradients bit for bit;
    the gradient's directional derivative equals central differences of the
    GPU's own areas on two pairs away from degeneracies."""
    import torch
    import segclip
    from segclip.autograd import clip_area_iou
    bases, ovs = _smallest(large_pairs, 8)
    bsz, csz = [len(b) for b in bases], [len(o) for o in ovs]
    w = np.linspace(0.5, 1.5, len(bases))
    out = []
    for rec in (False, True):
        print("[gpu-full] autograd recompute=%s" % rec, flush=True)
        bx = torch.tensor(np.concatenate(bases), dtype=torch.float64, device="cuda", requires_grad=True)
        ox = torch.tensor(np.concatenate(ovs), dtype=torch.float64, device="cuda", requires_grad=True)
        a, iou = clip_area_iou(bx, bsz, ox, csz, recompute=rec)
        (a * torch.tensor(w, dtype=torch.float64, device="cuda") + iou).sum().backward()
        out.append((a.detach().cpu().numpy(), iou.detach().cpu().numpy(), bx.grad.cpu().numpy(),
                    ox.grad.cpu().numpy()))
    # areas and IoU bit for bit; gradients to FP64 summation order (contributions
    # meet at input vertices through atomic adds, whose order varies)
    for x, y in zip(out[0][:2], out[1][:2]):
        assert _bits_equal(x.reshape(-1), y.reshape(-1))
    for x, y in zip(out[0][2:], out[1][2:]):
        assert np.max(np.abs(x - y)) <= 1e-12 * max(1.0, np.max(np.abs(x)))
    # directional central differences on two pairs where the area is smooth
    # along the direction: rings of 10k-200k vertices have consecutive triples
    # near the clean-up tolerance (|A| < 1e-9 removes a vertex), and a step
    # that flips such a decision is a jump, not a slope.  A pair qualifies when
    # the differences at h and h/2 agree (a jump within h makes them differ).
    rng = np.random.default_rng(5)
    gb = np.split(out[0][2], np.cumsum(bsz)[:-1])
    gc = np.split(out[0][3], np.cumsum(csz)[:-1])
    checked = 0
    for j in range(len(bases)):
        db, dc = rng.standard_normal(bases[j].shape), rng.standard_normal(ovs[j].shape)

        def loss(t):
            c2 = segclip.Clipper().set_pairs([bases[j] + t * db], [ovs[j] + t * dc]).clip_batch()
            x = c2.batch_area()
            return w[j] * x[0, 0] + x[3, 0]

        h = 1e-7
        fd = (loss(h) - loss(-h)) / (2 * h)
        fd2 = (loss(0.5 * h) - loss(-0.5 * h)) / h
        an = float(np.sum(gb[j] * db) + np.sum(gc[j] * dc))
        smooth = abs(fd - fd2) <= 1e-7 * max(1.0, abs(fd))
        print("[gpu-full] pair %d directional derivative %.12g fd %.12g (h/2: %.12g) smooth %s"
              % (j, an, fd, fd2, smooth), flush=True)
        if not smooth:
            continue
        assert abs(an - fd) <= 1e-6 * max(1.0, abs(fd)), (j, an, fd)
        checked += 1
        if checked == 2:
            break
    assert checked == 2
