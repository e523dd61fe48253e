"""GPU parity of batched clipping (rows A16, A18): a batch of jobs run in the
same launches gives, job by job, exactly the rings of clipping each job alone,
and those equal the oracle's.

One-to-many batches share the base ring (the reference's one-to-many mode);
paired batches give every pair its own base ring.
"""
import numpy as np
import pytest

import oracle_py as O
import segclip.synth as S
from fixtures import degenerate_cases, mixed_overlays, rings_equal, status_agrees

pytestmark = pytest.mark.gpu


def test_one_to_many_batch_equals_per_overlay(gpu):
    import segclip
    b = S.harmonic(4000, 31)
    ovs = mixed_overlays()
    print("[gpu-batch] one-to-many", len(b), [len(o) for o in ovs], flush=True)
    tr = O.Tree(b, ovs)
    cl = segclip.Clipper().set_polygons(b, ovs)
    single, status = [], []
    for k in range(1, len(ovs) + 1):
        cl.clip(k)
        single.append(cl.rings())
        status.append(cl.counts()["undefined"])
        r = tr.clip(k, canonical=True)
        err = rings_equal(single[-1], r.rings)
        assert err is None, "overlay %d vs oracle: %s" % (k, err)
    cl.clip_batch(1, len(ovs))
    br = cl.batch_rings()
    st = cl.batch_status()
    assert len(br) == len(ovs)
    for k in range(len(ovs)):
        err = rings_equal(br[k], single[k])
        assert err is None, "overlay %d batch vs single: %s" % (k + 1, err)
        assert (st[k] & 1) == (status[k] & 1)
    # a sub-range
    cl.clip_batch(3, 7)
    br = cl.batch_rings()
    for k in range(3, 8):
        err = rings_equal(br[k - 3], single[k - 1])
        assert err is None, "overlay %d in batch 3..7: %s" % (k, err)


def _pairs():
    bases, ovs, names = [], [], []
    for name, b, o, _ in degenerate_cases():
        bases.append(b)
        ovs.append(o[0])
        names.append(name)
    sb, so = S.small_pairs(24, 4100)
    bases += sb
    ovs += so
    names += ["small%d" % i for i in range(len(sb))]
    bases.append(S.harmonic(3000, 77))
    ovs.append(S.star(2000, 78, 0.1, 0.05, 0.4, 0.9))
    names.append("harmonic3k_star2k")
    return bases, ovs, names


def test_pairs_batch_equals_each_pair_alone(gpu):
    import segclip
    bases, ovs, names = _pairs()
    print("[gpu-batch] pairs", len(bases), flush=True)
    cl = segclip.Clipper().set_pairs(bases, ovs)
    cl.clip_batch()
    br = cl.batch_rings()
    st = cl.batch_status()
    assert len(br) == len(bases)
    for i, (b, o, name) in enumerate(zip(bases, ovs, names)):
        one = segclip.Clipper().set_polygons(b, [o]).clip(1)
        err = rings_equal(br[i], one.rings())
        assert err is None, "%s: batch vs alone: %s" % (name, err)
        assert (st[i] & 1) == (one.counts()["undefined"] & 1)
        pb, pc, pt = cl.pairs(i + 1)
        qb, qc, qt = one.pairs(1)
        assert np.array_equal(pb, qb) and np.array_equal(pc, qc), name
        r = O.Tree(b, [o]).clip(1, canonical=True)
        err = rings_equal(br[i], r.rings)
        assert err is None, "%s: batch vs oracle: %s" % (name, err)


def _rect(x0, y0, w, h, t):
    c, s = np.cos(t), np.sin(t)
    pts = np.array([(0.0, 0.0), (w, 0.0), (w, h), (0.0, h)], dtype=np.float64)
    return pts @ np.array([[c, s], [-s, c]]) + (x0, y0)


def _far_pairs(count, seed):
    """Small rings at longitude/latitude-like places (stars 1e-4 to 1e-2
    across); every fourth pair is two rectangles turned by about 1e-9 rad,
    whose long edges span about 1e-12 in x: below the spacing of the x keys
    once a pair's keys are shifted by up to count * 361."""
    rng = np.random.default_rng(seed)
    bases, ovs = [], []
    for i in range(count):
        cx, cy = rng.uniform(-180.0, 180.0), rng.uniform(-85.0, 85.0)
        s = 10.0 ** rng.uniform(-4.0, -2.0)
        if i % 4 == 0:
            t = 1e-9 * (1.0 + rng.uniform())
            bases.append(_rect(cx, cy, 0.5 * s, s, t))
            ovs.append(_rect(cx + 0.25 * s, cy + 0.3 * s, 0.5 * s, 0.2 * s, t))
        else:
            bases.append(S.star(int(rng.integers(8, 40)), 7000 + i, cx, cy, 0.5 * s, s))
            ovs.append(S.star(int(rng.integers(8, 40)), 9000 + i, cx + 0.3 * s, cy + 0.2 * s, 0.5 * s, s))
    return bases, ovs


def test_pairs_batch_far_apart_keys(gpu):
    """1,024 pairs spread like longitude/latitude data on the staged path (one
    tree for all pairs, each pair's x keys shifted into its own window): every
    pair's pair set, rings and status equal clipping it alone, also for the
    edges whose shifted x keys coincide although their x coordinates differ;
    its status agrees with the oracle's, and its rings equal the oracle's
    wherever the reference's result is defined (every intersection got a
    slot; 60 of these pairs have one that did not)."""
    import segclip
    bases, ovs = _far_pairs(1024, 77)
    # the shifted keys of some edges coincide: the case this test is about
    allx = np.concatenate([r[:, 0] for r in bases + ovs])
    span = (allx.max() - allx.min()) + 1.0
    merged = 0
    for k, pair in enumerate(zip(bases, ovs)):
        sh = k * span - allx.min()
        for r in pair:
            x0, x1 = r[:, 0], np.roll(r[:, 0], -1)
            merged += int(np.sum((x0 != x1) & (x0 + sh == x1 + sh)))
    assert merged > 100
    print("[gpu-batch] far-apart pairs", len(bases), "merged keys", merged, flush=True)
    cl = segclip.Clipper().set_pairs(bases, ovs).set_tuning("small_fused", 0)
    cl.clip_batch()
    br = cl.batch_rings()
    st = cl.batch_status()
    one = segclip.Clipper()
    for i, (b, o) in enumerate(zip(bases, ovs)):
        one.set_polygons(b, [o]).clip(1)
        err = rings_equal(br[i], one.rings())
        assert err is None, "pair %d: batch vs alone: %s" % (i, err)
        assert (st[i] & 1) == (one.counts()["undefined"] & 1), i
        pb, pc, _ = cl.pairs(i + 1)  # a batch keeps no per-pair types
        qb, qc, _ = one.pairs(1)
        assert np.array_equal(pb, qb) and np.array_equal(pc, qc), i
        r = O.Tree(b, [o]).clip(1, canonical=True)
        assert status_agrees(int(st[i]), r.status), (i, int(st[i]), r.status)
        if r.status & 1:
            continue
        err = rings_equal(br[i], r.rings)
        assert err is None, "pair %d: batch vs oracle: %s" % (i, err)


def test_pooled_allocator_release_and_rebuild(gpu):
    """Buffers from torch's caching allocator: a batch, a release, the same
    batch again in the same context, all equal to the hipMalloc path."""
    import segclip
    bases, ovs, names = _pairs()
    print("[gpu-batch] pooled allocator", len(bases), flush=True)
    ref = segclip.Clipper().set_pairs(bases, ovs)
    ref.clip_batch()
    want = ref.batch_rings()
    ref.release()
    assert not segclip._torch_alloc_on
    with segclip.torch_allocator():
        assert segclip._torch_alloc_on
        cl = segclip.Clipper().set_pairs(bases, ovs)
        for rep in range(2):
            cl.clip_batch()
            got = cl.batch_rings()
            for i, name in enumerate(names):
                err = rings_equal(got[i], want[i])
                assert err is None, "%s (pass %d): %s" % (name, rep, err)
            cl.release()
    assert not segclip._torch_alloc_on
    cl.clip_batch()  # buffers of this pass come from hipMalloc again
    got = cl.batch_rings()
    for i, name in enumerate(names):
        assert rings_equal(got[i], want[i]) is None, name
    cl.close()


def test_small_pairs_batch_vs_oracle(gpu):
    import segclip
    bases, ovs = S.small_pairs(512, 4000)
    print("[gpu-batch] small pairs", len(bases), flush=True)
    cl = segclip.Clipper().set_pairs(bases, ovs)
    cl.clip_batch()
    br = cl.batch_rings()
    for i, (b, o) in enumerate(zip(bases, ovs)):
        r = O.Tree(b, [o]).clip(1, canonical=True)
        err = rings_equal(br[i], r.rings)
        assert err is None, "pair %d: %s" % (i, err)


def test_one_to_many_scaled_batch(gpu):
    import segclip
    b, ovs = S.one_to_many(20000, 128, 3000)
    print("[gpu-batch] one-to-many scaled", len(b), len(ovs), flush=True)
    cl = segclip.Clipper().set_polygons(b, ovs)
    single = []
    for k in range(1, len(ovs) + 1):
        cl.clip(k)
        single.append(cl.rings())
    cl.clip_batch()
    br = cl.batch_rings()
    for k in range(len(ovs)):
        err = rings_equal(br[k], single[k])
        assert err is None, "overlay %d: %s" % (k + 1, err)
    tr = O.Tree(b, ovs)
    for k in (1, 2, 3, 64, 128):
        r = tr.clip(k, canonical=True)
        err = rings_equal(br[k - 1], r.rings)
        assert err is None, "overlay %d vs oracle: %s" % (k, err)


def test_one_to_many_degenerate_batch(gpu):
    """Degenerate overlays of one base in one batch (sparse base ring): shared
    edges and identical rings take the sequential route, a containing overlay
    the no-crossing route; each equals the overlay clipped alone and the oracle."""
    import numpy as np
    import segclip
    from fixtures import square
    b = square(0.0, 0.0, 2.0, 2.0)
    tri = np.array([[0.5, 0.5], [1.5, 0.5], [1.0, 0.0]])
    ovs = [square(1.0, 0.5, 3.0, 1.5), square(1.0, 0.0, 3.0, 2.0), square(0.0, 0.0, 2.0, 2.0), tri,
           square(5.0, 5.0, 6.0, 6.0), square(-1.0, -1.0, 3.0, 3.0), square(0.5, 0.5, 1.5, 1.5),
           np.array([[2.0, 1.0], [3.0, -1.0], [4.0, 1.0], [3.0, 3.0]])]
    print("[gpu-batch] one-to-many degenerate", len(ovs), flush=True)
    tr = O.Tree(b, ovs)
    cl = segclip.Clipper().set_polygons(b, ovs)
    single, status = [], []
    for k in range(1, len(ovs) + 1):
        cl.clip(k)
        single.append(cl.rings())
        status.append(cl.counts()["undefined"])
        r = tr.clip(k, canonical=True)
        assert rings_equal(single[-1], r.rings) is None, k
    cl.clip_batch(1, len(ovs))
    br, st = cl.batch_rings(), cl.batch_status()
    for k in range(len(ovs)):
        err = rings_equal(br[k], single[k])
        assert err is None, "overlay %d: %s" % (k + 1, err)
        assert st[k] == status[k], (k + 1, st[k], status[k])
