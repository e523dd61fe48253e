"""GPU parity of the persistent small-pair kernel (batched clipping of
independent pairs whose rings have at most 64 vertices, one launch).

Its rings, statuses, routes and areas equal the staged path's bit for bit
(tuning setting small_fused = 0) and the oracle's; a batch holding a pair it
does not cover (degenerate fixtures, a ring beyond 64 vertices) gives the
staged path's results as well; the 16,384-pair workload runs without
falling back.
"""
import numpy as np
import pytest

import oracle_py as O
import segclip.synth as S
from fixtures import degenerate_cases, rings_equal

pytestmark = pytest.mark.gpu


def _run(bases, ovs, fused):
    import segclip
    cl = segclip.Clipper().set_pairs(bases, ovs)
    cl.set_tuning("small_fused", 1 if fused else 0)
    cl.clip_batch()
    return cl, cl.batch_rings(), cl.batch_status(), cl.batch_area()


def _same(a, b):
    (_, ra, sa, aa), (_, rb, sb, ab) = a, b
    assert np.array_equal(sa, sb)
    assert np.array_equal(aa.view(np.uint64), ab.view(np.uint64))
    for j, (x, y) in enumerate(zip(ra, rb)):
        err = rings_equal(x, y)
        assert err is None, "pair %d: %s" % (j, err)


def test_fused_equals_staged_and_oracle(gpu):
    bases, ovs = S.small_pairs(2048, 4700)
    print("[gpu-small] 2048 pairs fused vs staged", flush=True)
    f = _run(bases, ovs, True)
    assert f[0].get_tuning("small_fallback") == 0
    s = _run(bases, ovs, False)
    _same(f, s)
    r_f, _ = f[0].batch_routes()
    r_s, _ = s[0].batch_routes()
    assert np.array_equal(r_f, r_s)
    for j in range(0, len(bases), 64):
        r = O.Tree(bases[j], [ovs[j]]).clip(1, canonical=True)
        err = rings_equal(f[1][j], r.rings)
        assert err is None, "pair %d vs oracle: %s" % (j, err)


def test_batches_the_kernel_does_not_cover(gpu):
    # degenerate fixtures (chains, touching vertices, shared edges) among small pairs
    bases, ovs = S.small_pairs(64, 4800)
    for name, b, o, _ in degenerate_cases():
        bases.append(b)
        ovs.append(o[0])
    print("[gpu-small] degenerate fixtures in a small batch", flush=True)
    f = _run(bases, ovs, True)
    s = _run(bases, ovs, False)
    _same(f, s)
    for j, (b, o) in enumerate(zip(bases, ovs)):
        r = O.Tree(b, [o]).clip(1, canonical=True)
        if r.status & 1:
            continue
        err = rings_equal(f[1][j], r.rings)
        assert err is None, "pair %d vs oracle: %s" % (j, err)
    # a ring beyond 64 vertices: the batch is not eligible, results as staged
    bases2, ovs2 = S.small_pairs(32, 4900)
    bases2.append(S.star(200, 9, 0.0, 0.0, 0.5, 1.0))
    ovs2.append(S.star(150, 10, 0.2, 0.1, 0.5, 1.0))
    _same(_run(bases2, ovs2, True), _run(bases2, ovs2, False))


def test_workload_runs_fused(gpu):
    import segclip
    bases, ovs = S.small_pairs(16384, 4000)
    cl = segclip.Clipper().set_pairs(bases, ovs)
    cl.clip_batch()
    assert cl.get_tuning("small_fallback") == 0
    cl.clip_batch()  # a second batch on the same context gives the same rings
    a = cl.batch_area()
    cl.clip_batch()
    assert np.array_equal(cl.batch_area().view(np.uint64), a.view(np.uint64))
