"""GPU checks of many-to-many clipping of two layers (row F2): the bounding-box
join equals brute force over all (a, b) with the reference's closed MBR test
(getMBR / getCMBR, polyclip_time.h:95-124), boxes that only touch included;
every candidate pair clipped in one batch equals the oracle on that pair."""
import numpy as np
import pytest

import oracle_py as O
import segclip.synth as S
from fixtures import rings_equal

pytestmark = pytest.mark.gpu


def _layer(n, seed, span, rmax):
    u = S.uniform(4 * n, seed)
    out = []
    for i in range(n):
        R = 0.02 + rmax * u[4 * i]
        nv = 3 + int(u[4 * i + 1] * 60)
        out.append(S.star(nv, seed + 1 + i, span * u[4 * i + 2], span * u[4 * i + 3], 0.5 * R, R))
    return out


def _brute(a, b):
    ma = np.array([[r[:, 0].min(), r[:, 1].min(), r[:, 0].max(), r[:, 1].max()] for r in a])
    mb = np.array([[r[:, 0].min(), r[:, 1].min(), r[:, 0].max(), r[:, 1].max()] for r in b])
    ov = ~((ma[:, None, 0] > mb[None, :, 2]) | (ma[:, None, 2] < mb[None, :, 0]) |
           (ma[:, None, 1] > mb[None, :, 3]) | (ma[:, None, 3] < mb[None, :, 1]))
    return np.argwhere(ov).astype(np.int32)


def test_mbr_join_equals_brute_force(gpu):
    from segclip.layers import mbr_join
    a = _layer(700, 61000, 10.0, 0.6)
    b = _layer(900, 62000, 10.0, 0.4)
    # boxes that touch along an edge and at a corner, and a duplicate box
    sq = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    a += [sq + 20.0, sq + 30.0]
    b += [sq + [21.0, 20.0], sq + 31.0, sq + 20.0]
    print("[gpu-layers] join", len(a), len(b), flush=True)
    got = mbr_join(a, b)
    want = _brute(a, b)
    assert len(want) > 1000
    assert np.array_equal(got, want)
    # an empty join
    far = [r + 1000.0 for r in b[:5]]
    assert len(mbr_join(a[:5], far)) == 0


def test_clip_layers_equals_oracle(gpu):
    from segclip.layers import clip_layers
    a = _layer(150, 63000, 5.0, 0.7)
    b = _layer(200, 64000, 5.0, 0.5)
    print("[gpu-layers] clip", len(a), len(b), flush=True)
    pairs, cl = clip_layers(a, b)
    assert len(pairs) > 100
    rings = cl.batch_rings()
    st = cl.batch_status()
    for q in range(0, len(pairs), 7):
        ia, ib = pairs[q]
        r = O.Tree(a[ia], [b[ib]]).clip(1, canonical=True)
        assert int(st[q]) == r.status
        err = rings_equal(rings[q], r.rings)
        assert err is None, (q, err)
