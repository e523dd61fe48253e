"""GPU parity of the SegTree facade (include/segtree.h): the reference's
multicore call sequence, replayed by tests/facade/segtree_check through the
facade, yields the oracle's augmented rings, neighbour links, alpha values and
initial labels bit for bit (rows A11-A15 behind the reference's interface)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_py as O
import segclip.synth as S

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "facade", "segtree_check")


def _run(b, ovs):
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([len(b), len(ovs)] + [len(o) for o in ovs], np.int32).tofile(f)
            np.ascontiguousarray(b, np.float64).tofile(f)
            np.ascontiguousarray(np.concatenate(ovs), np.float64).tofile(f)
        r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        raw = open(fout, "rb").read()
    out, at = [], 0
    for _ in ovs:
        nba, nca = np.frombuffer(raw, np.int32, 2, at)
        at += 8
        xb = np.frombuffer(raw, np.float64, 2 * nba, at).reshape(-1, 2)
        at += 16 * nba
        xc = np.frombuffer(raw, np.float64, 2 * nca, at).reshape(-1, 2)
        at += 16 * nca
        arrs = []
        for n in (nba, nca, nba, nca, nba, nca):
            arrs.append(np.frombuffer(raw, np.int32, n, at))
            at += 4 * n
        out.append((xb, xc) + tuple(arrs))
    return out


@pytest.mark.parametrize("name", ["star1k", "harmonic3k_star2k", "one_to_many", "rotated_squares"])
def test_facade_matches_oracle(gpu, name):
    import math
    if name == "star1k":
        b, ovs = S.star1k()
    elif name == "harmonic3k_star2k":
        b, ovs = S.harmonic(3000, 77), [S.star(2000, 78, 0.1, 0.05, 0.4, 0.9)]
    elif name == "one_to_many":
        b = S.harmonic(4000, 31)
        ovs = [S.star(n, 40 + i, cx, cy, 0.1, 0.4) for i, (n, cx, cy) in
               enumerate([(50, 0.2, 0.1), (300, -0.5, 0.2), (12, 0.9, -0.3)])]
    else:
        th = math.radians(30.0)
        sq = np.array([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]])
        R = lambda a: np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])
        b, ovs = sq @ R(th).T, [sq @ R(-0.4).T + np.array([0.3, 0.2])]
    print("[gpu-facade]", name, flush=True)
    got = _run(b, ovs)
    tr = O.Tree(b, ovs)
    for k in range(1, len(ovs) + 1):
        r = tr.clip(k, canonical=True)
        xb, xc, nbb, nbc, ab, ac, lb, lc = got[k - 1]
        assert np.array_equal(xb.view(np.uint64), r.aug_b.view(np.uint64)), k
        assert np.array_equal(xc.view(np.uint64), r.aug_c.view(np.uint64)), k
        assert np.array_equal(nbb, r.nbr_b) and np.array_equal(nbc, r.nbr_c), k
        assert np.array_equal(ab, r.alphav_b) and np.array_equal(ac, r.alphav_c), k
        inter_b = r.alphav_b != -100
        inter_c = r.nbr_c != -100
        assert np.array_equal(lb[inter_b], r.label_b[inter_b]), k
        assert np.array_equal(lc[inter_c], r.label_c[inter_c]), k
