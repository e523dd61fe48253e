"""The per-job golden summaries of the batched workloads (tests/golden/*.npz,
written by tests/golden/make_batch_goldens.py) recomputed by the oracle on a
sample of jobs, including jobs with status bits: the fixtures the GPU tests
read are the oracle's results on these inputs."""
import os

import numpy as np
import pytest

import oracle_py as O
import segclip.synth as S
from fixtures import rings_area, rings_digest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _check(name, base_of, ovs, jobs):
    z = np.load(os.path.join(GOLDEN, name))
    assert len(z["status"]) == len(ovs)
    for j in jobs:
        r = O.Tree(base_of(j), [ovs[j]]).clip(1, canonical=True)
        assert r.status == z["status"][j], j
        assert len(r.rings) == z["rings"][j] and int(r.ring_off[-1]) == z["verts"][j], j
        assert rings_digest(r.rings) == int(z["digest"][j]), j
        assert rings_area(r.rings) == z["area"][j], j


def test_small_pairs_golden():
    bases, ovs = S.small_pairs(16384, 4000)
    _check("small_pairs_16384.npz", lambda j: bases[j], ovs, range(0, 16384, 256))


def test_one_to_many_golden():
    base, ovs = S.one_to_many(100000, 4096, 3000)
    z = np.load(os.path.join(GOLDEN, "one_to_many_100k_x4096.npz"))
    flagged = np.nonzero(z["status"])[0]
    assert set(np.unique(z["status"][flagged]).tolist()) >= {16, 17, 20}
    # one job of each status value, and a few plain ones
    pick = [int(flagged[np.nonzero(z["status"][flagged] == v)[0][0]]) for v in (16, 17, 20)] + [0, 1, 2048]
    _check("one_to_many_100k_x4096.npz", lambda j: base, ovs, pick)


def test_large_pairs_golden():
    bases, ovs = S.large_pairs(256, 5000)
    order = np.argsort([len(b) + len(o) for b, o in zip(bases, ovs)], kind="stable")
    _check("large_pairs_256.npz", lambda j: bases[j], ovs, [int(order[0]), int(order[1])])
