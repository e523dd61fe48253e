"""GPU parity of the union (SEGCLIP_OP_UNION, the UNION switch of
foster/polyclip.h:39): output rings bit-exact to the oracle's union (itself
pinned to the reference's clipper run with -union in test_oracle.py) for the
small workloads, the degenerate fixtures, a paired batch and a one-to-many
batch; bin/seg -union writes the same rings; the area entry points refuse a
union (its output may hold holes)."""
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O
import segclip.io as IO
import segclip.synth as S
from fixtures import degenerate_cases, rings_equal, workloads_small

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SEG = os.path.join(ROOT, "segment-tree-based-clipping_amd", "bin", "seg")


def _cases():
    return [(n, b, o) for n, b, o in workloads_small()] + [(n, b, o) for n, b, o, _ in degenerate_cases()]


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_union_rings_bit_exact(gpu, case):
    import segclip
    name, b, ovs = case
    print("[gpu-union]", name, flush=True)
    tr = O.Tree(b, ovs)
    cl = segclip.Clipper().set_polygons(b, ovs).set_operation(union=True)
    for k in range(1, len(ovs) + 1):
        r = tr.clip(k, canonical=True, union=True)
        cl.clip(k)
        assert (cl.counts()["undefined"] & 21) == (r.status & 21)
        if r.status & 1:
            continue
        err = rings_equal(cl.rings(), r.rings)
        assert err is None, err


def test_union_batches(gpu):
    import segclip
    from segclip._lib import SegclipError, ESTATE
    bases, ovs = S.small_pairs(256, 5300)
    print("[gpu-union] paired batch", flush=True)
    cl = segclip.Clipper().set_pairs(bases, ovs).set_operation(union=True).clip_batch()
    rings, st = cl.batch_rings(), cl.batch_status()
    for j, (b, o) in enumerate(zip(bases, ovs)):
        r = O.Tree(b, [o]).clip(1, canonical=True, union=True)
        assert int(st[j]) == r.status, j
        if r.status == 0:
            assert rings_equal(rings[j], r.rings) is None, j
    with pytest.raises(SegclipError) as e:
        cl.batch_area()
    assert e.value.code == ESTATE
    # back to the intersection on the same context
    cl.set_operation(union=False).clip_batch()
    r = O.Tree(bases[0], [ovs[0]]).clip(1, canonical=True)
    assert rings_equal(cl.batch_rings()[0], r.rings) is None
    # one-to-many
    b = S.harmonic(4000, 31)
    ovs2 = [S.star(n, 90 + i, cx, cy, 0.1, 0.4) for i, (n, cx, cy) in
            enumerate([(50, 0.2, 0.1), (300, -0.5, 0.2), (12, 0.9, -0.3), (25, 6.0, 6.0)])]
    print("[gpu-union] one-to-many batch", flush=True)
    cl2 = segclip.Clipper().set_polygons(b, ovs2).set_operation(union=True).clip_batch(1, len(ovs2))
    tr = O.Tree(b, ovs2)
    for k, rr in enumerate(cl2.batch_rings()):
        r = tr.clip(k + 1, canonical=True, union=True)
        assert rings_equal(rr, r.rings) is None, k


def test_seg_cli_union(gpu, tmp_path):
    b, ovs = S.star1k()
    run_dir = tmp_path / "run"
    run_dir.mkdir()
    (tmp_path / "results").mkdir()
    IO.write_count_first(str(run_dir / "b.txt"), b)
    IO.write_count_first(str(run_dir / "c.txt"), ovs[0])
    env = dict(os.environ, SEGCLIP_SAVE="1", SEGCLIP_PRECISION="17")
    r = subprocess.run([SEG, "-union", "b.txt", "c.txt"], capture_output=True, text=True, timeout=120, cwd=run_dir,
                       env=env)
    assert r.returncode == 0, r.stderr
    got = IO.read_results(str(tmp_path / "results" / "Fig-def-R1.txt"))
    want = O.Tree(b, ovs).clip(1, canonical=True, union=True).rings
    assert len(got) == len(want) > 1
    assert all(np.array_equal(g.view(np.uint64), w.view(np.uint64)) for g, w in zip(got, want))
