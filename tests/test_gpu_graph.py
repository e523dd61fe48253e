"""GPU checks of the small-pair forward replayed as a hipGraph
(segclip_pairs_area_graph): two replays give the same bits, and those equal the
eager path (clip_batch + batch_area) bit for bit; after the caller updates its
rings in place a replay gives the eager result for the new rings, with the
program's own eager work and synchronisations between replays; a batch with
pairs the persistent kernel does not cover gives the staged path's areas."""
import numpy as np
import pytest

import segclip.synth as S
from fixtures import degenerate_cases

pytestmark = pytest.mark.gpu


def _eager(bases, ovs):
    import segclip
    cl = segclip.Clipper().set_pairs(bases, ovs).clip_batch()
    return cl.batch_area()


def _dev(bases, ovs):
    import torch
    bx = torch.tensor(np.concatenate(bases), dtype=torch.float64, device="cuda")
    ox = torch.tensor(np.concatenate(ovs), dtype=torch.float64, device="cuda")
    return bx, ox


def test_graph_replays_equal_eager(gpu):
    import torch
    import segclip
    bases, ovs = S.small_pairs(4096, 5100)
    J = len(bases)
    bx, ox = _dev(bases, ovs)
    cl = segclip.Clipper().set_pairs_device(bx.data_ptr(), [len(b) for b in bases], ox.data_ptr(),
                                            [len(o) for o in ovs])
    out = torch.empty((4, J), dtype=torch.float64, device="cuda")
    print("[gpu-graph] capture + replay", flush=True)
    cl.pairs_area_graph(bx.data_ptr(), ox.data_ptr(), out.data_ptr())
    first = out.cpu().numpy().copy()
    want = _eager(bases, ovs)
    assert np.array_equal(first.view(np.uint64), want.view(np.uint64))
    for rep in range(3):
        out.fill_(-1.0)  # eager work between replays
        torch.cuda.synchronize()
        cl.pairs_area_graph(bx.data_ptr(), ox.data_ptr(), out.data_ptr())
        assert np.array_equal(out.cpu().numpy().view(np.uint64), first.view(np.uint64)), rep
    # the caller moves its rings in place: the replay reads the new coordinates
    shift = torch.linspace(-0.05, 0.05, bx.shape[0], dtype=torch.float64, device="cuda")[:, None]
    bx += shift
    torch.cuda.synchronize()
    cl.pairs_area_graph(bx.data_ptr(), ox.data_ptr(), out.data_ptr())
    nb = np.split(bx.cpu().numpy(), np.cumsum([len(b) for b in bases])[:-1])
    want2 = _eager(nb, ovs)
    assert np.array_equal(out.cpu().numpy().view(np.uint64), want2.view(np.uint64))
    assert not np.array_equal(want2, want)


def test_graph_batch_with_uncovered_pairs(gpu):
    import torch
    import segclip
    bases, ovs = S.small_pairs(64, 5200)
    for name, b, o, _ in degenerate_cases():
        if len(b) <= 64 and len(o[0]) <= 64:  # shared edges (chains), touching vertices, squares
            bases.append(b)
            ovs.append(o[0])
    J = len(bases)
    bx, ox = _dev(bases, ovs)
    cl = segclip.Clipper().set_pairs_device(bx.data_ptr(), [len(b) for b in bases], ox.data_ptr(),
                                            [len(o) for o in ovs])
    out = torch.empty((4, J), dtype=torch.float64, device="cuda")
    print("[gpu-graph] degenerate fixtures", flush=True)
    for _ in range(2):
        cl.pairs_area_graph(bx.data_ptr(), ox.data_ptr(), out.data_ptr())
        got = out.cpu().numpy()
        want = _eager(bases, ovs)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_graph_rejects_large_rings(gpu):
    import torch
    import segclip
    from segclip._lib import SegclipError, EINVAL
    bases = [S.star(100, 3), S.star(20, 4)]
    ovs = [S.star(30, 5, 0.1, 0.0), S.star(20, 6, 0.1, 0.0)]
    bx, ox = _dev(bases, ovs)
    cl = segclip.Clipper().set_pairs_device(bx.data_ptr(), [100, 20], ox.data_ptr(), [30, 20])
    out = torch.empty((4, 2), dtype=torch.float64, device="cuda")
    with pytest.raises(SegclipError) as e:
        cl.pairs_area_graph(bx.data_ptr(), ox.data_ptr(), out.data_ptr())
    assert e.value.code == EINVAL
