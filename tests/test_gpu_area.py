"""GPU parity of the differentiable area / IoU (row A19).

Forward areas against the CPU shoelace over the oracle's rings; gradients
against the CPU reference gradient (area_ref.grad, itself pinned to finite
differences of the oracle in test_area_ref.py) and against central finite
differences of the GPU's own areas (h = 1e-7) on pairs whose minimum
vertex-to-edge distance exceeds 1e-4 and whose consecutive vertex triples
stay clear of the clean-up tolerance (area_ref.fd_safe); the torch.autograd
binding with and without recompute in backward.
"""
import numpy as np
import pytest

import area_ref as R
import segclip.synth as S
from fixtures import degenerate_cases

pytestmark = pytest.mark.gpu

FD_H = 1e-7
FD_REL = 1e-6


def _generic(count, seed, limit):
    bases, ovs = S.small_pairs(count, seed)
    out = []
    for b, o in zip(bases, ovs):
        if R.fd_safe(b, o):
            out.append((b, o))
        if len(out) == limit:
            break
    return out


def _gpu_grad(bases, ovs, g_area=None, g_iou=None):
    import torch
    import segclip
    cl = segclip.Clipper().set_pairs(bases, ovs).set_provenance(True).clip_batch()
    J = len(bases)
    ga = torch.tensor(np.ones(J) if g_area is None else g_area, dtype=torch.float64, device="cuda")
    gi = torch.tensor(np.zeros(J) if g_iou is None else g_iou, dtype=torch.float64, device="cuda")
    gb = torch.empty((sum(len(b) for b in bases), 2), dtype=torch.float64, device="cuda")
    gc = torch.empty((sum(len(o) for o in ovs), 2), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    cl.batch_area_backward_device(ga.data_ptr(), gi.data_ptr(), gb.data_ptr(), gc.data_ptr())
    return cl.batch_area(), gb.cpu().numpy(), gc.cpu().numpy()


def _split(a, sizes):
    return np.split(a, np.cumsum(sizes)[:-1])


def test_batch_area_matches_cpu(gpu):
    import segclip
    bases, ovs = [], []
    for name, b, o, _ in degenerate_cases():
        bases.append(b)
        ovs.append(o[0])
    sb, so = S.small_pairs(64, 4300)
    bases += sb
    ovs += so
    bases.append(S.harmonic(3000, 77))
    ovs.append(S.star(2000, 78, 0.1, 0.05, 0.4, 0.9))
    print("[gpu-area] forward", len(bases), flush=True)
    cl = segclip.Clipper().set_pairs(bases, ovs).clip_batch()
    a = cl.batch_area()
    for j, (b, o) in enumerate(zip(bases, ovs)):
        i, iou = R.area_iou(b, o)
        assert a[0, j] == pytest.approx(i, rel=1e-12, abs=1e-13), j
        assert a[1, j] == pytest.approx(abs(0.5 * R.shoelace2(b)), rel=1e-12), j  # touch_inside: clockwise
        assert a[2, j] == pytest.approx(abs(0.5 * R.shoelace2(o)), rel=1e-12), j
        assert a[3, j] == pytest.approx(iou, rel=1e-11, abs=1e-13), j
    # the same bits on a second run (fixed-order reductions)
    cl.clip_batch()
    assert np.array_equal(cl.batch_area(), a)


def test_gradient_matches_cpu_reference(gpu):
    pairs = _generic(64, 4400, 24)
    bases, ovs = [p[0] for p in pairs], [p[1] for p in pairs]
    print("[gpu-area] gradient vs cpu", len(pairs), flush=True)
    _, gb, gc = _gpu_grad(bases, ovs)
    for j, (b, o, xb, xc) in enumerate(zip(bases, ovs, _split(gb, [len(x) for x in bases]),
                                          _split(gc, [len(x) for x in ovs]))):
        rb, rc = R.grad(b, o)
        scale = max(1.0, np.max(np.abs(rb)), np.max(np.abs(rc)))
        assert np.max(np.abs(xb - rb)) <= 1e-12 * scale, j
        assert np.max(np.abs(xc - rc)) <= 1e-12 * scale, j


def test_gradient_vs_gpu_finite_differences(gpu):
    import segclip
    pairs = _generic(256, 4500, 64)
    bases, ovs = [p[0] for p in pairs], [p[1] for p in pairs]
    print("[gpu-area] gradient vs finite differences", len(pairs), flush=True)
    rng = np.random.default_rng(7)
    db = [rng.standard_normal(b.shape) for b in bases]
    dc = [rng.standard_normal(o.shape) for o in ovs]
    gi = rng.standard_normal(len(bases))
    a0, gb, gc = _gpu_grad(bases, ovs, np.ones(len(bases)), gi)
    gb = _split(gb, [len(x) for x in bases])
    gc = _split(gc, [len(x) for x in ovs])

    def loss(sign):
        cl = segclip.Clipper().set_pairs([b + sign * FD_H * d for b, d in zip(bases, db)],
                                         [o + sign * FD_H * d for o, d in zip(ovs, dc)]).clip_batch()
        a = cl.batch_area()
        return a[0] + gi * a[3]

    fd = (loss(1.0) - loss(-1.0)) / (2 * FD_H)
    for j in range(len(bases)):
        an = np.sum(gb[j] * db[j]) + np.sum(gc[j] * dc[j])
        assert abs(an - fd[j]) <= FD_REL * max(1.0, abs(fd[j])), (j, an, fd[j])


def test_one_to_many_gradient(gpu):
    import torch
    import segclip
    b = S.star(300, 5, 0.0, 0.0, 0.6, 1.0)
    ovs = [S.star(40, 50 + i, cx, cy, 0.2, 0.5) for i, (cx, cy) in enumerate([(0.5, 0.1), (-0.4, 0.3), (0.0, -0.7)])]
    print("[gpu-area] one-to-many gradient", flush=True)
    cl = segclip.Clipper().set_polygons(b, ovs).set_provenance(True).clip_batch()
    J = len(ovs)
    ga = torch.ones(J, dtype=torch.float64, device="cuda")
    gb = torch.empty((len(b), 2), dtype=torch.float64, device="cuda")
    gc = torch.empty((sum(len(o) for o in ovs), 2), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    cl.batch_area_backward_device(ga.data_ptr(), 0, gb.data_ptr(), gc.data_ptr())
    gb, gc = gb.cpu().numpy(), gc.cpu().numpy()
    wb = np.zeros_like(b)
    for o, xc in zip(ovs, _split(gc, [len(o) for o in ovs])):
        rb, rc = R.grad(b, o)
        wb += rb
        assert np.max(np.abs(xc - rc)) <= 1e-12 * max(1.0, np.max(np.abs(rc)))
    assert np.max(np.abs(gb - wb)) <= 1e-12 * max(1.0, np.max(np.abs(wb)))


@pytest.mark.parametrize("recompute", [False, True])
def test_autograd_binding(gpu, recompute):
    import torch
    from segclip.autograd import clip_area_iou
    pairs = _generic(64, 4600, 16)
    bases, ovs = [p[0] for p in pairs], [p[1] for p in pairs]
    print("[gpu-area] autograd recompute=%s" % recompute, flush=True)
    bx = torch.tensor(np.concatenate(bases), dtype=torch.float64, device="cuda", requires_grad=True)
    ox = torch.tensor(np.concatenate(ovs), dtype=torch.float64, device="cuda", requires_grad=True)
    area, iou = clip_area_iou(bx, [len(b) for b in bases], ox, [len(o) for o in ovs], recompute=recompute)
    w = torch.linspace(0.5, 1.5, len(bases), dtype=torch.float64, device="cuda")
    (area * w + 2.0 * iou).sum().backward()
    a0, gb, gc = _gpu_grad(bases, ovs, w.cpu().numpy(), np.full(len(bases), 2.0))
    assert np.array_equal(area.detach().cpu().numpy(), a0[0])
    assert np.allclose(bx.grad.cpu().numpy(), gb, rtol=0, atol=1e-13)
    assert np.allclose(ox.grad.cpu().numpy(), gc, rtol=0, atol=1e-13)


def test_clockwise_rings(gpu):
    """Clockwise base and/or overlay rings: areas and IoU equal the CPU
    reference's (orientation-free), the gradient equals the CPU reference
    gradient and central finite differences of the GPU's own areas."""
    import segclip
    b0, o0 = S.star(300, 5, 0.0, 0.0, 0.6, 1.0), S.star(200, 6, 0.3, 0.1, 0.3, 0.8)
    pairs = [(b0, o0)] + _generic(64, 4400, 6)
    bases, ovs = [], []
    for i, (b, o) in enumerate(pairs):
        for fb, fo in ((True, False), (False, True), (True, True)):
            bases.append(b[::-1].copy() if fb else b)
            ovs.append(o[::-1].copy() if fo else o)
    print("[gpu-area] clockwise rings", len(bases), flush=True)
    cl = segclip.Clipper().set_pairs(bases, ovs).clip_batch()
    a = cl.batch_area()
    for j, (b, o) in enumerate(zip(bases, ovs)):
        i, iou = R.area_iou(b, o)
        assert a[0, j] == pytest.approx(i, rel=1e-12, abs=1e-13), j
        assert a[1, j] == pytest.approx(abs(0.5 * R.shoelace2(b)), rel=1e-12), j
        assert a[2, j] == pytest.approx(abs(0.5 * R.shoelace2(o)), rel=1e-12), j
        assert a[3, j] == pytest.approx(iou, rel=1e-11, abs=1e-13), j
    rng = np.random.default_rng(11)
    gi = rng.standard_normal(len(bases))
    _, gb, gc = _gpu_grad(bases, ovs, np.ones(len(bases)), gi)
    gbs, gcs = _split(gb, [len(x) for x in bases]), _split(gc, [len(x) for x in ovs])
    for j in range(3, len(bases)):  # the fd-safe pairs: CPU reference gradient of the area alone
        _, xb, xc = _gpu_grad([bases[j]], [ovs[j]])
        rb, rc = R.grad(bases[j], ovs[j])
        scale = max(1.0, np.max(np.abs(rb)), np.max(np.abs(rc)))
        assert np.max(np.abs(xb - rb)) <= 1e-12 * scale, j
        assert np.max(np.abs(xc - rc)) <= 1e-12 * scale, j
    db = [rng.standard_normal(b.shape) for b in bases[3:]]
    dc = [rng.standard_normal(o.shape) for o in ovs[3:]]

    def loss(sign):
        c2 = segclip.Clipper().set_pairs([b + sign * FD_H * d for b, d in zip(bases[3:], db)],
                                         [o + sign * FD_H * d for o, d in zip(ovs[3:], dc)]).clip_batch()
        x = c2.batch_area()
        return x[0] + gi[3:] * x[3]

    fd = (loss(1.0) - loss(-1.0)) / (2 * FD_H)
    for j in range(len(db)):
        an = np.sum(gbs[j + 3] * db[j]) + np.sum(gcs[j + 3] * dc[j])
        assert abs(an - fd[j]) <= FD_REL * max(1.0, abs(fd[j])), (j, an, fd[j])
