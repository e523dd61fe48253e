"""The CPU area / gradient reference (row A19) pinned without the GPU: the
analytic gradient over the oracle's output equals central finite differences
of the oracle's area, and the areas of the closed-form cases are exact."""
import numpy as np
import pytest

import area_ref as R
import segclip.synth as S
from fixtures import degenerate_cases


def test_closed_form_areas():
    for name, b, ovs, want in degenerate_cases():
        if want is None:
            continue
        # magnitudes: the touch_inside triangle is clockwise
        assert abs(R.area(b, ovs[0])) == pytest.approx(want, rel=1e-12, abs=1e-12), name
    # IoU of the squares: I = 1, P = 4, Q = 2
    name, b, ovs, _ = degenerate_cases()[0]
    i, iou = R.area_iou(b, ovs[0])
    assert iou == pytest.approx(1.0 / 5.0, rel=1e-12)


def test_contained_area_is_overlay_area():
    name, b, ovs, _ = degenerate_cases()[1]
    assert R.area(b, ovs[0]) == pytest.approx(0.5 * R.shoelace2(ovs[0]), rel=1e-12)


def _generic_pairs(count, seed):
    bases, ovs = S.small_pairs(count, seed)
    out = []
    for b, o in zip(bases, ovs):
        if R.fd_safe(b, o) and R.area(b, o) > 1e-3:
            out.append((b, o))
    return out


def test_analytic_gradient_matches_finite_differences():
    pairs = _generic_pairs(24, 4200)[:6]
    assert len(pairs) >= 4
    for b, o in pairs:
        gb, go = R.grad(b, o)
        fb, fo = R.fd_grad(b.copy(), o.copy())
        scale = max(np.max(np.abs(fb)), np.max(np.abs(fo)))
        assert np.max(np.abs(gb - fb)) <= 1e-6 * scale
        assert np.max(np.abs(go - fo)) <= 1e-6 * scale


def test_planted_sign_error_is_caught():
    """A flipped sign in the alpha term must fail the finite-difference check."""
    b, o = _generic_pairs(24, 4200)[0]
    gb, go = R.grad(b, o)
    fb, fo = R.fd_grad(b.copy(), o.copy())
    orig = R._dA
    try:
        R._dA = lambda p, q, r: tuple(-v for v in orig(p, q, r))
        gb2, go2 = R.grad(b, o)
    finally:
        R._dA = orig
    scale = max(np.max(np.abs(fb)), np.max(np.abs(fo)))
    assert max(np.max(np.abs(gb2 - fb)), np.max(np.abs(go2 - fo))) > 1e-3 * scale


def test_clockwise_rings_same_area_and_gradient_matches_fd():
    """Area and IoU do not depend on ring orientation; with clockwise rings the
    output mixes orientations, and the gradient (carrying each ring's sign)
    still equals central finite differences."""
    b0, o0 = S.star(300, 5, 0.0, 0.0, 0.6, 1.0), S.star(200, 6, 0.3, 0.1, 0.3, 0.8)
    want = R.area_iou(b0, o0)
    for fb in (False, True):
        for fo in (False, True):
            b = b0[::-1].copy() if fb else b0
            o = o0[::-1].copy() if fo else o0
            got = R.area_iou(b, o)
            assert got[0] == pytest.approx(want[0], rel=1e-12)
            assert got[1] == pytest.approx(want[1], rel=1e-12)
    # the signed sum differs from the area for the clockwise base: mixed orientations
    rings = R.clip_rings(b0[::-1].copy(), o0).rings
    assert abs(0.5 * sum(R.shoelace2(r) for r in rings) - want[0]) > 1e-3
    for b, o in _generic_pairs(24, 4200)[:2]:
        b, o = b[::-1].copy(), o[::-1].copy()
        gb, go = R.grad(b, o)
        fb, fo = R.fd_grad(b.copy(), o.copy())
        scale = max(np.max(np.abs(fb)), np.max(np.abs(fo)))
        assert np.max(np.abs(gb - fb)) <= 1e-6 * scale
        assert np.max(np.abs(go - fo)) <= 1e-6 * scale
