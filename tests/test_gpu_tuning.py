"""GPU checks of the tuning settings: every value of "stab_tiles" (64-column
tiles per stab work unit) and of "stab_units_per_wave" (consecutive units one
wave runs) gives bit-identical pairs, intersections and rings,
equal to the oracle's; only the number of stab waves changes.  Out-of-range
values and unknown names are refused with SEGCLIP_EINVAL."""
import math

import numpy as np
import pytest

import oracle_py as O
import segclip.synth as S
from fixtures import rings_equal

pytestmark = pytest.mark.gpu

TILES = (1, 2, 3, 4, 8, 16)
# (stab_tiles, stab_units_per_wave) combinations beside the defaults (2, 2)
COMBOS = ((2, 1), (2, 2), (2, 3), (2, 4), (2, 16), (2, 32), (4, 4), (1, 2))


def _snapshot(cl, k=1):
    pb, pc, pt = cl.pairs(k)
    xy, qa, qb = cl.intersection_points()
    cnt = cl.counts()
    return {"pairs": (pb.copy(), pc.copy(), pt.copy()), "ix": xy.copy(), "rings": cl.rings(),
            "tests": cnt["tests"], "units": cnt["stab_units"]}


def test_stab_tiles_do_not_change_results(gpu):
    import segclip
    b = S.harmonic(20000, 2001)
    o = S.harmonic(60000, 2002, 0.15, -0.1, 1.0, 17.0 * math.pi / 180.0)
    print("[gpu-tuning] harmonic 20k x 60k, stab_tiles", TILES, flush=True)
    cl = segclip.Clipper().set_polygons(b, [o])
    assert cl.get_tuning("stab_tiles") == 2
    snaps = {}
    for t in TILES:
        print("[gpu-tuning] stab_tiles=%d" % t, flush=True)
        cl.set_tuning("stab_tiles", t)
        assert cl.get_tuning("stab_tiles") == t
        cl.reset()
        cl.clip(1)
        assert cl.counts()["undefined"] & 1 == 0
        snaps[t] = _snapshot(cl)
    cl.set_tuning("stab_tiles", 2)
    combo = {}
    for t, u in COMBOS:
        print("[gpu-tuning] stab_tiles=%d stab_units_per_wave=%d" % (t, u), flush=True)
        cl.set_tuning("stab_tiles", t)
        cl.set_tuning("stab_units_per_wave", u)
        assert cl.get_tuning("stab_units_per_wave") == u
        cl.reset()
        cl.clip(1)
        combo[(t, u)] = _snapshot(cl)
    ref = snaps[2]
    assert len(ref["pairs"][0]) > 1000
    for (t, u), s in combo.items():
        for a, r in zip(s["pairs"], ref["pairs"]):
            assert np.array_equal(a, r), "stab_tiles=%d units_per_wave=%d: pairs differ" % (t, u)
        assert np.array_equal(s["ix"].view(np.uint64), ref["ix"].view(np.uint64)), (t, u)
        err = rings_equal(s["rings"], ref["rings"])
        assert err is None, "stab_tiles=%d units_per_wave=%d: %s" % (t, u, err)
    for t in TILES:
        s = snaps[t]
        for a, r in zip(s["pairs"], ref["pairs"]):
            assert np.array_equal(a, r), "stab_tiles=%d: pairs differ" % t
        assert np.array_equal(s["ix"].view(np.uint64), ref["ix"].view(np.uint64)), t
        err = rings_equal(s["rings"], ref["rings"])
        assert err is None, "stab_tiles=%d: %s" % (t, err)
        assert s["tests"] == ref["tests"], t
    # more tiles per unit -> fewer (longer) waves
    units = [snaps[t]["units"] for t in TILES]
    assert all(u0 >= u1 for u0, u1 in zip(units, units[1:])), units
    assert units[0] > units[-1]
    # and the rings equal the oracle's
    r = O.Tree(b, [o]).clip(1, canonical=True)
    err = rings_equal(ref["rings"], r.rings)
    assert err is None, err


def test_stab_tiles_batched_pairs(gpu):
    import segclip
    bases, ovs = S.small_pairs(256, 4300)
    bases.append(S.harmonic(5000, 71))
    ovs.append(S.star(3000, 72, 0.1, 0.05, 0.4, 0.9))
    print("[gpu-tuning] paired batch", len(bases), flush=True)
    got = {}
    for t, u in ((1, 1), (2, 1), (8, 1), (2, 4), (4, 8)):
        cl = segclip.Clipper().set_pairs(bases, ovs)
        cl.set_tuning("stab_tiles", t)
        cl.set_tuning("stab_units_per_wave", u)
        cl.clip_batch()
        got[(t, u)] = cl.batch_rings()
        cl.close()
    for t in ((1, 1), (8, 1), (2, 4), (4, 8)):
        for i in range(len(bases)):
            err = rings_equal(got[t][i], got[(2, 1)][i])
            assert err is None, "settings %s pair %d: %s" % (t, i, err)


def test_tuning_refuses_bad_settings(gpu):
    import segclip
    from segclip._lib import SegclipError
    cl = segclip.Clipper()
    for v in (0, 17, -1):
        with pytest.raises(SegclipError):
            cl.set_tuning("stab_tiles", v)
    for v in (0, 33):
        with pytest.raises(SegclipError):
            cl.set_tuning("stab_units_per_wave", v)
    assert cl.get_tuning("stab_units_per_wave") == 2
    with pytest.raises(SegclipError):
        cl.set_tuning("no_such_setting", 2)
    with pytest.raises(SegclipError):
        cl.get_tuning("no_such_setting")
    assert cl.get_tuning("stab_tiles") == 2
    cl.close()
