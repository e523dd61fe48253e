"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/segclip.h declares, and its host-only helpers work (no GPU calls)."""
import os
import re

import numpy as np

import segclip
from segclip._lib import SYMBOLS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "segclip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(segclip_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_are_exported():
    L = segclip.lib()
    declared = _declared()
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(SYMBOLS) == declared


def test_synth_is_deterministic_and_simple():
    a = segclip.synth.harmonic(5000, 7)
    b = segclip.synth.harmonic(5000, 7)
    assert np.array_equal(a, b)
    assert segclip.synth.count_vertical(a) == 0
    # star-shaped about the centre with increasing angle
    ang = np.unwrap(np.arctan2(a[:, 1], a[:, 0]))
    assert np.all(np.diff(ang) > 0)


def test_create_without_gpu_fails_cleanly():
    import ctypes as C
    h = C.c_void_p()
    rc = segclip.lib().segclip_create(C.byref(h), 0, None)
    if rc == 0:  # a GPU is present: nothing to check here
        segclip.lib().segclip_destroy(h)
    else:
        assert rc in (segclip._lib.EHIP, segclip._lib.EINVAL)
