"""Many-to-many clipping of two polygon layers: candidate pairs by bounding-box
overlap on the GPU (segclip_mbr_join), then every candidate pair clipped in
one batch (segclip_set_pairs + segclip_clip_batch).

    pairs = mbr_join(layer_a, layer_b)          # (n, 2) int32, sorted by (a, b)
    pairs, clipper = clip_layers(layer_a, layer_b)
    rings = clipper.batch_rings()               # per candidate pair
    areas = clipper.batch_area()                # 4 x n: intersection, a, b, IoU

A layer is a list of (n, 2) float64 rings (no repeated closing vertex).  The
reference joins layers through an R-tree of polygon MBRs
(lib/optimizedFostersAlgorithm/readShapefile.cpp) before clipping each pair;
pairs with disjoint boxes have an empty intersection and are never clipped.
"""
import ctypes as C

import numpy as np

from ._lib import ESPACE, OK, SegclipError, as_ptr, lib


def _device_rings(rings):
    import torch
    sizes = np.array([len(r) for r in rings], dtype=np.int32)
    xy = torch.tensor(np.concatenate([np.asarray(r, dtype=np.float64).reshape(-1, 2) for r in rings]),
                      dtype=torch.float64, device="cuda")
    return xy, sizes


def mbr_join(layer_a, layer_b):
    """(a, b) index pairs whose bounding boxes overlap (closed), sorted."""
    import torch
    ax, asz = _device_rings(layer_a)
    bx, bsz = _device_rings(layer_b)
    torch.cuda.synchronize()
    L = lib()
    need = C.c_int64(0)
    cap = 0
    out = np.zeros(2, np.int32)
    for _ in range(2):
        rc = L.segclip_mbr_join(C.c_void_p(ax.data_ptr()), as_ptr(asz, C.c_int32), len(asz),
                                C.c_void_p(bx.data_ptr()), as_ptr(bsz, C.c_int32), len(bsz), as_ptr(out, C.c_int32),
                                cap, C.byref(need), None)
        if rc == OK:
            return out[:2 * need.value].reshape(-1, 2).copy()
        if rc != ESPACE:
            raise SegclipError(rc, "segclip_mbr_join failed")
        cap = need.value
        out = np.zeros(2 * max(cap, 1), np.int32)
    raise SegclipError(ESPACE, "segclip_mbr_join: output size changed between calls")


def clip_layers(layer_a, layer_b, provenance=False):
    """Clip every candidate pair of two layers in one batch.  Returns the pairs
    and the Clipper holding the batch (rings, status, areas per pair)."""
    from . import Clipper
    pairs = mbr_join(layer_a, layer_b)
    cl = Clipper()
    if len(pairs) == 0:
        return pairs, cl
    cl.set_pairs([layer_a[a] for a, _ in pairs], [layer_b[b] for _, b in pairs])
    cl.set_provenance(provenance)
    cl.clip_batch()
    return pairs, cl
