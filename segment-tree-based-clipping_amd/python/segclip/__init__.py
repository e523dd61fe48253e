"""segclip -- Python view of libsegclip, the MI355X segment-tree polygon clipper.

`Clipper` drives the C ABI stage by stage (the same stages as the reference's
SegTree methods) and copies intermediate arrays back as numpy arrays; `synth`
builds the seeded workload rings.  The compute runs only in libsegclip.so on
the GPU: if the library or a GPU is missing the calls raise.
"""
import ctypes as C

import numpy as np

from ._lib import (Counts, SegclipError, as_ptr, lib, OK, ESPACE)  # noqa: F401
from . import synth  # noqa: F401

__all__ = ["Clipper", "SegclipError", "synth", "lib", "memory", "use_torch_allocator", "torch_allocator"]


def memory(reset_peak=False):
    """(current, peak) device bytes held by libsegclip buffers, all contexts."""
    cur, pk = C.c_int64(0), C.c_int64(0)
    lib().segclip_memory(C.byref(cur), C.byref(pk), 1 if reset_peak else 0)
    return cur.value, pk.value


_ALLOC_FN = C.CFUNCTYPE(C.c_void_p, C.c_size_t, C.c_void_p)
_FREE_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p)
_hooks = None
_torch_alloc_on = False


def use_torch_allocator(on=True):
    """Allocate libsegclip's device buffers from torch's caching allocator
    (on=True) or with hipMalloc (on=False), for buffers allocated from now on.
    Released buffers then return to torch's pool, so a clip after a release
    (for instance the recomputing backward) reuses them without driver calls,
    and torch can use them in between."""
    global _hooks, _torch_alloc_on
    import torch
    if not on:
        _check_rc(lib().segclip_set_allocator(None, None, None))
        _torch_alloc_on = False
        return  # the old callbacks stay referenced: buffers from them may still be live

    def _alloc(nbytes, _user):
        try:
            return torch.cuda.caching_allocator_alloc(int(nbytes), torch.cuda.current_device())
        except Exception:  # reported to the library as out of memory
            return None

    def _free(ptr, _user):
        try:
            torch.cuda.caching_allocator_delete(ptr)
        except Exception:  # interpreter shutdown
            pass

    if _hooks is None:
        _hooks = (_ALLOC_FN(_alloc), _FREE_FN(_free))
    _check_rc(lib().segclip_set_allocator(C.cast(_hooks[0], C.c_void_p), C.cast(_hooks[1], C.c_void_p), None))
    _torch_alloc_on = True


class torch_allocator:
    """Context manager: libsegclip buffers allocated inside come from torch's
    caching allocator; the previous setting is restored on exit."""

    def __enter__(self):
        self._was = _torch_alloc_on
        use_torch_allocator(True)
        return self

    def __exit__(self, *exc):
        use_torch_allocator(self._was)
        return False


def _check_rc(rc):
    if rc != OK:
        raise SegclipError(rc, "segclip_set_allocator failed")


class Clipper:
    """One clipping problem (base ring + K overlay rings) resident on one GPU."""

    def __init__(self, device=0, stream=None):
        L = lib()
        h = C.c_void_p()
        rc = L.segclip_create(C.byref(h), int(device), C.c_void_p(stream) if stream else None)
        if rc != OK:
            raise SegclipError(rc, "segclip_create failed (no usable HIP device?)")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().segclip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------------
    def _check(self, rc):
        if rc != OK:
            raise SegclipError(rc, lib().segclip_last_error(self._h).decode())

    def set_polygons(self, base, overlays):
        base = np.ascontiguousarray(base, dtype=np.float64).reshape(-1, 2)
        ovs = [np.ascontiguousarray(o, dtype=np.float64).reshape(-1, 2) for o in overlays]
        sizes = np.array([len(o) for o in ovs], dtype=np.int32)
        cxy = np.ascontiguousarray(np.concatenate(ovs, axis=0))
        self.nb, self.K, self.sizes = len(base), len(ovs), sizes
        self._check(lib().segclip_set_polygons(self._h, as_ptr(base, C.c_double), len(base),
                                               as_ptr(cxy, C.c_double), as_ptr(sizes, C.c_int32), len(ovs)))
        return self

    def set_pairs(self, bases, overlays):
        """P independent pairs: pair k clips bases[k] against overlays[k]."""
        bs = [np.ascontiguousarray(b, dtype=np.float64).reshape(-1, 2) for b in bases]
        ovs = [np.ascontiguousarray(o, dtype=np.float64).reshape(-1, 2) for o in overlays]
        if len(bs) != len(ovs):
            raise ValueError("set_pairs: %d bases for %d overlays" % (len(bs), len(ovs)))
        bsz = np.array([len(b) for b in bs], dtype=np.int32)
        csz = np.array([len(o) for o in ovs], dtype=np.int32)
        bxy = np.ascontiguousarray(np.concatenate(bs, axis=0))
        cxy = np.ascontiguousarray(np.concatenate(ovs, axis=0))
        self.nb, self.K, self.sizes = int(bsz.sum()), len(ovs), csz
        self._check(lib().segclip_set_pairs(self._h, as_ptr(bxy, C.c_double), as_ptr(bsz, C.c_int32),
                                            as_ptr(cxy, C.c_double), as_ptr(csz, C.c_int32), len(ovs)))
        return self

    def set_pairs_device(self, base_ptr, base_sizes, ov_ptr, ov_sizes):
        """Pairs already in device memory (pointers to x,y doubles)."""
        bsz = np.ascontiguousarray(base_sizes, dtype=np.int32)
        csz = np.ascontiguousarray(ov_sizes, dtype=np.int32)
        self.nb, self.K, self.sizes = int(bsz.sum()), len(csz), csz
        self._bsz = bsz
        self._check(lib().segclip_set_pairs_device(self._h, C.c_void_p(base_ptr), as_ptr(bsz, C.c_int32),
                                                   C.c_void_p(ov_ptr), as_ptr(csz, C.c_int32), len(csz)))
        return self

    def set_operation(self, union=False):
        """Intersection (default) or union of the following clips."""
        self._check(lib().segclip_set_operation(self._h, 1 if union else 0))
        return self

    def set_provenance(self, on=True):
        self._check(lib().segclip_set_provenance(self._h, 1 if on else 0))
        return self

    def release(self):
        self._check(lib().segclip_release(self._h))
        return self

    def batch_area(self):
        """Per job of the last batch: (intersection area, base area, overlay area, IoU) arrays."""
        L = lib()
        n = self._need(L.segclip_batch_area, None, 0)
        out = np.zeros(n, np.float64)
        self._check(L.segclip_batch_area(self._h, as_ptr(out, C.c_double), n, None))
        return out.reshape(4, -1)

    def batch_area_device(self, out_ptr):
        self._check(lib().segclip_batch_area_device(self._h, C.c_void_p(out_ptr)))

    def pairs_area_graph(self, base_ptr, ov_ptr, out_ptr):
        """Areas / IoU of every pair into out (device, 4 x P doubles) through the
        captured graph; base_ptr / ov_ptr: device rings laid out as for set_pairs."""
        self._check(lib().segclip_pairs_area_graph(self._h, C.c_void_p(base_ptr), C.c_void_p(ov_ptr),
                                                   C.c_void_p(out_ptr)))
        return self

    def batch_area_backward_device(self, g_area_ptr, g_iou_ptr, grad_b_ptr, grad_c_ptr):
        self._check(lib().segclip_batch_area_backward_device(
            self._h, C.c_void_p(g_area_ptr) if g_area_ptr else None, C.c_void_p(g_iou_ptr) if g_iou_ptr else None,
            C.c_void_p(grad_b_ptr), C.c_void_p(grad_c_ptr)))

    def build_tree(self):
        self._check(lib().segclip_build_tree(self._h))
        return self

    def materialize_end_lists(self):
        self._check(lib().segclip_materialize_end_lists(self._h))
        return self

    def find_candidates(self):
        self._check(lib().segclip_find_candidates(self._h))
        return self

    def intersections(self, k=1):
        self._check(lib().segclip_intersections(self._h, int(k)))
        return self

    def label_traverse(self):
        self._check(lib().segclip_label_traverse(self._h))
        return self

    def clip(self, k=1):
        self._check(lib().segclip_clip(self._h, int(k)))
        return self

    def clip_batch(self, k0=1, k1=None):
        """Overlays (or pairs) k0..k1 as one batch of launches."""
        k1 = self.K if k1 is None else k1
        self._check(lib().segclip_clip_batch(self._h, int(k0), int(k1)))
        return self

    def reset(self):
        self._check(lib().segclip_reset(self._h))
        return self

    def set_polygons_device(self, base_ptr, nb, ov_ptr, sizes):
        """Polygons already in device memory (pointers to x,y doubles)."""
        sizes = np.ascontiguousarray(sizes, dtype=np.int32)
        self.nb, self.K, self.sizes = int(nb), len(sizes), sizes
        self._check(lib().segclip_set_polygons_device(self._h, C.c_void_p(base_ptr), int(nb), C.c_void_p(ov_ptr),
                                                      as_ptr(sizes, C.c_int32), len(sizes)))
        return self

    def counts(self):
        c = Counts()
        self._check(lib().segclip_get_counts(self._h, C.byref(c)))
        return c.as_dict()

    def set_timing(self, on=True):
        self._check(lib().segclip_set_timing(self._h, 1 if on else 0))

    def set_tuning(self, name, value):
        """A tuning setting of this context (see segclip.tune); results do not depend on it."""
        self._check(lib().segclip_set_tuning(self._h, name.encode(), int(value)))
        return self

    def get_tuning(self, name):
        v = C.c_int(0)
        self._check(lib().segclip_get_tuning(self._h, name.encode(), C.byref(v)))
        return v.value

    def apply_tuning(self, settings):
        """Apply a dict of settings, e.g. the 'settings' of a segclip.tune JSON file."""
        for k, v in settings.items():
            self.set_tuning(k, v)
        return self

    def stage_times(self):
        a = (C.c_float * 8)()
        self._check(lib().segclip_get_stage_times(self._h, a))
        names = ["elementary", "tree", "endlist", "cover", "candidates", "stab_kernel", "post", "label"]
        return dict(zip(names, list(a)))

    # ---- copy-out ------------------------------------------------------
    def _need(self, fn, *args):
        n = C.c_int64(0)
        rc = fn(self._h, *args, C.byref(n))
        if rc not in (OK, ESPACE):
            self._check(rc)
        return n.value

    def elementary(self):
        L = lib()
        n = self._need(L.segclip_copy_elementary, None, 0)
        out = np.zeros(n, np.float64)
        self._check(L.segclip_copy_elementary(self._h, as_ptr(out, C.c_double), n, None))
        return out

    def node_intervals(self):
        L = lib()
        n = self._need(L.segclip_copy_node_intervals, None, 0)
        out = np.zeros(n, np.float64)
        self._check(L.segclip_copy_node_intervals(self._h, as_ptr(out, C.c_double), n, None))
        return out

    def cover_lists(self):
        L = lib()
        n = self._need(L.segclip_copy_cover_lists, None, None, 0)
        T = self.counts()["tree_size"]
        off = np.zeros(T + 1, np.uint32)
        e = np.zeros(max(n, 1), np.uint32)
        self._check(L.segclip_copy_cover_lists(self._h, as_ptr(off, C.c_uint32), as_ptr(e, C.c_uint32), n, None))
        return off.astype(np.int64), e[:n].astype(np.int64)

    def end_lists(self):
        L = lib()
        n = self._need(L.segclip_copy_end_lists, None, None, 0)
        T = self.counts()["tree_size"]
        off = np.zeros(T + 1, np.uint32)
        e = np.zeros(max(n, 1), np.uint32)
        self._check(L.segclip_copy_end_lists(self._h, as_ptr(off, C.c_uint32), as_ptr(e, C.c_uint32), n, None))
        return off.astype(np.int64), e[:n].astype(np.int64)

    def leaf_targets(self):
        L = lib()
        n = self._need(L.segclip_copy_leaf_targets, None, None, 0)
        a = np.zeros(n, np.int32)
        b = np.zeros(n, np.int32)
        self._check(L.segclip_copy_leaf_targets(self._h, as_ptr(a, C.c_int32), as_ptr(b, C.c_int32), n, None))
        return a, b

    def pairs(self, k=1):
        L = lib()
        n = self._need(L.segclip_copy_pairs, int(k), None, None, None, 0)
        b = np.zeros(max(n, 1), np.int32)
        c = np.zeros(max(n, 1), np.int32)
        t = np.zeros(max(n, 1), np.int32)
        self._check(L.segclip_copy_pairs(self._h, int(k), as_ptr(b, C.c_int32), as_ptr(c, C.c_int32),
                                         as_ptr(t, C.c_int32), n, None))
        return b[:n], c[:n], t[:n]

    def intersection_points(self):
        L = lib()
        n = self._need(L.segclip_copy_intersections, None, None, None, 0)
        xy = np.zeros((max(n, 1), 2), np.float64)
        qa = np.zeros(max(n, 1), np.int32)
        qb = np.zeros(max(n, 1), np.int32)
        self._check(L.segclip_copy_intersections(self._h, as_ptr(xy, C.c_double), as_ptr(qa, C.c_int32),
                                                 as_ptr(qb, C.c_int32), n, None))
        return xy[:n], qa[:n], qb[:n]

    def augmented(self, side):
        L = lib()
        n = self._need(L.segclip_copy_augmented, int(side), None, None, None, None, 0)
        xy = np.zeros((max(n, 1), 2), np.float64)
        nbr = np.zeros(max(n, 1), np.int32)
        av = np.zeros(max(n, 1), np.int32)
        lab = np.zeros(max(n, 1), np.int32)
        self._check(L.segclip_copy_augmented(self._h, int(side), as_ptr(xy, C.c_double), as_ptr(nbr, C.c_int32),
                                             as_ptr(av, C.c_int32), as_ptr(lab, C.c_int32), n, None))
        return xy[:n], nbr[:n], av[:n], lab[:n]

    def rings(self):
        L = lib()
        need = (C.c_int64 * 2)()
        rc = L.segclip_copy_rings(self._h, None, 0, None, 0, need)
        if rc not in (OK, ESPACE):
            self._check(rc)
        off = np.zeros(need[0], np.int32)
        xy = np.zeros((max(need[1], 1), 2), np.float64)
        self._check(L.segclip_copy_rings(self._h, as_ptr(off, C.c_int32), need[0], as_ptr(xy, C.c_double),
                                         need[1], need))
        xy = xy[:need[1]]
        return [xy[off[i]:off[i + 1]] for i in range(len(off) - 1)]

    def batch_ring_offsets(self):
        """Per job of the last batch: ring offsets (jobs+1) into rings()."""
        L = lib()
        n = self._need(L.segclip_copy_batch_rings, None, 0)
        off = np.zeros(n, np.int32)
        self._check(L.segclip_copy_batch_rings(self._h, as_ptr(off, C.c_int32), n, None))
        return off

    def batch_status(self):
        L = lib()
        n = self._need(L.segclip_copy_batch_status, None, 0)
        st = np.zeros(n, np.int32)
        self._check(L.segclip_copy_batch_status(self._h, as_ptr(st, C.c_int32), n, None))
        return st

    def batch_routes(self):
        """Per job of the last batch: (route, reason bits) of the labelling stage."""
        L = lib()
        n = self._need(L.segclip_copy_batch_routes, None, None, 0)
        r = np.zeros(n, np.int32)
        w = np.zeros(n, np.int32)
        self._check(L.segclip_copy_batch_routes(self._h, as_ptr(r, C.c_int32), as_ptr(w, C.c_int32), n, None))
        return r, w

    def batch_rings(self):
        """Output rings of every job of the last batch: a list (per job) of lists of (n,2) arrays."""
        r = self.rings()
        off = self.batch_ring_offsets()
        return [r[off[j]:off[j + 1]] for j in range(len(off) - 1)]
