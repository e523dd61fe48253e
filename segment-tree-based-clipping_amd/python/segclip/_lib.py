"""ctypes binding of libsegclip.so (the C ABI in include/segclip.h).

The library is loaded from the package directory; if it is missing the import
of any compute entry point fails loudly (there is no CPU fallback).
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LIB_PATH = os.environ.get("SEGCLIP_LIB", os.path.join(PKG_DIR, "libsegclip.so"))

OK, EINVAL, EHIP, ENOMEM, ESTATE, ERANGE, EUNDEF, ESPACE = range(8)

SYMBOLS = [
    "segclip_create", "segclip_destroy", "segclip_last_error", "segclip_version", "segclip_set_polygons",
    "segclip_set_polygons_device", "segclip_build_elementary", "segclip_build_node_intervals",
    "segclip_build_end_lists", "segclip_build_cover_lists", "segclip_materialize_end_lists", "segclip_build_tree",
    "segclip_find_candidates", "segclip_intersections", "segclip_label_traverse", "segclip_clip", "segclip_reset",
    "segclip_get_counts", "segclip_copy_elementary", "segclip_copy_node_intervals", "segclip_copy_cover_lists",
    "segclip_copy_end_lists", "segclip_copy_leaf_targets", "segclip_copy_pairs", "segclip_copy_intersections",
    "segclip_copy_augmented", "segclip_copy_rings", "segclip_stable_sort_by_key_f64",
    "segclip_stable_sort_by_key_i32", "segclip_stable_sort_by_key_i64", "segclip_exclusive_scan_i32",
    "segclip_exclusive_scan_i64", "segclip_synth_star", "segclip_synth_harmonic", "segclip_synth_convex",
    "segclip_synth_uniform", "segclip_count_vertical", "segclip_set_timing", "segclip_get_stage_times",
    "segclip_set_pairs", "segclip_set_pairs_device", "segclip_clip_batch", "segclip_copy_batch_rings",
    "segclip_copy_batch_status", "segclip_set_provenance", "segclip_batch_area", "segclip_batch_area_device",
    "segclip_batch_area_backward_device", "segclip_release", "segclip_memory", "segclip_set_allocator", "segclip_save_intersections_ids",
    "segclip_sort_intersections_at_edges", "segclip_copy_intersections_arrays", "segclip_init_labels_arrays",
    "segclip_copy_batch_routes", "segclip_set_tuning", "segclip_get_tuning",
    "segclip_stable_sort_by_key_f64_noalloc", "segclip_pairs_area_graph", "segclip_mbr_join",
    "segclip_set_operation",
]


class SegclipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("segclip error %d: %s" % (code, msg))
        self.code = code


class Counts(C.Structure):
    _fields_ = [("num_edges", C.c_int64), ("epc", C.c_int64), ("tree_size", C.c_int64),
                ("cover_entries", C.c_int64), ("leaf_end_entries", C.c_int64), ("end_entries", C.c_int64),
                ("tests", C.c_int64), ("candidates", C.c_int64), ("pairs", C.c_int64), ("overlay", C.c_int32),
                ("overlay_pairs", C.c_int32), ("nb_aug", C.c_int32), ("nc_aug", C.c_int32), ("rings", C.c_int32),
                ("undefined", C.c_int32), ("out_vertices", C.c_int64), ("stab_units", C.c_int64),
                ("near_pairs", C.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)
_i64p = C.POINTER(C.c_int64)
_vp = C.c_void_p


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libsegclip.so not built at %s (run __graft_entry__.build())" % LIB_PATH)
    # A process can hold one HIP/HSA runtime only.  PyTorch bundles its own
    # copy with the same soname as ROCm's (libamdhip64.so.7): loading torch
    # first makes the dynamic loader reuse that copy for libsegclip, so torch
    # tensors and this library share one runtime (loading libsegclip first
    # would leave torch without a usable GPU).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    sig = {
        "segclip_create": (C.c_int, [C.POINTER(_vp), C.c_int, _vp]),
        "segclip_destroy": (None, [_vp]),
        "segclip_last_error": (C.c_char_p, [_vp]),
        "segclip_version": (C.c_char_p, []),
        "segclip_set_polygons": (C.c_int, [_vp, _dp, C.c_int, _dp, _ip, C.c_int]),
        "segclip_set_polygons_device": (C.c_int, [_vp, _vp, C.c_int, _vp, _ip, C.c_int]),
        "segclip_set_pairs": (C.c_int, [_vp, _dp, _ip, _dp, _ip, C.c_int]),
        "segclip_set_pairs_device": (C.c_int, [_vp, _vp, _ip, _vp, _ip, C.c_int]),
        "segclip_clip_batch": (C.c_int, [_vp, C.c_int, C.c_int]),
        "segclip_copy_batch_rings": (C.c_int, [_vp, _ip, C.c_int64, _i64p]),
        "segclip_copy_batch_status": (C.c_int, [_vp, _ip, C.c_int64, _i64p]),
        "segclip_copy_batch_routes": (C.c_int, [_vp, _ip, _ip, C.c_int64, _i64p]),
        "segclip_set_provenance": (C.c_int, [_vp, C.c_int]),
        "segclip_batch_area": (C.c_int, [_vp, _dp, C.c_int64, _i64p]),
        "segclip_batch_area_device": (C.c_int, [_vp, _vp]),
        "segclip_batch_area_backward_device": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
        "segclip_memory": (C.c_int, [_i64p, _i64p, C.c_int]),
        "segclip_set_allocator": (C.c_int, [_vp, _vp, _vp]),
        "segclip_save_intersections_ids": (C.c_int, [_vp, _ip, _ip, C.c_int, _dp, _ip, _ip, _ip]),
        "segclip_sort_intersections_at_edges": (C.c_int, [_vp, _ip, _ip, C.c_int, _ip, _ip, C.c_int, _ip, C.c_int,
                                                          C.c_int, _ip]),
        "segclip_copy_intersections_arrays": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _ip, _ip, C.c_int, _ip, _ip,
                                                        _ip, _ip, _dp, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _ip, _ip,
                                                        _ip, _ip]),
        "segclip_init_labels_arrays": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp, _ip, _ip, _ip, _ip]),
        "segclip_intersections": (C.c_int, [_vp, C.c_int]),
        "segclip_clip": (C.c_int, [_vp, C.c_int]),
        "segclip_get_counts": (C.c_int, [_vp, C.POINTER(Counts)]),
        "segclip_copy_elementary": (C.c_int, [_vp, _dp, C.c_int64, _i64p]),
        "segclip_copy_node_intervals": (C.c_int, [_vp, _dp, C.c_int64, _i64p]),
        "segclip_copy_cover_lists": (C.c_int, [_vp, _up, _up, C.c_int64, _i64p]),
        "segclip_copy_end_lists": (C.c_int, [_vp, _up, _up, C.c_int64, _i64p]),
        "segclip_copy_leaf_targets": (C.c_int, [_vp, _ip, _ip, C.c_int64, _i64p]),
        "segclip_copy_pairs": (C.c_int, [_vp, C.c_int, _ip, _ip, _ip, C.c_int64, _i64p]),
        "segclip_copy_intersections": (C.c_int, [_vp, _dp, _ip, _ip, C.c_int64, _i64p]),
        "segclip_copy_augmented": (C.c_int, [_vp, C.c_int, _dp, _ip, _ip, _ip, C.c_int64, _i64p]),
        "segclip_copy_rings": (C.c_int, [_vp, _ip, C.c_int64, _dp, C.c_int64, _i64p]),
        "segclip_stable_sort_by_key_f64": (C.c_int, [_vp, _vp, C.c_int64, _vp]),
        "segclip_stable_sort_by_key_f64_noalloc": (C.c_int, [_vp, _vp, C.c_int64, _vp]),
        "segclip_pairs_area_graph": (C.c_int, [_vp, _vp, _vp, _vp]),
        "segclip_set_operation": (C.c_int, [_vp, C.c_int]),
        "segclip_mbr_join": (C.c_int, [_vp, _ip, C.c_int, _vp, _ip, C.c_int, _ip, C.c_int64, _i64p, _vp]),
        "segclip_stable_sort_by_key_i32": (C.c_int, [_vp, _vp, C.c_int64, _vp]),
        "segclip_stable_sort_by_key_i64": (C.c_int, [_vp, _vp, C.c_int64, _vp]),
        "segclip_exclusive_scan_i32": (C.c_int, [_vp, C.c_int64, _vp]),
        "segclip_exclusive_scan_i64": (C.c_int, [_vp, C.c_int64, _vp]),
        "segclip_synth_star": (C.c_int, [_dp, C.c_int, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_double,
                                         C.c_double]),
        "segclip_synth_harmonic": (C.c_int, [_dp, C.c_int, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                             C.c_double]),
        "segclip_synth_convex": (C.c_int, [_dp, C.c_int, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                           C.c_double]),
        "segclip_synth_uniform": (None, [_dp, C.c_int, C.c_uint64]),
        "segclip_count_vertical": (C.c_int, [_dp, C.c_int]),
        "segclip_set_timing": (C.c_int, [_vp, C.c_int]),
        "segclip_get_stage_times": (C.c_int, [_vp, C.POINTER(C.c_float)]),
        "segclip_set_tuning": (C.c_int, [_vp, C.c_char_p, C.c_int]),
        "segclip_get_tuning": (C.c_int, [_vp, C.c_char_p, _ip]),
    }
    for name in SYMBOLS:
        fn = getattr(L, name)  # AttributeError here = a declared symbol is not exported
        res, args = sig.get(name, (C.c_int, [_vp]))
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def as_ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))
