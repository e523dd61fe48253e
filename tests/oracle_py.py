"""ctypes view of the CPU oracle (oracle/libsegoracle.so).

Test infrastructure only: the oracle is the checker, never the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "libsegoracle.so")
REF_POLYCLIP = os.path.join(ORACLE_DIR, "_ref", "polyclip")
# the same program with its result rings written at 17 significant digits (oracle/polyclip_full.cpp)
REF_POLYCLIP17 = os.path.join(ORACLE_DIR, "_ref", "polyclip17")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR, "libsegoracle.so"])
        L = C.CDLL(LIB_PATH)
        vp, ip, dp, i64p = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int64)
        L.so_build.restype = vp
        L.so_build.argtypes = [dp, C.c_int, dp, ip, C.c_int]
        L.so_free.argtypes = [vp]
        for n in ["so_num_edges", "so_epc", "so_tree_size"]:
            getattr(L, n).restype = C.c_int
            getattr(L, n).argtypes = [vp]
        for n in ["so_elementary", "so_node_intervals", "so_bpoly", "so_cpoly"]:
            getattr(L, n).restype = dp
            getattr(L, n).argtypes = [vp]
        for n in ["so_cover_off", "so_end_off"]:
            getattr(L, n).restype = i64p
            getattr(L, n).argtypes = [vp]
        for n in ["so_cover_edges", "so_end_edges", "so_after_leaf", "so_before_leaf", "so_coff"]:
            getattr(L, n).restype = ip
            getattr(L, n).argtypes = [vp]
        L.so_edges.restype = vp
        L.so_edges.argtypes = [vp]
        L.so_clip.restype = vp
        L.so_clip.argtypes = [vp, C.c_int, C.c_int]
        L.so_clip_op.restype = vp
        L.so_clip_op.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.so_result_free.argtypes = [vp]
        for n in ["so_res_status", "so_res_num_pairs", "so_res_nb_aug", "so_res_nc_aug", "so_res_num_rings"]:
            getattr(L, n).restype = C.c_int
            getattr(L, n).argtypes = [vp]
        L.so_res_num_candidates.restype = C.c_int64
        L.so_res_num_candidates.argtypes = [vp]
        for n in ["so_res_pair_b", "so_res_pair_c", "so_res_pair_type", "so_res_alpha_b", "so_res_alpha_c",
                  "so_res_nbr_b", "so_res_nbr_c", "so_res_alphav_b", "so_res_alphav_c", "so_res_label_b",
                  "so_res_label_c", "so_res_ring_off"]:
            getattr(L, n).restype = ip
            getattr(L, n).argtypes = [vp]
        for n in ["so_res_intersections", "so_res_aug_b", "so_res_aug_c", "so_res_ring_xy"]:
            getattr(L, n).restype = dp
            getattr(L, n).argtypes = [vp]
        L.so_brute_pairs.restype = C.c_int64
        L.so_brute_pairs.argtypes = [vp, C.c_int, ip, ip, ip, C.c_int64]
        L.so_lsmf.restype = C.c_int
        L.so_lsmf.argtypes = [dp]
        L.so_classify.restype = C.c_int
        L.so_classify.argtypes = [dp, dp, dp]
        L.so_rings_area.restype = C.c_double
        L.so_rings_area.argtypes = [ip, dp, C.c_int]
        L.so_set_threads.argtypes = [C.c_int]
        L.so_get_threads.restype = C.c_int
        _lib = L
    return _lib


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops (results do not depend on it)."""
    lib().so_set_threads(int(n))


def get_threads():
    return lib().so_get_threads()


EDGE_DTYPE = np.dtype([("id", np.int32), ("type", np.int32), ("start", np.float64), ("end", np.float64),
                       ("y", np.float64)])


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    ct = np.ctypeslib.as_ctypes_type(dtype)
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,)).copy()


class Tree:
    """Oracle tree over one base ring and K overlay rings (A1-A6)."""

    def __init__(self, base, overlays):
        L = lib()
        base = np.ascontiguousarray(base, dtype=np.float64).reshape(-1, 2)
        ovs = [np.ascontiguousarray(o, dtype=np.float64).reshape(-1, 2) for o in overlays]
        csz = np.array([len(o) for o in ovs], dtype=np.int32)
        cxy = np.ascontiguousarray(np.concatenate(ovs, axis=0))
        self.nb, self.K = len(base), len(ovs)
        self.csz = csz
        self._h = L.so_build(base.ctypes.data_as(C.POINTER(C.c_double)), len(base),
                             cxy.ctypes.data_as(C.POINTER(C.c_double)), csz.ctypes.data_as(C.POINTER(C.c_int)),
                             len(ovs))
        if not self._h:
            raise ValueError("oracle rejected the input")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().so_free(self._h)
            self._h = None

    @property
    def num_edges(self):
        return lib().so_num_edges(self._h)

    def edges(self):
        n = self.num_edges
        buf = (C.c_char * (n * EDGE_DTYPE.itemsize)).from_address(lib().so_edges(self._h))
        return np.frombuffer(bytes(buf), dtype=EDGE_DTYPE)

    def elementary(self):
        return _arr(lib().so_elementary(self._h), lib().so_epc(self._h), np.float64)

    @property
    def tree_size(self):
        return lib().so_tree_size(self._h)

    def node_intervals(self):
        return _arr(lib().so_node_intervals(self._h), 2 * self.tree_size, np.float64)

    def cover(self):
        T = self.tree_size
        off = _arr(lib().so_cover_off(self._h), T + 1, np.int64)
        return off, _arr(lib().so_cover_edges(self._h), int(off[-1]), np.int32)

    def end_lists(self):
        T = self.tree_size
        off = _arr(lib().so_end_off(self._h), T + 1, np.int64)
        return off, _arr(lib().so_end_edges(self._h), int(off[-1]), np.int32)

    def leaf_targets(self):
        n = self.num_edges
        return (_arr(lib().so_after_leaf(self._h), n, np.int32), _arr(lib().so_before_leaf(self._h), n, np.int32))

    def clip(self, k=1, canonical=True, union=False):
        return Result(self, k, canonical, union)

    def brute_pairs(self, k=1):
        L = lib()
        n = L.so_brute_pairs(self._h, k, None, None, None, 0)
        b = np.zeros(max(n, 1), np.int32)
        c = np.zeros(max(n, 1), np.int32)
        t = np.zeros(max(n, 1), np.int32)
        ip = C.POINTER(C.c_int)
        L.so_brute_pairs(self._h, k, b.ctypes.data_as(ip), c.ctypes.data_as(ip), t.ctypes.data_as(ip), n)
        return b[:n], c[:n], t[:n]


class Result:
    """Oracle clip of the base against overlay k (A8-A17)."""

    def __init__(self, tree, k, canonical, union=False):
        L = lib()
        self._h = L.so_clip_op(tree._h, k, 1 if canonical else 0, 1 if union else 0)
        if not self._h:
            raise ValueError("oracle clip failed")
        h = self._h
        self.status = L.so_res_status(h)
        self.num_candidates = L.so_res_num_candidates(h)
        I = L.so_res_num_pairs(h)
        self.pair_b = _arr(L.so_res_pair_b(h), I, np.int32)
        self.pair_c = _arr(L.so_res_pair_c(h), I, np.int32)
        self.pair_type = _arr(L.so_res_pair_type(h), I, np.int32)
        self.intersections = _arr(L.so_res_intersections(h), 2 * I, np.float64).reshape(-1, 2)
        self.alpha_b = _arr(L.so_res_alpha_b(h), I, np.int32)
        self.alpha_c = _arr(L.so_res_alpha_c(h), I, np.int32)
        nb, nc = L.so_res_nb_aug(h), L.so_res_nc_aug(h)
        self.aug_b = _arr(L.so_res_aug_b(h), 2 * nb, np.float64).reshape(-1, 2)
        self.aug_c = _arr(L.so_res_aug_c(h), 2 * nc, np.float64).reshape(-1, 2)
        self.nbr_b = _arr(L.so_res_nbr_b(h), nb, np.int32)
        self.nbr_c = _arr(L.so_res_nbr_c(h), nc, np.int32)
        self.alphav_b = _arr(L.so_res_alphav_b(h), nb, np.int32)
        self.alphav_c = _arr(L.so_res_alphav_c(h), nc, np.int32)
        self.label_b = _arr(L.so_res_label_b(h), nb, np.int32)
        self.label_c = _arr(L.so_res_label_c(h), nc, np.int32)
        nr = L.so_res_num_rings(h)
        self.ring_off = _arr(L.so_res_ring_off(h), nr + 1, np.int32)
        self.ring_xy = _arr(L.so_res_ring_xy(h), 2 * int(self.ring_off[-1]), np.float64).reshape(-1, 2)
        L.so_result_free(h)
        self._h = None

    @property
    def rings(self):
        return [self.ring_xy[self.ring_off[i]:self.ring_off[i + 1]] for i in range(len(self.ring_off) - 1)]

    def area(self):
        a = 0.0
        for r in self.rings:
            x, y = r[:, 0], r[:, 1]
            a += 0.5 * float(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y))
        return a


def classify(p8):
    p = np.ascontiguousarray(p8, dtype=np.float64)
    a = C.c_double()
    b = C.c_double()
    t = lib().so_classify(p.ctypes.data_as(C.POINTER(C.c_double)), C.byref(a), C.byref(b))
    return t, a.value, b.value


def lsmf(p8):
    p = np.ascontiguousarray(p8, dtype=np.float64)
    return lib().so_lsmf(p.ctypes.data_as(C.POINTER(C.c_double)))


def write_poly_file(path, ring):
    """.poly text for the reference's sequential Foster clipper: 'x y,' lines, ';' closes."""
    with open(path, "w") as f:
        n = len(ring)
        for i, (x, y) in enumerate(ring):
            f.write("%.17g %.17g%s\n" % (x, y, ";" if i == n - 1 else ","))


def read_poly_file(path):
    rings, cur = [], []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            end = line.endswith(";")
            xy = line.rstrip(",;").split()
            cur.append((float(xy[0]), float(xy[1])))
            if end:
                rings.append(np.array(cur))
                cur = []
    return rings


def run_reference_foster(base, overlay, workdir, union=False, full_precision=False):
    """Run the reference's sequential Foster clipper (oracle/_ref/polyclip) and
    return its output rings (6 significant digits, as it prints them; 17 with
    full_precision, oracle/_ref/polyclip17)."""
    exe = REF_POLYCLIP17 if full_precision else REF_POLYCLIP
    if not os.path.exists(exe):
        return None
    p = os.path.join(workdir, "P.poly")
    q = os.path.join(workdir, "Q.poly")
    r = os.path.join(workdir, "R.poly")
    write_poly_file(p, base)
    write_poly_file(q, overlay)
    subprocess.run([exe] + (["-union"] if union else []) + [p, q, r], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, cwd=workdir)
    return read_poly_file(r)
