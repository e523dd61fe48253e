"""GPU check of segclip_stable_sort_by_key_f64_noalloc, the allocation-free
variant of segclip_stable_sort_by_key_f64: through the C ABI, two calls in a
row (the second on a smaller input, so the shared scratch array holds the
first call's leftovers) give exactly the keys and values of the allocating
entry point (stable: ties keep their input order; -0.0 equals 0.0), and so
does a call at the size limit; a size beyond the limit is rejected."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIMIT = 1 << 20


def _keys(n, seed):
    rng = np.random.default_rng(seed)
    k = np.round(rng.standard_normal(n) * 40) / 8.0  # many ties, negatives
    if n > 4:
        k[:4] = [0.0, -0.0, 0.0, -0.0]
    return k


def _sort(name, keys):
    import torch
    from segclip import lib
    dk = torch.from_numpy(np.ascontiguousarray(keys)).cuda()
    dv = torch.arange(len(keys), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    rc = getattr(lib(), name)(C.c_void_p(dk.data_ptr()), C.c_void_p(dv.data_ptr()), len(keys), None)
    assert rc == 0, (name, rc)
    torch.cuda.synchronize()
    return dk.cpu().numpy(), dv.cpu().numpy()


@pytest.mark.parametrize("sizes", [(300001, 1000), (LIMIT, 4097)])
def test_noalloc_equals_allocating_sort(gpu, sizes):
    for i, n in enumerate(sizes):
        keys = _keys(n, 100 * n + i)
        print("[gpu-noalloc] n=%d" % n, flush=True)
        k0, v0 = _sort("segclip_stable_sort_by_key_f64", keys)
        k1, v1 = _sort("segclip_stable_sort_by_key_f64_noalloc", keys)
        assert np.array_equal(v1, v0)
        assert np.array_equal(k1.view(np.uint64), k0.view(np.uint64))
        assert np.array_equal(v1, np.argsort(keys, kind="stable").astype(np.int32))


def test_noalloc_rejects_beyond_limit(gpu):
    from segclip import lib
    from segclip._lib import ERANGE
    # rejected by the size check before any device work: no buffer is touched
    assert lib().segclip_stable_sort_by_key_f64_noalloc(C.c_void_p(16), C.c_void_p(16), LIMIT + 1, None) == ERANGE
