"""GPU checks of the thrust_func.h drop-ins (stable sort by key, exclusive
scan) against numpy's stable sort and cumulative sum: ties keep their input
order, -0.0 equals 0.0, negative keys, empty and one-element inputs, sizes
spanning many radix tiles."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _sort(name, keys):
    import torch
    from segclip import lib
    n = len(keys)
    vals = np.arange(n, dtype=np.int32)
    dk, dv = _dev(keys), _dev(vals)
    torch.cuda.synchronize()
    rc = getattr(lib(), name)(C.c_void_p(dk.data_ptr()), C.c_void_p(dv.data_ptr()), n, None)
    assert rc == 0
    torch.cuda.synchronize()
    return dk.cpu().numpy(), dv.cpu().numpy()


@pytest.mark.parametrize("n", [0, 1, 2, 1000, 3072, 3073, 4096, 4097, 300001])
def test_stable_sort_f64(gpu, n):
    rng = np.random.default_rng(n)
    keys = np.round(rng.standard_normal(n) * 50) / 4.0  # many ties, negatives
    if n > 4:
        keys[:4] = [0.0, -0.0, 0.0, -0.0]
    k, v = _sort("segclip_stable_sort_by_key_f64", keys)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(v, order.astype(np.int32))
    assert np.array_equal(k, keys[order])


@pytest.mark.parametrize("n", [0, 1, 5000, 200003, (1 << 21) + 7])
def test_stable_sort_i32_i64(gpu, n):
    rng = np.random.default_rng(100 + n)
    k32 = rng.integers(-1000, 1000, n).astype(np.int32)
    k, v = _sort("segclip_stable_sort_by_key_i32", k32)
    assert np.array_equal(v, np.argsort(k32, kind="stable").astype(np.int32))
    k64 = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    k64[::7] = 12345
    k, v = _sort("segclip_stable_sort_by_key_i64", k64)
    assert np.array_equal(v, np.argsort(k64, kind="stable").astype(np.int32))


@pytest.mark.parametrize("n", [0, 1, 777, 2048, 2049, 1 << 20, 5_000_011])
def test_exclusive_scan(gpu, n):
    import torch
    from segclip import lib
    rng = np.random.default_rng(200 + n)
    a32 = rng.integers(-50, 100, n).astype(np.int32)
    d = _dev(a32)
    torch.cuda.synchronize()
    assert lib().segclip_exclusive_scan_i32(C.c_void_p(d.data_ptr()), n, None) == 0
    torch.cuda.synchronize()
    want = np.concatenate([[0], np.cumsum(a32, dtype=np.int64)[:-1]]).astype(np.int32) if n else a32
    assert np.array_equal(d.cpu().numpy(), want)
    a64 = rng.integers(-(1 << 33), 1 << 33, n).astype(np.int64)
    d = _dev(a64)
    torch.cuda.synchronize()
    assert lib().segclip_exclusive_scan_i64(C.c_void_p(d.data_ptr()), n, None) == 0
    torch.cuda.synchronize()
    want = np.concatenate([[0], np.cumsum(a64)[:-1]]).astype(np.int64) if n else a64
    assert np.array_equal(d.cpu().numpy(), want)


def test_exclusive_scan_streams_back_to_back(gpu):
    """Multi-tile scans on two streams in turn, a dozen on each: every result
    equals numpy's (each stream keeps its own tile status words and tickets,
    so the tickets of one stream's scans never count the other's)."""
    import torch
    from segclip import lib
    rng = np.random.default_rng(7)
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    inputs, outs = [], []
    torch.cuda.synchronize()
    for k in range(24):
        n = int(rng.integers(2049, 300000))
        a = rng.integers(0, 1000, n).astype(np.int32)
        s = streams[k % 2]
        with torch.cuda.stream(s):
            d = torch.from_numpy(a).to("cuda", non_blocking=False)
        s.synchronize()
        assert lib().segclip_exclusive_scan_i32(C.c_void_p(d.data_ptr()), n, C.c_void_p(s.cuda_stream)) == 0
        inputs.append(a)
        outs.append(d)
    torch.cuda.synchronize()
    for a, d in zip(inputs, outs):
        want = np.concatenate([[0], np.cumsum(a, dtype=np.int64)[:-1]]).astype(np.int32)
        assert np.array_equal(d.cpu().numpy(), want)
