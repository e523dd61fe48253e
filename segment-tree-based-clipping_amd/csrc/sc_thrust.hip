// sc_thrust.hip -- the primitives of the reference's thrust_func.h
// (p_seg_tree/util/thrust_func.h:4-12, p_seg_tree/thrust_func.cu:19-42) on
// libsegclip's own radix sort and scan: stable sort-by-key of REAL / int /
// LONG keys with int payload, exclusive scan of int / long, all in place on
// device pointers.  Exposed through the C ABI (segclip_*) and, for C++
// callers compiled against the reference header, through the reference's
// overloaded function names.
#include "../../include/segclip.h"
#include "../../include/thrust_func.h"
#include "sc_common.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

constexpr int BLK = 256;
inline unsigned grid_for(int64_t n) { return unsigned(sc::ceil_div(n > 0 ? n : 1, BLK)); }

// per-thread scratch, intentionally never freed (see sc_prims.hip)
struct Scratch {
    sc::RadixScratch rs;
    sc::DBuf<uint64_t> keys, tmp;
    sc::DBuf<uint32_t> idx;
};
Scratch &scratch() {
    thread_local Scratch *p = new Scratch();
    return *p;
}

__global__ void k_keys_f64(const double *k, int64_t n, uint64_t *o, uint32_t *idx) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        o[i] = sc::f64_order_key(k[i]);
        idx[i] = uint32_t(i);
    }
}
__global__ void k_keys_i32(const int32_t *k, int64_t n, uint64_t *o, uint32_t *idx) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        o[i] = uint64_t(uint32_t(k[i]) ^ 0x80000000u);
        idx[i] = uint32_t(i);
    }
}
__global__ void k_keys_i64(const int64_t *k, int64_t n, uint64_t *o, uint32_t *idx) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        o[i] = uint64_t(k[i]) ^ 0x8000000000000000ull;
        idx[i] = uint32_t(i);
    }
}
// gather keys and values through the sorted permutation into tmp, then copy back
template <class K>
__global__ void k_gather(const K *keys, const int32_t *vals, const uint32_t *idx, int64_t n, K *tk, int32_t *tv) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        tk[i] = keys[idx[i]];
        tv[i] = vals[idx[i]];
    }
}

template <class K, class KeyK>
int sort_by_key(K *keys, int32_t *vals, int64_t n, void *stream, KeyK keyk, int bits) {
    if (n < 0 || (n > 0 && (!keys || !vals))) return SEGCLIP_EINVAL;
    if (n <= 1) return SEGCLIP_OK;
    try {
        hipStream_t s = static_cast<hipStream_t>(stream);
        Scratch &sc_ = scratch();
        uint64_t *kk = sc_.keys.ensure(size_t(n));
        uint32_t *ix = sc_.idx.ensure(size_t(n));
        keyk<<<grid_for(n), BLK, 0, s>>>(keys, n, kk, ix);
        SC_HIP(hipGetLastError());
        sc::radix_sort_u64(kk, ix, n, 0, bits, sc_.rs, s);
        // temporary space for the gathered keys and values
        size_t words = (size_t(n) * sizeof(K) + size_t(n) * sizeof(int32_t) + 7) / 8 + 1;
        uint64_t *tmp = sc_.tmp.ensure(words);
        K *tk = reinterpret_cast<K *>(tmp);
        int32_t *tv = reinterpret_cast<int32_t *>(reinterpret_cast<char *>(tmp) + size_t(n) * sizeof(K));
        k_gather<K><<<grid_for(n), BLK, 0, s>>>(keys, vals, ix, n, tk, tv);
        SC_HIP(hipGetLastError());
        SC_HIP(hipMemcpyAsync(keys, tk, sizeof(K) * size_t(n), hipMemcpyDeviceToDevice, s));
        SC_HIP(hipMemcpyAsync(vals, tv, sizeof(int32_t) * size_t(n), hipMemcpyDeviceToDevice, s));
        SC_HIP(hipStreamSynchronize(s));
        return SEGCLIP_OK;
    } catch (const sc::Error &e) {
        return e.code;
    }
}

} // namespace

extern "C" {

int segclip_stable_sort_by_key_f64(double *keys, int32_t *vals, int64_t n, void *stream) {
    return sort_by_key<double>(keys, vals, n, stream, k_keys_f64, 64);
}
int segclip_stable_sort_by_key_i32(int32_t *keys, int32_t *vals, int64_t n, void *stream) {
    return sort_by_key<int32_t>(keys, vals, n, stream, k_keys_i32, 32);
}
int segclip_stable_sort_by_key_i64(int64_t *keys, int32_t *vals, int64_t n, void *stream) {
    return sort_by_key<int64_t>(keys, vals, n, stream, k_keys_i64, 64);
}
int segclip_exclusive_scan_i32(int32_t *data, int64_t n, void *stream) {
    if (n < 0 || (n > 0 && !data)) return SEGCLIP_EINVAL;
    try {
        hipStream_t s = static_cast<hipStream_t>(stream);
        sc::scan_exclusive_u32(reinterpret_cast<uint32_t *>(data), reinterpret_cast<uint32_t *>(data), n, nullptr, s);
        SC_HIP(hipStreamSynchronize(s));
        return SEGCLIP_OK;
    } catch (const sc::Error &e) {
        return e.code;
    }
}
int segclip_exclusive_scan_i64(int64_t *data, int64_t n, void *stream) {
    if (n < 0 || (n > 0 && !data)) return SEGCLIP_EINVAL;
    try {
        hipStream_t s = static_cast<hipStream_t>(stream);
        sc::scan_exclusive_u64(reinterpret_cast<uint64_t *>(data), reinterpret_cast<uint64_t *>(data), n, nullptr, s);
        SC_HIP(hipStreamSynchronize(s));
        return SEGCLIP_OK;
    } catch (const sc::Error &e) {
        return e.code;
    }
}

} // extern "C"

// Overloads with the reference's names and signatures (REAL = double,
// LONG = long), default stream, for callers built against thrust_func.h.
// A failed call throws std::runtime_error (Thrust throws
// thrust::system_error, also a std::runtime_error), so a sort that did not
// happen is never silent.
namespace {
void throw_if(int rc, const char *what) {
    if (rc != SEGCLIP_OK)
        throw std::runtime_error(std::string(what) + " failed with libsegclip status " + std::to_string(rc));
}
} // namespace

void thrustDeviceStableSortByKey(double *arr, int *arrIdx, int size) {
    throw_if(segclip_stable_sort_by_key_f64(arr, arrIdx, size, nullptr), "thrustDeviceStableSortByKey(REAL)");
}
void thrustDeviceStableSortByKey(int *arr, int *arrIdx, int size) {
    throw_if(segclip_stable_sort_by_key_i32(arr, arrIdx, size, nullptr), "thrustDeviceStableSortByKey(int)");
}
void thrustDeviceStableSortByKey(long *arr, int *arrIdx, int size) {
    throw_if(segclip_stable_sort_by_key_i64(reinterpret_cast<int64_t *>(arr), arrIdx, size, nullptr),
             "thrustDeviceStableSortByKey(LONG)");
}
void thrustDeviceExclusiveScan(int *arr, int size) {
    throw_if(segclip_exclusive_scan_i32(arr, size, nullptr), "thrustDeviceExclusiveScan(int)");
}
void thrustDeviceExclusiveScan(long *arr, int size) {
    throw_if(segclip_exclusive_scan_i64(reinterpret_cast<int64_t *>(arr), size, nullptr),
             "thrustDeviceExclusiveScan(long)");
}
// Host stable sort of (REAL key, int payload), thrust_func.cu:40-42: the
// payload follows its key, ties keep their order.
void thrustHostStableSortByKey(double *arr, int *arrIdx, int size) {
    if (size <= 1) return;
    std::vector<int> perm(static_cast<size_t>(size));
    for (int i = 0; i < size; ++i) perm[size_t(i)] = i;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return arr[a] < arr[b]; });
    std::vector<double> k(static_cast<size_t>(size));
    std::vector<int> v(static_cast<size_t>(size));
    for (int i = 0; i < size; ++i) {
        k[size_t(i)] = arr[perm[size_t(i)]];
        v[size_t(i)] = arrIdx[perm[size_t(i)]];
    }
    std::copy(k.begin(), k.end(), arr);
    std::copy(v.begin(), v.end(), arrIdx);
}
