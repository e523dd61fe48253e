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
#include <mutex>
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
Scratch &thread_scratch() {
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
        Scratch &sc_ = thread_scratch();
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

// ---------------------------------------------------------------------------
// Allocation-free variant of the f64 stable sort (no counterpart in the
// reference): every buffer beyond the caller's keys and values lives in one
// file-scope device array, `scratch`, of a fixed size for NOALLOC_MAX keys.
// Each GPU holds its own copy of the array until the process ends.  Host code
// never passes the array's name to a kernel (on the host it is a stand-in,
// not a device address): the device address is looked up once per device
// with hipGetSymbolAddress, with that device current, and kept per device.
// Every call on a GPU shares that GPU's array and leaves its contents
// behind; what has to start at zero (histograms, tile counters, look-back
// status words) is cleared by fill kernels at the start of each call.  Calls
// that overlap on one GPU race on the array.
// ---------------------------------------------------------------------------
namespace {

constexpr int64_t NOALLOC_MAX = int64_t(1) << 20;
// layout, bytes: keys and alternates (8+8), values and alternates (4+4),
// gathered keys and values (8+4), per key; then the radix work words
constexpr size_t NA_KEYS = 0;
constexpr size_t NA_KALT = NA_KEYS + 8 * size_t(NOALLOC_MAX);
constexpr size_t NA_IDX = NA_KALT + 8 * size_t(NOALLOC_MAX);
constexpr size_t NA_VALT = NA_IDX + 4 * size_t(NOALLOC_MAX);
constexpr size_t NA_TK = NA_VALT + 4 * size_t(NOALLOC_MAX);
constexpr size_t NA_TV = NA_TK + 8 * size_t(NOALLOC_MAX);
constexpr size_t NA_WORK = NA_TV + 4 * size_t(NOALLOC_MAX);
// radix work words for NOALLOC_MAX keys: 8 x 256 histograms, then 256
// two-word status words per radix tile
constexpr size_t NA_WORK_WORDS = 8 * 256 + 2 * size_t(NOALLOC_MAX / sc::RADIX_TILE_KEYS) * 256;
constexpr size_t NA_BYTES = NA_WORK + 4 * NA_WORK_WORDS;

constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mu;
unsigned char *na_ptr[NA_MAX_DEVICES] = {};

} // namespace

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {
// device address of this device's copy of `scratch` (looked up on the first call per device)
unsigned char *noalloc_scratch(int dev) {
    SC_REQUIRE(dev >= 0 && dev < NA_MAX_DEVICES, sc::SC_EINVAL, "device index");
    std::lock_guard<std::mutex> g(na_mu);
    if (!na_ptr[dev]) {
        void *ptr = nullptr;
        SC_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(::scratch)));
        SC_REQUIRE(ptr != nullptr, sc::SC_EHIP, "hipGetSymbolAddress gave no address for the scratch array");
        na_ptr[dev] = static_cast<unsigned char *>(ptr);
    }
    return na_ptr[dev];
}
} // namespace

extern "C" {

int segclip_stable_sort_by_key_f64_noalloc(double *keys, int32_t *vals, int64_t n, void *stream) {
    if (n < 0 || (n > 0 && (!keys || !vals))) return SEGCLIP_EINVAL;
    if (n > NOALLOC_MAX) return SEGCLIP_ERANGE;
    if (n <= 1) return SEGCLIP_OK;
    try {
        hipStream_t s = static_cast<hipStream_t>(stream);
        int dev = 0;
        SC_HIP(hipGetDevice(&dev));
        unsigned char *base = noalloc_scratch(dev);
        SC_REQUIRE(sc::radix_fixed_words(n) <= NA_WORK_WORDS, sc::SC_ERANGE, "radix work beyond the scratch array");
        uint64_t *kk = reinterpret_cast<uint64_t *>(base + NA_KEYS);
        uint64_t *kalt = reinterpret_cast<uint64_t *>(base + NA_KALT);
        uint32_t *ix = reinterpret_cast<uint32_t *>(base + NA_IDX);
        uint32_t *valt = reinterpret_cast<uint32_t *>(base + NA_VALT);
        double *tk = reinterpret_cast<double *>(base + NA_TK);
        int32_t *tv = reinterpret_cast<int32_t *>(base + NA_TV);
        uint32_t *work = reinterpret_cast<uint32_t *>(base + NA_WORK);
        k_keys_f64<<<grid_for(n), BLK, 0, s>>>(keys, n, kk, ix);
        SC_HIP(hipGetLastError());
        if (sc::radix_sort_u64_fixed(kk, ix, kalt, valt, n, 0, 64, work, s)) ix = valt;
        k_gather<double><<<grid_for(n), BLK, 0, s>>>(keys, vals, ix, n, tk, tv);
        SC_HIP(hipGetLastError());
        SC_HIP(hipMemcpyAsync(keys, tk, sizeof(double) * size_t(n), hipMemcpyDeviceToDevice, s));
        SC_HIP(hipMemcpyAsync(vals, tv, sizeof(int32_t) * size_t(n), hipMemcpyDeviceToDevice, s));
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
