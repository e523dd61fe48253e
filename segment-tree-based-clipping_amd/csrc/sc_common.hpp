// sc_common.hpp -- shared host/device helpers of libsegclip (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace sc {

// Error raised by any failing HIP call or violated precondition; the C-ABI
// layer turns it into a status code plus a message (segclip_last_error).
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
    SC_OK = 0,
    SC_EINVAL = 1,    // bad argument / input
    SC_EHIP = 2,      // HIP runtime failure
    SC_ENOMEM = 3,    // device allocation failed
    SC_ESTATE = 4,    // call out of order
    SC_ERANGE = 5,    // size beyond what the kernels support
    SC_EUNDEF = 6,    // the reference's behaviour is undefined on this input
};

#define SC_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw ::sc::Error(::sc::SC_EHIP, std::string(#call) + ": " + hipGetErrorString(e_) + \
                                                 " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

#define SC_REQUIRE(cond, code, msg)                   \
    do {                                              \
        if (!(cond)) throw ::sc::Error((code), (msg)); \
    } while (0)

// Device bytes held by the library's buffers (all contexts) and their peak.
inline std::atomic<int64_t> &dev_bytes() {
    static std::atomic<int64_t> v{0};
    return v;
}
inline std::atomic<int64_t> &dev_peak() {
    static std::atomic<int64_t> v{0};
    return v;
}
inline void dev_account(int64_t delta) {
    int64_t now = dev_bytes().fetch_add(delta) + delta;
    int64_t pk = dev_peak().load();
    while (now > pk && !dev_peak().compare_exchange_weak(pk, now)) {
    }
}

// Optional device allocator supplied by the embedding program (for instance
// a framework's caching allocator), used for every buffer allocated while it
// is set.  The library waits for all of its device work before handing a
// block back, so the allocator may reuse it at once on any stream.
struct AllocHooks {
    void *(*alloc)(size_t bytes, void *user) = nullptr;
    void (*free)(void *ptr, void *user) = nullptr;
    void *user = nullptr;
};
inline AllocHooks &alloc_hooks() {
    static AllocHooks h;
    return h;
}

// Growable device buffer; contents are not preserved across growth.
template <class T> struct DBuf {
    T *p = nullptr;
    size_t cap = 0;
    AllocHooks from; // allocator of p (alloc == nullptr: hipMalloc)
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) {
            if (from.alloc) {
                (void)hipDeviceSynchronize(); // work still reading p may be queued
                from.free(p, from.user);
            } else {
                (void)hipFree(p);
            }
            dev_account(-int64_t(cap * sizeof(T)));
        }
        p = nullptr;
        cap = 0;
    }
    T *ensure(size_t n) {
        if (n == 0) n = 1;
        if (n > cap) {
            release();
            size_t want = n + n / 8;
            const AllocHooks &h = alloc_hooks();
            if (h.alloc) {
                p = static_cast<T *>(h.alloc(want * sizeof(T), h.user));
                if (!p) throw Error(SC_ENOMEM, "allocation of " + std::to_string(want * sizeof(T)) + " bytes failed");
                from = h;
            } else {
                hipError_t e = hipMalloc(&p, want * sizeof(T));
                if (e != hipSuccess) {
                    p = nullptr;
                    (void)hipGetLastError();
                    throw Error(SC_ENOMEM, "hipMalloc of " + std::to_string(want * sizeof(T)) + " bytes failed");
                }
                from = AllocHooks();
            }
            cap = want;
            dev_account(int64_t(want * sizeof(T)));
        }
        return p;
    }
    // exchange storage (pointer, capacity, allocator) with another buffer
    void swap(DBuf &o) {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
        std::swap(from, o.from);
    }
    operator T *() const { return p; }
};

constexpr int WAVE = 64;

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

inline int next_pow2_int(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}
inline int ilog2_exact(int64_t n) {
    int l = 0;
    while ((int64_t(1) << l) < n) l++;
    return l;
}

// Index of the segment holding i: the last j in [0,n) with off[j] <= i, for
// non-decreasing offsets (empty segments are skipped).
__device__ inline int seg_of(const int *off, int n, int64_t i) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// primitives (sc_prims.hip)
// ---------------------------------------------------------------------------

// Exclusive prefix sum of n values into out (in == out allowed); returns the
// total through *d_total when non-null (device pointer).
void scan_exclusive_u32(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *d_total, hipStream_t s);
void scan_exclusive_u64(const uint64_t *in, uint64_t *out, int64_t n, uint64_t *d_total, hipStream_t s);
// Sum of n values into *d_total (device pointer).
void sum_u64(const uint64_t *in, int64_t n, uint64_t *d_total, hipStream_t s);
// Inclusive running maximum of n values (in == out allowed).
void scan_inclusive_max_u32(const uint32_t *in, uint32_t *out, int64_t n, hipStream_t s);
// two independent arrays in the launches of one scan (same results as two calls)
void scan_exclusive_u32_2(const uint32_t *in0, uint32_t *out0, int64_t n0, uint32_t *t0, const uint32_t *in1,
                          uint32_t *out1, int64_t n1, uint32_t *t1, hipStream_t s);
void scan_inclusive_max_u32_2(const uint32_t *in0, uint32_t *out0, int64_t n0, const uint32_t *in1,
                              uint32_t *out1, int64_t n1, hipStream_t s);

// Several 32-bit spans set to a value each, in one launch: add() queues a
// span and launches when FILL_MAX are queued, flush() launches the rest (the
// destructor does not: call flush() before the spans are read).
constexpr int FILL_MAX = 8;
struct FillBatch {
    hipStream_t s;
    int count = 0;
    void *p[FILL_MAX];
    int64_t n[FILL_MAX];
    uint32_t v[FILL_MAX];
    explicit FillBatch(hipStream_t stream) : s(stream) {}
    void add(void *ptr, int64_t words, uint32_t value) {
        if (words <= 0) return;
        p[count] = ptr;
        n[count] = words;
        v[count] = value;
        if (++count == FILL_MAX) flush();
    }
    void flush();
};

// Stable LSD radix sort of (keys, vals) by key bits [bit_lo, bit_hi).
// vals may be null (keys only).  The sorted data ends in keys/vals; the
// alternate buffers are scratch of the same length.
struct RadixScratch {
    DBuf<uint32_t> global_hist;            // per pass 256 digit counts and a tile counter
    DBuf<unsigned long long> tile_status;  // chained scan: per tile and digit, epoch | flag | count
    size_t status_clean = 0;               // tile_status entries known to hold no current epoch
    uint32_t epoch = 0;                    // tag of the last pass's status words
    DBuf<uint64_t> keys_alt;
    DBuf<uint32_t> vals_alt;
};
void radix_sort_u64(uint64_t *keys, uint32_t *vals, int64_t n, int bit_lo, int bit_hi, RadixScratch &rs,
                    hipStream_t s);
// The same sort on keys/vals held in buffers, with alternate buffers of the
// caller's own (ensured here to n entries): when the pass count is odd the
// sorted data ends in the alternates, and each is exchanged with its partner
// instead of copied back.  Raw pointers taken from any of the four before the
// call are stale after it.  A call site that ensures keys beyond n ensures
// keys_alt to the same size, so neither partner has to grow again.
void radix_sort_u64_buf(DBuf<uint64_t> &keys, DBuf<uint64_t> &keys_alt, DBuf<uint32_t> *vals,
                        DBuf<uint32_t> *vals_alt, int64_t n, int bit_lo, int bit_hi, RadixScratch &rs, hipStream_t s);

// The same sort on caller-provided memory only (no allocation): work holds
// radix_fixed_words(n) 32-bit words, cleared here by fill kernels; kalt/valt
// are scratch of n entries.  Returns true when the sorted data ends in
// kalt/valt.  No host wait.
// keys per tile of one radix pass (a block of 256 threads, 16 keys each)
constexpr int RADIX_TILE_KEYS = 4096;
size_t radix_fixed_words(int64_t n);
bool radix_sort_u64_fixed(uint64_t *keys, uint32_t *vals, uint64_t *kalt, uint32_t *valt, int64_t n, int bit_lo,
                          int bit_hi, uint32_t *work, hipStream_t s);

// Order-preserving map of an IEEE double to uint64 (-0.0 and +0.0 map to the
// same key, so a stable sort on it orders ties by original position as a
// comparison sort with operator< does).
__host__ __device__ inline uint64_t f64_order_key(double x) {
    if (x == 0.0) x = 0.0;
    uint64_t u;
    __builtin_memcpy(&u, &x, 8);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}

} // namespace sc
