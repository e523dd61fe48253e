// sc_prims.hip -- device-wide scan and stable LSD radix sort for gfx950.
//
// Scan: u32 sums and maxima in two launches per scan (tile totals, then
// each tile combines the totals of the tiles before it itself), up to four
// arrays per launch pair; u64 scans as reduce / scan-of-block-sums /
// downsweep.  Wave64 shuffles inside a block.
//
// Radix sort: 8-bit digits, 4096 keys per tile, one launch per pass.  One
// pass over the keys first builds every digit histogram at once, so a pass
// whose digit is the same for all keys only copies its tiles.  The scatter ranks keys stably inside a tile with
// a 64-lane match (eight ballots per key) and per-wave digit counts in LDS.
#include "sc_common.hpp"

namespace sc {

namespace {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

template <class T> __device__ inline T wave_incl_scan(T x, int lane) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        T y = __shfl_up(x, d, WAVE);
        if (lane >= d) x += y;
    }
    return x;
}

// Exclusive scan of one thread value over the block; returns the block total.
template <class T> __device__ inline T block_excl_scan(T x, T *s_wave, T &total) {
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    T inc = wave_incl_scan(x, lane);
    if (lane == WAVE - 1) s_wave[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        T w = lane < SCAN_BLOCK / WAVE ? s_wave[lane] : T(0);
        T wi = wave_incl_scan(w, lane);
        if (lane < SCAN_BLOCK / WAVE) s_wave[lane] = wi - w;
        if (lane == SCAN_BLOCK / WAVE - 1) s_wave[SCAN_BLOCK / WAVE] = wi;
    }
    __syncthreads();
    T r = s_wave[wid] + inc - x;
    total = s_wave[SCAN_BLOCK / WAVE];
    __syncthreads();
    return r;
}

template <class T> __global__ __launch_bounds__(SCAN_BLOCK) void k_tile_reduce(const T *in, int64_t n, T *sums) {
    __shared__ T s_wave[SCAN_BLOCK / WAVE + 1];
    int64_t base = int64_t(blockIdx.x) * SCAN_TILE;
    T acc = 0;
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        int64_t i = base + int64_t(it) * SCAN_BLOCK + threadIdx.x;
        if (i < n) acc += in[i];
    }
    T total;
    (void)block_excl_scan(acc, s_wave, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// single block: exclusive scan of 'n' block sums in place, total -> *d_total
template <class T> __global__ __launch_bounds__(SCAN_BLOCK) void k_scan_sums(T *sums, int64_t n, T *d_total) {
    __shared__ T s_wave[SCAN_BLOCK / WAVE + 1];
    T carry = 0;
    for (int64_t base = 0; base < n; base += SCAN_BLOCK) {
        int64_t i = base + threadIdx.x;
        T v = i < n ? sums[i] : T(0);
        T total;
        T ex = block_excl_scan(v, s_wave, total);
        if (i < n) sums[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0 && d_total) *d_total = carry;
}

template <class T>
__global__ __launch_bounds__(SCAN_BLOCK) void k_tile_scan(const T *in, T *out, int64_t n, const T *offsets,
                                                          T *d_total_single) {
    __shared__ T s_data[SCAN_TILE];
    __shared__ T s_wave[SCAN_BLOCK / WAVE + 1];
    int64_t base = int64_t(blockIdx.x) * SCAN_TILE;
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        int64_t i = base + int64_t(it) * SCAN_BLOCK + threadIdx.x;
        s_data[it * SCAN_BLOCK + threadIdx.x] = i < n ? in[i] : T(0);
    }
    __syncthreads();
    T loc[SCAN_ITEMS];
    T acc = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        loc[j] = s_data[threadIdx.x * SCAN_ITEMS + j];
        acc += loc[j];
    }
    T total;
    T ex = block_excl_scan(acc, s_wave, total);
    T off = offsets ? offsets[blockIdx.x] : T(0);
    T run = off + ex;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        T v = loc[j];
        s_data[threadIdx.x * SCAN_ITEMS + j] = run;
        run += v;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        int64_t i = base + int64_t(it) * SCAN_BLOCK + threadIdx.x;
        if (i < n) out[i] = s_data[it * SCAN_BLOCK + threadIdx.x];
    }
    if (d_total_single && threadIdx.x == 0) *d_total_single = total;
}

template <class T> struct ScanScratch {
    DBuf<T> sums, sums2;
};

template <class T> void scan_impl(const T *in, T *out, int64_t n, T *d_total, hipStream_t s, ScanScratch<T> &ss) {
    if (n <= 0) {
        if (d_total) SC_HIP(hipMemsetAsync(d_total, 0, sizeof(T), s));
        return;
    }
    int64_t nb = ceil_div(n, SCAN_TILE);
    if (nb == 1) {
        k_tile_scan<T><<<1, SCAN_BLOCK, 0, s>>>(in, out, n, nullptr, d_total);
        SC_HIP(hipGetLastError());
        return;
    }
    SC_REQUIRE(nb < (int64_t(1) << 31), SC_ERANGE, "scan too long");
    T *sums = ss.sums.ensure(size_t(nb));
    k_tile_reduce<T><<<unsigned(nb), SCAN_BLOCK, 0, s>>>(in, n, sums);
    SC_HIP(hipGetLastError());
    k_scan_sums<T><<<1, SCAN_BLOCK, 0, s>>>(sums, nb, d_total);
    SC_HIP(hipGetLastError());
    k_tile_scan<T><<<unsigned(nb), SCAN_BLOCK, 0, s>>>(in, out, n, sums, nullptr);
    SC_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// two-launch scan of u32 (exclusive sum, inclusive max)
// ---------------------------------------------------------------------------
// The first launch reduces each 2048-item tile to one value; in the second,
// each block combines the values of the tiles before its own (at most
// PREF_TILES of them, read from L2) into its prefix and scans its tile.  This
// drops a single-block scan of the tile values, a launch whose cost is all
// latency at these sizes.  Longer inputs scan the tile values themselves
// first (two more launches of the same kernels).
constexpr int64_t PREF_TILES = 2048;

struct OpSum {
    static __device__ uint32_t id() { return 0u; }
    static __device__ uint32_t f(uint32_t a, uint32_t b) { return a + b; }
};
struct OpMax {
    static __device__ uint32_t id() { return 0u; }
    static __device__ uint32_t f(uint32_t a, uint32_t b) { return a > b ? a : b; }
};

template <class Op> __device__ inline uint32_t wave_incl_op(uint32_t x, int lane) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, WAVE);
        if (lane >= d) x = Op::f(y, x);
    }
    return x;
}

// exclusive scan of one value per thread over the block; total = the block's aggregate
template <class Op> __device__ inline uint32_t block_excl_op(uint32_t x, uint32_t *s_wave, uint32_t &total) {
    constexpr int NW = SCAN_BLOCK / WAVE;
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const uint32_t inc = wave_incl_op<Op>(x, lane);
    uint32_t ex = __shfl_up(inc, 1, WAVE);
    if (lane == 0) ex = Op::id();
    if (lane == WAVE - 1) s_wave[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        const uint32_t w = lane < NW ? s_wave[lane] : Op::id();
        const uint32_t wi = wave_incl_op<Op>(w, lane);
        uint32_t we = __shfl_up(wi, 1, WAVE);
        if (lane == 0) we = Op::id();
        if (lane < NW) s_wave[lane] = we;
        if (lane == NW - 1) s_wave[NW] = wi;
    }
    __syncthreads();
    const uint32_t r = Op::f(s_wave[wid], ex);
    total = s_wave[NW];
    __syncthreads();
    return r;
}

template <class Op>
__device__ inline void tile_reduce_body(const uint32_t *in, int64_t n, uint32_t *sums, uint32_t tile,
                                        uint32_t *s_wave) {
    const int64_t base = int64_t(tile) * SCAN_TILE;
    uint32_t acc = Op::id();
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int64_t i = base + int64_t(it) * SCAN_BLOCK + threadIdx.x;
        if (i < n) acc = Op::f(acc, in[i]);
    }
    uint32_t total;
    (void)block_excl_op<Op>(acc, s_wave, total);
    if (threadIdx.x == 0) sums[tile] = total;
}

// INCL: inclusive output (the max scan), else exclusive (the sum scan).
// In place (in == out) is safe: a block reads only its own tile's inputs.
// The tiles before this one come either as their values (sums, at most
// PREF_TILES of them, combined here) or already combined (pre[tile]).
template <class Op, bool INCL>
__device__ inline void tile_scan_body(const uint32_t *in, uint32_t *out, int64_t n, const uint32_t *sums,
                                      const uint32_t *pre, uint32_t *d_total, uint32_t tile, uint32_t *s_data,
                                      uint32_t *s_wave) {
    const int64_t base = int64_t(tile) * SCAN_TILE;
    uint32_t pacc = Op::id();
    if (!pre)
        for (uint32_t i = threadIdx.x; i < tile; i += SCAN_BLOCK) pacc = Op::f(pacc, sums[i]);
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int64_t i = base + int64_t(it) * SCAN_BLOCK + threadIdx.x;
        s_data[it * SCAN_BLOCK + threadIdx.x] = i < n ? in[i] : Op::id();
    }
    uint32_t prefix; // the tiles before this one
    if (pre) {
        prefix = pre[tile];
        __syncthreads();
    } else {
        (void)block_excl_op<Op>(pacc, s_wave, prefix);
    }
    uint32_t loc[SCAN_ITEMS];
    uint32_t acc = Op::id();
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        loc[j] = s_data[threadIdx.x * SCAN_ITEMS + j];
        acc = Op::f(acc, loc[j]);
    }
    uint32_t total;
    const uint32_t ex = block_excl_op<Op>(acc, s_wave, total);
    uint32_t run = Op::f(prefix, ex);
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        const uint32_t v = loc[j];
        if (INCL) run = Op::f(run, v);
        s_data[threadIdx.x * SCAN_ITEMS + j] = run;
        if (!INCL) run = Op::f(run, v);
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int64_t i = base + int64_t(it) * SCAN_BLOCK + threadIdx.x;
        if (i < n) out[i] = s_data[it * SCAN_BLOCK + threadIdx.x];
    }
    if (d_total && threadIdx.x == 0 && base + SCAN_TILE >= n) *d_total = Op::f(prefix, total);
}

template <class Op>
__global__ __launch_bounds__(SCAN_BLOCK) void k_tile_reduce_op(const uint32_t *in, int64_t n, uint32_t *sums) {
    __shared__ uint32_t s_wave[SCAN_BLOCK / WAVE + 1];
    tile_reduce_body<Op>(in, n, sums, blockIdx.x, s_wave);
}

template <class Op, bool INCL>
__global__ __launch_bounds__(SCAN_BLOCK) void k_tile_scan_op(const uint32_t *in, uint32_t *out, int64_t n,
                                                              const uint32_t *sums, const uint32_t *pre,
                                                              uint32_t *d_total) {
    __shared__ uint32_t s_data[SCAN_TILE];
    __shared__ uint32_t s_wave[SCAN_BLOCK / WAVE + 1];
    tile_scan_body<Op, INCL>(in, out, n, sums, pre, d_total, blockIdx.x, s_data, s_wave);
}

// Several independent arrays scanned by the same two launches: blockIdx.y
// picks the array, blocks past its last tile return at once.
constexpr int SCAN_SET = 4;
struct ScanItem {
    const uint32_t *in;
    uint32_t *out;
    int64_t n;
    uint32_t *d_total; // or null
    uint32_t *sums;    // its tiles' values
};
struct ScanSet {
    ScanItem it[SCAN_SET];
};

template <class Op> __global__ __launch_bounds__(SCAN_BLOCK) void k_tile_reduce_set(ScanSet S) {
    __shared__ uint32_t s_wave[SCAN_BLOCK / WAVE + 1];
    const ScanItem q = S.it[blockIdx.y];
    if (int64_t(blockIdx.x) * SCAN_TILE >= q.n) return;
    tile_reduce_body<Op>(q.in, q.n, q.sums, blockIdx.x, s_wave);
}

template <class Op, bool INCL> __global__ __launch_bounds__(SCAN_BLOCK) void k_tile_scan_set(ScanSet S) {
    __shared__ uint32_t s_data[SCAN_TILE];
    __shared__ uint32_t s_wave[SCAN_BLOCK / WAVE + 1];
    const ScanItem q = S.it[blockIdx.y];
    if (int64_t(blockIdx.x) * SCAN_TILE >= q.n) return;
    tile_scan_body<Op, INCL>(q.in, q.out, q.n, q.sums, nullptr, q.d_total, blockIdx.x, s_data, s_wave);
}

// Per host thread and stream, intentionally never freed: releasing device
// memory from thread-exit destructors can run after the HIP runtime has shut
// down.  Keying by stream keeps scans queued on different streams from
// sharing tile sums.
struct StreamScanScratch {
    hipStream_t stream = nullptr;
    ScanScratch<uint32_t> s32;
    ScanScratch<uint64_t> s64;
};

StreamScanScratch &stream_scratch(hipStream_t s) {
    thread_local std::vector<StreamScanScratch *> *v = new std::vector<StreamScanScratch *>();
    for (StreamScanScratch *p : *v)
        if (p->stream == s) return *p;
    StreamScanScratch *p = new StreamScanScratch();
    p->stream = s;
    v->push_back(p);
    return *p;
}

template <class T> ScanScratch<T> &scan_scratch(hipStream_t s);
template <> ScanScratch<uint32_t> &scan_scratch<uint32_t>(hipStream_t s) { return stream_scratch(s).s32; }
template <> ScanScratch<uint64_t> &scan_scratch<uint64_t>(hipStream_t s) { return stream_scratch(s).s64; }

template <class Op, bool INCL>
void scan_u32_op(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *d_total, hipStream_t s) {
    if (n <= 0) {
        if (d_total) SC_HIP(hipMemsetAsync(d_total, 0, sizeof(uint32_t), s));
        return;
    }
    const int64_t nb = ceil_div(n, SCAN_TILE);
    SC_REQUIRE(nb <= PREF_TILES * SCAN_TILE, SC_ERANGE, "scan too long");
    if (nb == 1) {
        k_tile_scan_op<Op, INCL><<<1, SCAN_BLOCK, 0, s>>>(in, out, n, nullptr, nullptr, d_total);
        SC_HIP(hipGetLastError());
        return;
    }
    ScanScratch<uint32_t> &ss = scan_scratch<uint32_t>(s);
    uint32_t *sums = ss.sums.ensure(size_t(nb));
    k_tile_reduce_op<Op><<<unsigned(nb), SCAN_BLOCK, 0, s>>>(in, n, sums);
    SC_HIP(hipGetLastError());
    if (nb <= PREF_TILES) {
        k_tile_scan_op<Op, INCL><<<unsigned(nb), SCAN_BLOCK, 0, s>>>(in, out, n, sums, nullptr, d_total);
        SC_HIP(hipGetLastError());
        return;
    }
    // more tiles than one block combines: exclusive scan of the tile values in place
    const int64_t nb2 = ceil_div(nb, SCAN_TILE); // <= PREF_TILES
    uint32_t *sums2 = ss.sums2.ensure(size_t(nb2));
    k_tile_reduce_op<Op><<<unsigned(nb2), SCAN_BLOCK, 0, s>>>(sums, nb, sums2);
    SC_HIP(hipGetLastError());
    k_tile_scan_op<Op, false><<<unsigned(nb2), SCAN_BLOCK, 0, s>>>(sums, sums, nb, sums2, nullptr, nullptr);
    SC_HIP(hipGetLastError());
    k_tile_scan_op<Op, INCL><<<unsigned(nb), SCAN_BLOCK, 0, s>>>(in, out, n, nullptr, sums, d_total);
    SC_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// radix sort
// ---------------------------------------------------------------------------
constexpr int RS_BLOCK = 256;
constexpr int RS_ITEMS = RADIX_TILE_KEYS / RS_BLOCK;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;
constexpr int RS_WAVES = RS_BLOCK / WAVE;
constexpr int RS_MAXPASS = 8;

__global__ __launch_bounds__(RS_BLOCK) void k_global_hist(const uint64_t *keys, int64_t n, int bit_lo, int bit_hi,
                                                          int npass, uint32_t *ghist) {
    __shared__ uint32_t h[RS_MAXPASS][256];
    for (int i = threadIdx.x; i < RS_MAXPASS * 256; i += RS_BLOCK) (&h[0][0])[i] = 0;
    __syncthreads();
    for (int64_t i = int64_t(blockIdx.x) * RS_BLOCK + threadIdx.x; i < n; i += int64_t(gridDim.x) * RS_BLOCK) {
        uint64_t k = keys[i];
        for (int p = 0; p < npass; ++p) {
            int sh = bit_lo + 8 * p;
            int bits = bit_hi - sh < 8 ? bit_hi - sh : 8;
            uint32_t d = uint32_t(k >> sh) & ((1u << bits) - 1u);
            atomicAdd(&h[p][d], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < npass * 256; i += RS_BLOCK) {
        uint32_t v = (&h[0][0])[i];
        if (v) atomicAdd(&ghist[i], v);
    }
}

// One pass of the LSD radix sort in one launch (a chained scan over tiles).
// Tile t is block t: a tile waits only for tiles of lower index, and every
// XCD dispatches its blocks in index order, so the lowest block not yet
// dispatched waits for nothing that does not finish (a ticket counter
// instead costs one contended atomic per tile; a wait is also bounded, see
// the look-back below).  Wave w of the block owns
// the RS_TILE/RS_WAVES consecutive keys starting at w*RS_TILE/RS_WAVES, 64 at
// a time; keys are ranked within their wave and digit (peers found with
// eight ballots, a per-wave digit counter in LDS) and a block scan over
// (digit, wave) counters gives every key its place in the tile's digit
// order.  The tile publishes its per-digit counts and is sorted into LDS in
// digit order; then thread d walks back over earlier tiles' published words
// until one carries an inclusive prefix, and publishes its own, and each
// digit's run is written out contiguously at global digit base + prefix.
// Within a digit keys keep their input order (the sort is stable).
constexpr int RS_PER_WAVE = RS_TILE / RS_WAVES;
constexpr int RS_IT = RS_PER_WAVE / WAVE; // = RS_ITEMS
static_assert(RS_BLOCK == 256, "one thread per 8-bit digit");
static_assert(RS_TILE <= 65535, "16-bit tile offsets");
// status word: epoch (8 bits) | flag (2 bits: aggregate, inclusive prefix) | count (54 bits);
// a word of another epoch is not yet written for this pass
constexpr int ST_EPOCH_SHIFT = 56;
constexpr unsigned long long ST_AGG = 1ull << 54, ST_PRE = 2ull << 54, ST_VAL = (1ull << 54) - 1;
constexpr int RS_SPIN_LIMIT = 1 << 20; // polls of one predecessor word before giving up
constexpr int RS_LOOK = 8;             // predecessor words read per round trip

struct SweepArgs {
    const uint64_t *kin;
    const uint32_t *vin;
    uint64_t *kout;
    uint32_t *vout;
    int64_t n;
    int shift;
    uint32_t mask;
    const uint32_t *ghist;       // this pass's 256 digit counts over all keys
    unsigned long long *status;  // ntiles * 256 words, none of them tagged with this epoch
    unsigned long long epoch;    // 1..255, shifted into place
};

__device__ inline uint32_t block_excl_sum256(uint32_t x, uint32_t *s_wsum, uint32_t &total) {
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), w = tid / WAVE;
    uint32_t inc = x;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t y = __shfl_up(inc, d, WAVE);
        if (lane >= d) inc += y;
    }
    if (lane == WAVE - 1) s_wsum[w] = inc;
    __syncthreads();
    uint32_t ex = inc - x;
    total = 0;
    for (int q = 0; q < RS_WAVES; ++q) {
        if (q < w) ex += s_wsum[q];
        total += s_wsum[q];
    }
    __syncthreads();
    return ex;
}

template <bool VALS> __device__ __forceinline__ void onesweep_tile(const SweepArgs &a) {
    __shared__ uint64_t s_k[RS_TILE];
    __shared__ uint32_t s_v[VALS ? RS_TILE : 1];
    // index digit*RS_WAVES + wave; counts and tile offsets stay below RS_TILE, so
    // 16 bits suffice (the smaller footprint lets three blocks share a CU)
    __shared__ uint16_t s_cnt[256 * RS_WAVES];
    __shared__ uint32_t s_gbase[256];          // output offset of the digit minus its tile-local start
    __shared__ uint32_t s_wsum[RS_WAVES];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), w = tid / WAVE;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (WAVE - lane));
    for (int i = tid; i < 256 * RS_WAVES; i += RS_BLOCK) s_cnt[i] = 0;
    // a pass in which every key has the same digit keeps the order: copy the tile
    const bool trivial = __syncthreads_or(uint64_t(a.ghist[tid]) == uint64_t(a.n));
    const uint32_t tile = blockIdx.x;
    const int64_t base = int64_t(tile) * RS_TILE;
    if (trivial) {
        for (int64_t i = base + tid; i < a.n && i < base + RS_TILE; i += RS_BLOCK) {
            a.kout[i] = a.kin[i];
            if (VALS) a.vout[i] = a.vin[i];
        }
        return;
    }
    uint64_t k[RS_IT];
    uint32_t v[VALS ? RS_IT : 1], rk[RS_IT];
#pragma unroll
    for (int it = 0; it < RS_IT; ++it) {
        const int64_t i = base + w * RS_PER_WAVE + it * WAVE + lane;
        k[it] = i < a.n ? a.kin[i] : 0ull;
        if (VALS) v[it] = i < a.n ? a.vin[i] : 0u;
    }
#pragma unroll
    for (int it = 0; it < RS_IT; ++it) {
        const int64_t i = base + w * RS_PER_WAVE + it * WAVE + lane;
        const bool valid = i < a.n;
        const uint32_t d = uint32_t(k[it] >> a.shift) & a.mask;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t r = __popcll(m & lt);
        uint32_t before = 0;
        if (valid) before = s_cnt[d * RS_WAVES + w];
        rk[it] = before + r;
        if (valid && r == 0) s_cnt[d * RS_WAVES + w] = uint16_t(before + uint32_t(__popcll(m)));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // thread d: digit d's counts per wave and in the tile
    uint32_t c[RS_WAVES], cnt = 0;
#pragma unroll
    for (int q = 0; q < RS_WAVES; ++q) {
        c[q] = s_cnt[tid * RS_WAVES + q];
        cnt += c[q];
    }
    unsigned long long *my = a.status + size_t(tile) * 256 + tid;
    __hip_atomic_store(my, a.epoch | (tile == 0 ? ST_PRE : ST_AGG) | cnt, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    uint32_t tot;
    const uint32_t ex = block_excl_sum256(cnt, s_wsum, tot);
    {
        uint32_t run = ex;
#pragma unroll
        for (int q = 0; q < RS_WAVES; ++q) {
            s_cnt[tid * RS_WAVES + q] = uint16_t(run);
            run += c[q];
        }
    }
    // global base of the digit: exclusive scan of the pass histogram
    const uint32_t gx = block_excl_sum256(a.ghist[tid], s_wsum, tot);
    // the tile in digit order in LDS (the places inside the tile do not
    // depend on earlier tiles, and the keys' registers are free during the
    // look-back)
#pragma unroll
    for (int it = 0; it < RS_IT; ++it) {
        const int64_t i = base + w * RS_PER_WAVE + it * WAVE + lane;
        if (i < a.n) {
            const uint32_t d = uint32_t(k[it] >> a.shift) & a.mask;
            const uint32_t pos = s_cnt[d * RS_WAVES + w] + rk[it];
            s_k[pos] = k[it];
            if (VALS) s_v[pos] = v[it];
        }
    }
    // look-back over earlier tiles
    unsigned long long prefix = 0;
    if (tile > 0) {
        const unsigned long long EM = 0xffull << ST_EPOCH_SHIFT;
        int64_t t = int64_t(tile) - 1; // next predecessor to account for
        int spins = 0;
        bool done = false;
        while (!done) {
            // RS_LOOK predecessors per round trip: t, t-1, ...
            unsigned long long word[RS_LOOK];
#pragma unroll
            for (int j = 0; j < RS_LOOK; ++j)
                word[j] = t - j >= 0 ? __hip_atomic_load(a.status + size_t(t - j) * 256 + tid, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT)
                                     : 0ull;
            int used = 0;
            bool stalled = false;
#pragma unroll
            for (int j = 0; j < RS_LOOK; ++j) {
                if (done || stalled || t - j < 0) continue;
                const unsigned long long w = word[j];
                if ((w & EM) != a.epoch) {
                    stalled = true; // not yet published: wait for it
                    continue;
                }
                prefix += w & ST_VAL;
                ++used;
                if (w & ST_PRE) done = true;
            }
            t -= used;
            if (done) break;
            // A tile 0 word always carries a prefix, so t < 0 is a malformed
            // chain; a predecessor that stays unpublished past the spin limit
            // is treated the same way.  Either way this thread counts its
            // digit in the earlier tiles' input keys itself: exact and slow,
            // so a pass never scatters with a partial prefix (the keys are
            // used as indices before the host could see an error).
            const bool give_up = t < 0 || (used == 0 && ++spins > RS_SPIN_LIMIT);
            if (give_up) {
                uint64_t own = 0;
                for (int64_t i = 0; i < base; ++i) own += (uint32_t(a.kin[i] >> a.shift) & a.mask) == uint32_t(tid);
                prefix = own;
                break;
            }
            if (used == 0) __builtin_amdgcn_s_sleep(1);
            else spins = 0;
        }
        __hip_atomic_store(my, a.epoch | ST_PRE | (prefix + cnt), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_gbase[tid] = gx + uint32_t(prefix) - ex;
    __syncthreads();
    const int m = int(a.n - base < RS_TILE ? a.n - base : RS_TILE);
    for (int i = tid; i < m; i += RS_BLOCK) {
        const uint64_t key = s_k[i];
        const uint32_t o = s_gbase[uint32_t(key >> a.shift) & a.mask] + uint32_t(i);
        a.kout[o] = key;
        if (VALS) a.vout[o] = s_v[i];
    }
}

// keys with values; keys only without the values' LDS and registers, which
// lets four blocks share a CU (three with values)
__global__ __launch_bounds__(RS_BLOCK) void k_onesweep(SweepArgs a) { onesweep_tile<true>(a); }
__global__ __launch_bounds__(RS_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_onesweep_keys(SweepArgs a) {
    onesweep_tile<false>(a);
}

// up to SCAN_SET arrays of at most PREF_TILES tiles each in two launches (one
// when every array fits one tile); longer arrays take scan_u32_op
template <class Op, bool INCL>
void scan_u32_set(int count, const uint32_t *const *in, uint32_t *const *out, const int64_t *n,
                  uint32_t *const *d_total, hipStream_t s) {
    SC_REQUIRE(count >= 1 && count <= SCAN_SET, SC_EINVAL, "scan set size");
    ScanSet S{};
    int m = 0;
    int64_t maxnb = 0, nsum = 0;
    for (int k = 0; k < count; ++k) {
        const int64_t nb = ceil_div(n[k], SCAN_TILE);
        if (n[k] <= 0 || nb > PREF_TILES) {
            scan_u32_op<Op, INCL>(in[k], out[k], n[k], d_total[k], s);
            continue;
        }
        S.it[m].in = in[k];
        S.it[m].out = out[k];
        S.it[m].n = n[k];
        S.it[m].d_total = d_total[k];
        S.it[m].sums = reinterpret_cast<uint32_t *>(uintptr_t(nsum)); // offset, placed below
        nsum += nb;
        maxnb = nb > maxnb ? nb : maxnb;
        ++m;
    }
    if (m == 0) return;
    if (maxnb > 1) {
        uint32_t *sums = scan_scratch<uint32_t>(s).sums.ensure(size_t(nsum));
        for (int k = 0; k < m; ++k) S.it[k].sums = sums + uintptr_t(S.it[k].sums);
        k_tile_reduce_set<Op><<<dim3(unsigned(maxnb), unsigned(m)), SCAN_BLOCK, 0, s>>>(S);
        SC_HIP(hipGetLastError());
    } else {
        for (int k = 0; k < m; ++k) S.it[k].sums = nullptr; // one tile each: no earlier tiles
    }
    k_tile_scan_set<Op, INCL><<<dim3(unsigned(maxnb), unsigned(m)), SCAN_BLOCK, 0, s>>>(S);
    SC_HIP(hipGetLastError());
}

} // namespace

void scan_exclusive_u32(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *d_total, hipStream_t s) {
    scan_u32_op<OpSum, false>(in, out, n, d_total, s);
}
void scan_exclusive_u64(const uint64_t *in, uint64_t *out, int64_t n, uint64_t *d_total, hipStream_t s) {
    scan_impl<uint64_t>(in, out, n, d_total, s, scan_scratch<uint64_t>(s));
}
void sum_u64(const uint64_t *in, int64_t n, uint64_t *d_total, hipStream_t s) {
    if (n <= 0) {
        SC_HIP(hipMemsetAsync(d_total, 0, sizeof(uint64_t), s));
        return;
    }
    const int64_t nb = ceil_div(n, SCAN_TILE);
    SC_REQUIRE(nb < (int64_t(1) << 31), SC_ERANGE, "sum too long");
    uint64_t *sums = scan_scratch<uint64_t>(s).sums.ensure(size_t(nb));
    k_tile_reduce<uint64_t><<<unsigned(nb), SCAN_BLOCK, 0, s>>>(in, n, sums);
    SC_HIP(hipGetLastError());
    k_scan_sums<uint64_t><<<1, SCAN_BLOCK, 0, s>>>(sums, nb, d_total); // the total; the scanned sums go unused
    SC_HIP(hipGetLastError());
}

namespace {
struct FillArgs {
    uint32_t *p[FILL_MAX];
    int64_t n[FILL_MAX];
    uint32_t v[FILL_MAX];
};
// blockIdx.y picks the span
__global__ void k_fill_spans(FillArgs a) {
    const int q = blockIdx.y;
    uint32_t *p = a.p[q];
    const int64_t n = a.n[q];
    const uint32_t v = a.v[q];
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}
} // namespace

void FillBatch::flush() {
    if (count == 0) return;
    FillArgs a;
    int64_t mx = 0;
    for (int q = 0; q < count; ++q) {
        a.p[q] = static_cast<uint32_t *>(p[q]);
        a.n[q] = n[q];
        a.v[q] = v[q];
        mx = n[q] > mx ? n[q] : mx;
    }
    int64_t gx = ceil_div(mx, int64_t(256));
    if (gx > 2048) gx = 2048;
    k_fill_spans<<<dim3(unsigned(gx), unsigned(count)), 256, 0, s>>>(a);
    SC_HIP(hipGetLastError());
    count = 0;
}

void scan_inclusive_max_u32(const uint32_t *in, uint32_t *out, int64_t n, hipStream_t s) {
    if (n > 0) scan_u32_op<OpMax, true>(in, out, n, nullptr, s);
}
void scan_exclusive_u32_2(const uint32_t *in0, uint32_t *out0, int64_t n0, uint32_t *t0, const uint32_t *in1,
                          uint32_t *out1, int64_t n1, uint32_t *t1, hipStream_t s) {
    const uint32_t *in[2] = {in0, in1};
    uint32_t *out[2] = {out0, out1}, *tot[2] = {t0, t1};
    const int64_t n[2] = {n0, n1};
    scan_u32_set<OpSum, false>(2, in, out, n, tot, s);
}
void scan_inclusive_max_u32_2(const uint32_t *in0, uint32_t *out0, int64_t n0, const uint32_t *in1,
                              uint32_t *out1, int64_t n1, hipStream_t s) {
    const uint32_t *in[2] = {in0, in1};
    uint32_t *out[2] = {out0, out1}, *tot[2] = {nullptr, nullptr};
    const int64_t n[2] = {n0, n1};
    if (n0 <= 0 && n1 <= 0) return;
    scan_u32_set<OpMax, true>(2, in, out, n, tot, s);
}

namespace {
// the passes of the radix sort; true when the sorted data ends in kalt/valt
bool radix_passes(uint64_t *keys, uint32_t *vals, uint64_t *kalt, uint32_t *valt, int64_t n, int bit_lo,
                  int bit_hi, RadixScratch &rs, hipStream_t s) {
    SC_REQUIRE(n < (int64_t(1) << 32), SC_ERANGE, "radix sort longer than 2^32 keys");
    SC_REQUIRE(bit_lo >= 0 && bit_hi <= 64, SC_EINVAL, "radix bit range");
    int npass = int(ceil_div(bit_hi - bit_lo, 8));
    SC_REQUIRE(npass <= RS_MAXPASS, SC_EINVAL, "radix passes");
    // per pass 256 digit counts
    const size_t GH = size_t(RS_MAXPASS) * 256;
    uint32_t *gh = rs.global_hist.ensure(GH);
    SC_HIP(hipMemsetAsync(gh, 0, sizeof(uint32_t) * GH, s));
    int ghb = int(ceil_div(n, RS_BLOCK * 16));
    if (ghb > 1024) ghb = 1024;
    k_global_hist<<<ghb, RS_BLOCK, 0, s>>>(keys, n, bit_lo, bit_hi, npass, gh);
    SC_HIP(hipGetLastError());

    const int64_t ntiles = ceil_div(n, RS_TILE);
    SC_REQUIRE(ntiles < (int64_t(1) << 31), SC_ERANGE, "radix sort tiles");
    const size_t nst = size_t(ntiles) * 256;
    unsigned long long *const old_st = rs.tile_status.p;
    unsigned long long *st = rs.tile_status.ensure(nst);
    if (st != old_st) rs.status_clean = 0; // new memory: contents unknown
    uint64_t *kc = keys, *ko = kalt;
    uint32_t *vc = vals, *vo = valt;
    for (int p = 0; p < npass; ++p) {
        // a fresh epoch tag for this pass; zero the words when the tags wrap
        // or when the words in use were never cleared
        if (rs.epoch >= 255 || rs.status_clean < nst) {
            SC_HIP(hipMemsetAsync(st, 0, sizeof(unsigned long long) * rs.tile_status.cap, s));
            rs.status_clean = rs.tile_status.cap;
            rs.epoch = 0;
        }
        ++rs.epoch;
        const int sh = bit_lo + 8 * p;
        const int bits = bit_hi - sh < 8 ? bit_hi - sh : 8;
        SweepArgs a;
        a.kin = kc;
        a.vin = vc;
        a.kout = ko;
        a.vout = vo;
        a.n = n;
        a.shift = sh;
        a.mask = (1u << bits) - 1u;
        a.ghist = gh + size_t(p) * 256;
        a.status = st;
        a.epoch = (unsigned long long)rs.epoch << ST_EPOCH_SHIFT;
        if (vc) k_onesweep<<<unsigned(ntiles), RS_BLOCK, 0, s>>>(a);
        else k_onesweep_keys<<<unsigned(ntiles), RS_BLOCK, 0, s>>>(a);
        SC_HIP(hipGetLastError());
        std::swap(kc, ko);
        std::swap(vc, vo);
    }
    return kc != keys;
}
} // namespace

size_t radix_fixed_words(int64_t n) {
    return size_t(RS_MAXPASS) * 256 + 2 * size_t(ceil_div(n > 0 ? n : 1, RS_TILE)) * 256;
}

bool radix_sort_u64_fixed(uint64_t *keys, uint32_t *vals, uint64_t *kalt, uint32_t *valt, int64_t n, int bit_lo,
                          int bit_hi, uint32_t *work, hipStream_t s) {
    if (n <= 1 || bit_hi <= bit_lo) return false;
    SC_REQUIRE(n < (int64_t(1) << 32), SC_ERANGE, "radix sort longer than 2^32 keys");
    SC_REQUIRE(bit_lo >= 0 && bit_hi <= 64, SC_EINVAL, "radix bit range");
    const int npass = int(ceil_div(bit_hi - bit_lo, 8));
    SC_REQUIRE(npass <= RS_MAXPASS, SC_EINVAL, "radix passes");
    const size_t GH = size_t(RS_MAXPASS) * 256;
    const int64_t ntiles = ceil_div(n, RS_TILE);
    uint32_t *gh = work;
    // 8-byte aligned: GH words is even
    unsigned long long *st = reinterpret_cast<unsigned long long *>(work + GH);
    {
        // histograms and every status word start at zero; cleared by a
        // kernel (no memset node when a caller captures the stream)
        FillBatch fb(s);
        fb.add(gh, int64_t(GH), 0u);
        fb.add(st, 2 * ntiles * 256, 0u);
        fb.flush();
    }
    int ghb = int(ceil_div(n, RS_BLOCK * 16));
    if (ghb > 1024) ghb = 1024;
    k_global_hist<<<ghb, RS_BLOCK, 0, s>>>(keys, n, bit_lo, bit_hi, npass, gh);
    SC_HIP(hipGetLastError());
    uint64_t *kc = keys, *ko = kalt;
    uint32_t *vc = vals, *vo = valt;
    for (int p = 0; p < npass; ++p) {
        // epochs 1..npass: the words were zeroed above, so none is stale
        const int sh = bit_lo + 8 * p;
        const int bits = bit_hi - sh < 8 ? bit_hi - sh : 8;
        SweepArgs a;
        a.kin = kc;
        a.vin = vc;
        a.kout = ko;
        a.vout = vo;
        a.n = n;
        a.shift = sh;
        a.mask = (1u << bits) - 1u;
        a.ghist = gh + size_t(p) * 256;
        a.status = st;
        a.epoch = (unsigned long long)(p + 1) << ST_EPOCH_SHIFT;
        if (vc) k_onesweep<<<unsigned(ntiles), RS_BLOCK, 0, s>>>(a);
        else k_onesweep_keys<<<unsigned(ntiles), RS_BLOCK, 0, s>>>(a);
        SC_HIP(hipGetLastError());
        std::swap(kc, ko);
        std::swap(vc, vo);
    }
    return kc != keys;
}

void radix_sort_u64(uint64_t *keys, uint32_t *vals, int64_t n, int bit_lo, int bit_hi, RadixScratch &rs,
                    hipStream_t s) {
    if (n <= 1 || bit_hi <= bit_lo) return;
    uint64_t *kalt = rs.keys_alt.ensure(size_t(n));
    uint32_t *valt = vals ? rs.vals_alt.ensure(size_t(n)) : nullptr;
    if (radix_passes(keys, vals, kalt, valt, n, bit_lo, bit_hi, rs, s)) {
        SC_HIP(hipMemcpyAsync(keys, kalt, sizeof(uint64_t) * size_t(n), hipMemcpyDeviceToDevice, s));
        if (vals) SC_HIP(hipMemcpyAsync(vals, valt, sizeof(uint32_t) * size_t(n), hipMemcpyDeviceToDevice, s));
    }
}

void radix_sort_u64_buf(DBuf<uint64_t> &keys, DBuf<uint64_t> &keys_alt, DBuf<uint32_t> *vals,
                        DBuf<uint32_t> *vals_alt, int64_t n, int bit_lo, int bit_hi, RadixScratch &rs, hipStream_t s) {
    if (n <= 1 || bit_hi <= bit_lo) return;
    SC_REQUIRE(keys.cap >= size_t(n) && (!vals || vals->cap >= size_t(n)) && (!vals == !vals_alt), SC_EINVAL,
               "radix sort buffers");
    uint64_t *kalt = keys_alt.ensure(size_t(n));
    uint32_t *valt = vals ? vals_alt->ensure(size_t(n)) : nullptr;
    if (radix_passes(keys.p, vals ? vals->p : nullptr, kalt, valt, n, bit_lo, bit_hi, rs, s)) {
        keys.swap(keys_alt);
        if (vals) vals->swap(*vals_alt);
    }
}

} // namespace sc
