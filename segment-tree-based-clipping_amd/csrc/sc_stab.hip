// sc_stab.hip -- stab queries and candidate pairs (rows A8-A10).
//
// Reference: saveIntersectionsIdsMulticore (oacc_segment_tree.cpp:6010-6244)
// visits every tree node v with a non-empty cover list and tests, with LSMF
// then lineSegmentIntersection2, (i) cover entries of opposite polygons and
// (ii) every cover entry against the node's end list; pairs with a non-zero
// type are emitted, duplicates removed after a stable sort by base id
// (clip:368-493).
//
// Here the end list of v is not materialised: it is the union of the leaf end
// entries under v, a contiguous range of the (class, leaf)-sorted leaf entry
// array, with the copy of an edge whose "before" leaf is also under v skipped.
// Work is cut into units of (node, rectangle, <=RCH rows, 64 columns); one
// wave runs a unit with the 64 column edges in lanes and the rows broadcast,
// so the base and overlay classes are never tested against themselves.
// Hits are compacted with one ballot and one atomic per wave and row.  All
// overlays are handled in the same pass: a pair key carries the overlay.
#include "sc_ctx.hpp"
#include "sc_fp64.hpp"

namespace sc {

namespace {

constexpr int BLK = 256;
constexpr int RCH = 64; // rows per work unit (one per lane)

inline unsigned grid_for(int64_t n, int per = BLK) {
    int64_t g = ceil_div(n, per);
    if (g < 1) g = 1;
    return unsigned(g);
}

struct NodeShape {
    uint32_t off, nbase, nc;
    uint32_t eb0, ebn, ec0, ecn;
    int lo, hi; // leaf range [lo,hi)
};

__device__ inline NodeShape node_shape(int v, const uint32_t *cov_off, const uint32_t *cov_split,
                                       const uint32_t *lend_off, int nl, int D) {
    NodeShape s;
    s.off = cov_off[v];
    uint32_t n = cov_off[v + 1] - s.off;
    s.nbase = cov_split[v];
    s.nc = n - s.nbase;
    int d = 31 - __clz(v);
    int sh = D - 1 - d;
    s.lo = (v - (1 << d)) << sh;
    s.hi = s.lo + (1 << sh);
    s.eb0 = lend_off[s.lo];
    s.ebn = lend_off[s.hi] - s.eb0;
    s.ec0 = lend_off[nl + s.lo];
    s.ecn = lend_off[nl + s.hi] - s.ec0;
    return s;
}

__device__ inline uint32_t cdiv32(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

__global__ void k_units(const uint32_t *__restrict__ cov_off, const uint32_t *__restrict__ cov_split,
                        const uint32_t *__restrict__ lend_off, int nl, int D, int T, uint32_t *units,
                        uint64_t *tests) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= T) return;
    uint32_t u = 0;
    uint64_t w = 0;
    if (v >= 1 && cov_off[v + 1] > cov_off[v]) {
        NodeShape s = node_shape(v, cov_off, cov_split, lend_off, nl, D);
        if (s.nbase) u += cdiv32(s.nbase, RCH) * (cdiv32(s.nc, WAVE) + cdiv32(s.ecn, WAVE));
        if (s.nc) u += cdiv32(s.nc, RCH) * cdiv32(s.ebn, WAVE);
        w = uint64_t(s.nbase) * s.nc + uint64_t(s.nbase) * s.ecn + uint64_t(s.nc) * s.ebn;
    }
    units[v] = u;
    tests[v] = w;
}

struct StabArgs {
    const uint32_t *cov_off, *cov_split, *cov_e, *lend_off, *lend_e, *uoff;
    const int *before, *etype, *eoff;
    const double *x0, *y0, *x1, *y1;          // endpoints (classification)
    const double *bx0, *bx1, *by0, *by1;      // bounding boxes (LSMF)
    int nl, D, T;
    uint32_t nunits;
    int bits_b, bits_c;
    uint64_t *out;
    uint64_t cap;
    unsigned long long *counter;
};

// value of lane 'src' broadcast to the wave (src is uniform)
__device__ inline double bcast(double v, int src) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane(int(b), src);
    int hi = __builtin_amdgcn_readlane(int(b >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// One wave per unit: up to 64 rows (one per lane, broadcast in turn) x 64
// columns (one per lane).  LSMF is the bounding-box overlap test on the
// precomputed boxes (min/max of the endpoints, exactly as LSMF forms them);
// pairs that pass are classified with lineSegmentIntersection2.
__global__ __launch_bounds__(BLK) void k_stab(StabArgs a) {
    const int lane = threadIdx.x & (WAVE - 1);
    const uint32_t wg = blockIdx.x * (BLK / WAVE) + threadIdx.x / WAVE;
    if (wg >= a.nunits) return;
    int lo = 0, hi = a.T - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (a.uoff[mid] <= wg) lo = mid;
        else hi = mid - 1;
    }
    const int v = lo;
    NodeShape s = node_shape(v, a.cov_off, a.cov_split, a.lend_off, a.nl, a.D);
    uint32_t u = wg - a.uoff[v];
    const uint32_t rb = s.nbase ? cdiv32(s.nbase, RCH) : 0u;
    const uint32_t r1 = s.nbase ? rb * cdiv32(s.nc, WAVE) : 0u;
    const uint32_t r2 = s.nbase ? rb * cdiv32(s.ecn, WAVE) : 0u;
    int kind;
    uint32_t ncol, rows_begin, nrows, col_begin;
    const uint32_t *colsrc;
    if (u < r1) {
        kind = 1;
        uint32_t cc = cdiv32(s.nc, WAVE);
        uint32_t rc = u / cc, ci = u % cc;
        rows_begin = s.off + rc * RCH;
        nrows = min(uint32_t(RCH), s.nbase - rc * RCH);
        col_begin = s.off + s.nbase + ci * WAVE;
        ncol = min(uint32_t(WAVE), s.nc - ci * WAVE);
        colsrc = a.cov_e;
    } else if (u < r1 + r2) {
        kind = 2;
        u -= r1;
        uint32_t cc = cdiv32(s.ecn, WAVE);
        uint32_t rc = u / cc, ci = u % cc;
        rows_begin = s.off + rc * RCH;
        nrows = min(uint32_t(RCH), s.nbase - rc * RCH);
        col_begin = s.ec0 + ci * WAVE;
        ncol = min(uint32_t(WAVE), s.ecn - ci * WAVE);
        colsrc = a.lend_e;
    } else {
        kind = 3;
        u -= r1 + r2;
        uint32_t cc = cdiv32(s.ebn, WAVE);
        uint32_t rc = u / cc, ci = u % cc;
        rows_begin = s.off + s.nbase + rc * RCH;
        nrows = min(uint32_t(RCH), s.nc - rc * RCH);
        col_begin = s.eb0 + ci * WAVE;
        ncol = min(uint32_t(WAVE), s.ebn - ci * WAVE);
        colsrc = a.lend_e;
    }
    // this lane's column
    bool valid = uint32_t(lane) < ncol;
    uint32_t ce = 0;
    if (valid) {
        uint32_t raw = colsrc[col_begin + lane];
        ce = raw & 0x7fffffffu;
        if (kind != 1 && (raw & 0x80000000u)) {
            // "after" copy of an edge whose "before" leaf is under v as well
            int bl = a.before[ce];
            if (bl >= a.nl + s.lo && bl < a.nl + s.hi) valid = false;
        }
    }
    double cminx = 0, cmaxx = -1, cminy = 0, cmaxy = -1;
    if (valid) {
        cminx = a.bx0[ce];
        cmaxx = a.bx1[ce];
        cminy = a.by0[ce];
        cmaxy = a.by1[ce];
    }
    // this lane's row (broadcast to all lanes in the loop below)
    uint32_t re_l = 0;
    double rminx = 0, rmaxx = 0, rminy = 0, rmaxy = 0;
    if (uint32_t(lane) < nrows) {
        re_l = a.cov_e[rows_begin + lane];
        rminx = a.bx0[re_l];
        rmaxx = a.bx1[re_l];
        rminy = a.by0[re_l];
        rmaxy = a.by1[re_l];
    }
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (WAVE - lane));
    const int sb = a.bits_b + a.bits_c;
    for (uint32_t r = 0; r < nrows; ++r) {
        const double qminx = bcast(rminx, int(r)), qmaxx = bcast(rmaxx, int(r));
        const double qminy = bcast(rminy, int(r)), qmaxy = bcast(rmaxy, int(r));
        bool box = valid && !(qminx > cmaxx || qmaxx < cminx) && !(qminy > cmaxy || qmaxy < cminy);
        bool hit = false;
        uint64_t key = 0;
        if (box) {
            const uint32_t re = uint32_t(__builtin_amdgcn_readlane(int(re_l), int(r)));
            uint32_t be = kind == 3 ? ce : re, qe = kind == 3 ? re : ce;
            int t = classify_pre(a.x0[be], a.y0[be], a.x1[be], a.y1[be], a.x0[qe], a.y0[qe], a.x1[qe], a.y1[qe]);
            if (t != 0) {
                hit = true;
                int k = a.etype[qe];
                uint32_t local = qe - uint32_t(a.eoff[k - 1]);
                key = (uint64_t(k) << sb) | (uint64_t(be) << a.bits_c) | uint64_t(local);
            }
        }
        uint64_t m = __ballot(hit);
        if (m) {
            unsigned long long base = 0;
            int leader = __ffsll((long long)m) - 1;
            if (lane == leader) base = atomicAdd(a.counter, (unsigned long long)__popcll(m));
            base = __shfl(base, leader, WAVE);
            if (hit) {
                uint64_t pos = base + __popcll(m & lt);
                if (pos < a.cap) a.out[pos] = key;
            }
        }
    }
}

__global__ void k_diff_flags64(const uint64_t *__restrict__ k, int64_t n, uint32_t *f) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f[i] = (i == 0 || k[i] != k[i - 1]) ? 1u : 0u;
}

__global__ void k_compact_unique(const uint64_t *__restrict__ k, const uint32_t *__restrict__ pos, int64_t n,
                                 uint64_t *out) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // pos is the exclusive scan of the head flags: a head is where pos[i+1] > pos[i]
    if (pos[i + 1] != pos[i]) out[pos[i]] = k[i];
}

// first unique pair of every overlay k (pairs are sorted by overlay first)
__global__ void k_overlay_offsets(const uint64_t *__restrict__ pairs, int64_t n, int K, int sh, uint32_t *off) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > K + 1) return;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (int64_t(pairs[mid] >> sh) < k) lo = mid + 1;
        else hi = mid;
    }
    off[k] = uint32_t(lo);
}

int bits_for(int64_t n) { // bits to hold values in [0, n)
    int b = 0;
    while ((int64_t(1) << b) < n) b++;
    return b < 1 ? 1 : b;
}

} // namespace

void find_candidates(Ctx &c) {
    SC_REQUIRE(c.n_cov >= 0 && c.T > 0 && c.n_lend >= 0, SC_ESTATE, "tree not built");
    hipStream_t s = c.stream;
    const int T = c.T;
    c.bits_b = bits_for(c.nb);
    c.bits_c = bits_for(c.maxcsz);
    c.bits_k = bits_for(int64_t(c.K) + 1);
    SC_REQUIRE(c.bits_b + c.bits_c + c.bits_k <= 64, SC_ERANGE, "pair key does not fit in 64 bits");
    uint32_t *units = c.u32a.ensure(size_t(T) + 1);
    uint64_t *tests = c.u64b.ensure(size_t(T) + 1);
    k_units<<<grid_for(T), BLK, 0, s>>>(c.cov_off, c.cov_split, c.lend_off, c.nleaves, c.depth, T, units, tests);
    SC_HIP(hipGetLastError());
    SC_HIP(hipMemsetAsync(units + T, 0, sizeof(uint32_t), s));
    SC_HIP(hipMemsetAsync(tests + T, 0, sizeof(uint64_t), s));
    uint32_t *tot = c.d_u32_total.ensure(1);
    uint64_t *tot64 = c.d_u64_total.ensure(1);
    scan_exclusive_u32(units, units, int64_t(T) + 1, tot, s);
    scan_exclusive_u64(tests, tests, int64_t(T) + 1, tot64, s);
    uint32_t nunits = read_u32(c, tot);
    c.n_tests = int64_t(read_u64(c, tot64));

    if (c.cand_cap == 0) c.cand_cap = std::max<int64_t>(1 << 20, 4 * int64_t(c.ne));
    unsigned long long *counter = reinterpret_cast<unsigned long long *>(c.d_u64_total.p);
    for (int attempt = 0; attempt < 2; ++attempt) {
        c.cand.ensure(size_t(c.cand_cap));
        SC_HIP(hipMemsetAsync(counter, 0, sizeof(uint64_t), s));
        StabArgs a;
        a.cov_off = c.cov_off; a.cov_split = c.cov_split; a.cov_e = c.cov_e;
        a.lend_off = c.lend_off; a.lend_e = c.lend_e; a.uoff = units;
        a.before = c.before_leaf; a.etype = c.etype; a.eoff = c.d_eoff;
        a.x0 = c.ex0; a.y0 = c.ey0; a.x1 = c.ex1; a.y1 = c.ey1;
        a.bx0 = c.paired ? c.uxs.p : c.estart.p; a.bx1 = c.paired ? c.uxe.p : c.eend.p; a.by0 = c.eymin; a.by1 = c.eymax;
        a.nl = c.nleaves; a.D = c.depth; a.T = T; a.nunits = nunits;
        a.bits_b = c.bits_b; a.bits_c = c.bits_c;
        a.out = c.cand; a.cap = uint64_t(c.cand_cap); a.counter = counter;
        c.mark(8);
        if (nunits > 0) {
            k_stab<<<grid_for(nunits, BLK / WAVE), BLK, 0, s>>>(a);
            SC_HIP(hipGetLastError());
        }
        c.mark(9);
        c.n_cand = int64_t(read_u64(c, c.d_u64_total.p));
        if (c.n_cand <= c.cand_cap) break;
        c.cand_cap = c.n_cand + c.n_cand / 4 + 1024;
    }
    SC_REQUIRE(c.n_cand <= c.cand_cap, SC_EHIP, "candidate buffer overflow after resize");

    // unique pairs sorted by (overlay, base edge, overlay-local edge)
    const int kb = c.bits_k + c.bits_b + c.bits_c;
    const int64_t n = c.n_cand;
    radix_sort_u64(c.cand, nullptr, n, 0, kb, c.rs, s);
    uint32_t *flag = c.u32b.ensure(size_t(n) + 1);
    SC_HIP(hipMemsetAsync(flag, 0, sizeof(uint32_t) * (size_t(n) + 1), s));
    if (n > 0) {
        k_diff_flags64<<<grid_for(n), BLK, 0, s>>>(c.cand, n, flag);
        SC_HIP(hipGetLastError());
    }
    scan_exclusive_u32(flag, flag, n + 1, tot, s);
    c.n_pairs = read_u32(c, tot);
    c.pairs.ensure(size_t(c.n_pairs));
    uint32_t *koff = c.d_pair_off.ensure(size_t(c.K) + 2);
    if (n > 0) {
        k_compact_unique<<<grid_for(n), BLK, 0, s>>>(c.cand, flag, n, c.pairs);
        SC_HIP(hipGetLastError());
    }
    k_overlay_offsets<<<grid_for(int64_t(c.K) + 2), BLK, 0, s>>>(c.pairs, c.n_pairs, c.K, c.bits_b + c.bits_c, koff);
    SC_HIP(hipGetLastError());
    // koff[k] = first pair of overlay k (k = 1..K+1): pairs of overlay k are [koff[k], koff[k+1])
    std::vector<uint32_t> h(size_t(c.K) + 2);
    SC_HIP(hipMemcpyAsync(h.data(), koff, sizeof(uint32_t) * h.size(), hipMemcpyDeviceToHost, s));
    SC_HIP(hipStreamSynchronize(s));
    c.pair_off.assign(size_t(c.K) + 1, 0);
    for (int k = 1; k <= c.K; ++k) c.pair_off[k] = h[k + 1];
}

} // namespace sc
