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
constexpr int RCH = 256; // rows per work unit

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
    const double *x0, *y0, *x1, *y1;
    int nl, D, T;
    uint32_t nunits;
    int bits_b, bits_c;
    uint64_t *out;
    uint64_t cap;
    unsigned long long *counter;
};

__global__ __launch_bounds__(BLK) void k_stab(StabArgs a) {
    const int lane = threadIdx.x & (WAVE - 1);
    const uint32_t wg = blockIdx.x * (BLK / WAVE) + threadIdx.x / WAVE;
    if (wg >= a.nunits) return;
    // node owning unit wg: last v with uoff[v] <= wg
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
    bool valid = uint32_t(lane) < ncol;
    uint32_t ce = 0;
    if (valid) {
        uint32_t raw = colsrc[col_begin + lane];
        ce = raw & 0x7fffffffu;
        if (kind != 1 && (raw & 0x80000000u)) {
            // "after" copy of an edge whose "before" leaf is under v as well
            int bl = a.before[ce];
            int half = a.nl;
            if (bl >= half + s.lo && bl < half + s.hi) valid = false;
        }
    }
    double cx0 = 0, cy0 = 0, cx1 = 0, cy1 = 0;
    int ck = 0;
    if (valid) {
        cx0 = a.x0[ce];
        cy0 = a.y0[ce];
        cx1 = a.x1[ce];
        cy1 = a.y1[ce];
        if (kind != 3) ck = a.etype[ce];
    }
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (WAVE - lane));
    const int sb = a.bits_b + a.bits_c;
    for (uint32_t r = 0; r < nrows; ++r) {
        const uint32_t re = a.cov_e[rows_begin + r];
        const double rx0 = a.x0[re], ry0 = a.y0[re], rx1 = a.x1[re], ry1 = a.y1[re];
        bool hit = false;
        uint64_t key = 0;
        if (valid) {
            double bx0, by0, bx1, by1, qx0, qy0, qx1, qy1;
            uint32_t be, qe;
            int k;
            if (kind == 3) {
                bx0 = cx0; by0 = cy0; bx1 = cx1; by1 = cy1; be = ce;
                qx0 = rx0; qy0 = ry0; qx1 = rx1; qy1 = ry1; qe = re;
                k = a.etype[re];
            } else {
                bx0 = rx0; by0 = ry0; bx1 = rx1; by1 = ry1; be = re;
                qx0 = cx0; qy0 = cy0; qx1 = cx1; qy1 = cy1; qe = ce;
                k = ck;
            }
            if (lsmf(bx0, by0, bx1, by1, qx0, qy0, qx1, qy1)) {
                double al, be_;
                int t = classify(bx0, by0, bx1, by1, qx0, qy0, qx1, qy1, al, be_);
                if (t != 0) {
                    hit = true;
                    uint32_t local = qe - uint32_t(a.eoff[k - 1]);
                    key = (uint64_t(k) << sb) | (uint64_t(be) << a.bits_c) | uint64_t(local);
                }
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
                                 uint64_t *out, int sh, uint32_t *kcount) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // pos is the exclusive scan of the head flags: a head is where pos[i+1] > pos[i]
    if (pos[i + 1] != pos[i]) {
        out[pos[i]] = k[i];
        atomicAdd(&kcount[int(k[i] >> sh)], 1u);
    }
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
    uint32_t *kcount = c.d_pair_off.ensure(size_t(c.K) + 2);
    SC_HIP(hipMemsetAsync(kcount, 0, sizeof(uint32_t) * (size_t(c.K) + 2), s));
    if (n > 0) {
        k_compact_unique<<<grid_for(n), BLK, 0, s>>>(c.cand, flag, n, c.pairs, c.bits_b + c.bits_c, kcount);
        SC_HIP(hipGetLastError());
    }
    // kcount[k] = pairs of overlay k (k >= 1); offsets per overlay
    std::vector<uint32_t> h(size_t(c.K) + 2);
    SC_HIP(hipMemcpyAsync(h.data(), kcount, sizeof(uint32_t) * h.size(), hipMemcpyDeviceToHost, s));
    SC_HIP(hipStreamSynchronize(s));
    c.pair_off.assign(size_t(c.K) + 1, 0);
    for (int k = 1; k <= c.K; ++k) c.pair_off[k] = c.pair_off[k - 1] + h[k];
}

} // namespace sc
