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
// Work is cut into units of (node, rectangle, <=RCH rows, one or more
// 64-column tiles); a unit is found from a unit->node map.  One wave runs a
// unit with the column edges in lanes and the rows broadcast, so the base and
// overlay classes are never tested against themselves.  Hits are staged per
// wave in LDS and written with one atomic per flush.  All overlays are
// handled in the same pass: a pair key carries the overlay.
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

__host__ __device__ inline uint32_t cdiv32(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Work units of one rectangle (rows x cols): row blocks of RCH rows; a unit
// of a full row block covers g = ftpu 64-column tiles, a unit of the last
// (partial) row block covers enough tiles that it runs about ftpu*RCH row
// steps.  Several tiles per unit spread the unit's set-up (its record, its
// rows) over more tests; fewer give more, shorter waves.  ftpu is a tuning
// setting of the context (1..STAB_TILES_MAX); every value enumerates the same
// (row, column) rectangle, so the candidate set does not depend on it.
constexpr uint32_t STAB_TILES_MAX = 16;
__host__ __device__ inline uint32_t tiles_per_unit(uint32_t rows_last, uint32_t ftpu) {
    return ftpu * cdiv32(RCH, rows_last);
}
__host__ __device__ inline uint32_t rect_units(uint32_t rows, uint32_t cols, uint32_t ftpu) {
    if (!rows || !cols) return 0;
    const uint32_t rb = cdiv32(rows, RCH), ct = cdiv32(cols, WAVE);
    const uint32_t last = rows - (rb - 1) * RCH;
    return (rb - 1) * cdiv32(ct, ftpu) + cdiv32(ct, tiles_per_unit(last, ftpu));
}
// unit u of a rectangle -> first row, row count, tile range [t0,t1)
__host__ __device__ inline void rect_unit(uint32_t rows, uint32_t cols, uint32_t u, uint32_t ftpu, uint32_t &r0,
                                          uint32_t &nr, uint32_t &t0, uint32_t &t1) {
    const uint32_t rb = cdiv32(rows, RCH), ct = cdiv32(cols, WAVE);
    const uint32_t upb = cdiv32(ct, ftpu); // units per full row block
    const uint32_t full = (rb - 1) * upb;
    uint32_t g;
    if (u < full) {
        r0 = (u / upb) * RCH;
        nr = RCH;
        t0 = (u % upb) * ftpu;
        g = ftpu;
    } else {
        const uint32_t last = rows - (rb - 1) * RCH;
        g = tiles_per_unit(last, ftpu);
        r0 = (rb - 1) * RCH;
        nr = last;
        t0 = (u - full) * g;
    }
    t1 = t0 + g < ct ? t0 + g : ct;
}

__global__ void k_units(const uint32_t *__restrict__ cov_off, const uint32_t *__restrict__ cov_split,
                        const uint32_t *__restrict__ lend_off, int nl, int D, int T, uint32_t ftpu,
                        uint32_t *units, uint64_t *tests) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= T) return;
    uint32_t u = 0;
    uint64_t w = 0;
    if (v >= 1 && cov_off[v + 1] > cov_off[v]) {
        NodeShape s = node_shape(v, cov_off, cov_split, lend_off, nl, D);
        u = rect_units(s.nbase, s.nc, ftpu) + rect_units(s.nbase, s.ecn, ftpu) + rect_units(s.nc, s.ebn, ftpu);
        w = uint64_t(s.nbase) * s.nc + uint64_t(s.nbase) * s.ecn + uint64_t(s.nc) * s.ebn;
    }
    units[v] = u;
    tests[v] = w;
}

// unit -> node: each node with units marks its first unit; a max-scan spreads it
__global__ void k_unit_heads(int T, const uint32_t *__restrict__ uoff, uint32_t *head) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= T) return;
    if (uoff[v + 1] > uoff[v]) head[uoff[v]] = uint32_t(v);
}

// A work unit as the stab reads it: one uniform 32-byte record.
struct __align__(16) UnitDesc {
    uint32_t row0;   // first row entry (cover list position)
    uint32_t col0;   // first column entry of the rectangle
    uint32_t cols;   // columns of the rectangle
    uint32_t t0, t1; // 64-column tiles [t0, t1)
    uint32_t nk;     // rows | kind << 8
    uint32_t blo, bhi; // "before" leaves [blo, bhi) of the node's subtree (end-entry dedupe)
};

// descriptor of every unit from its node's shape (one thread per unit)
__global__ void k_unit_desc(uint32_t nunits, const uint32_t *__restrict__ unode, const uint32_t *__restrict__ uoff,
                            const uint32_t *__restrict__ cov_off, const uint32_t *__restrict__ cov_split,
                            const uint32_t *__restrict__ lend_off, int nl, int D, uint32_t ftpu,
                            UnitDesc *desc) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nunits) return;
    const int v = int(unode[w]);
    const NodeShape s = node_shape(v, cov_off, cov_split, lend_off, nl, D);
    uint32_t u = w - uoff[v];
    const uint32_t n1 = rect_units(s.nbase, s.nc, ftpu), n2 = rect_units(s.nbase, s.ecn, ftpu);
    uint32_t kind, rows, cols, row_begin, col_begin;
    if (u < n1) {
        kind = 1; // base cover x overlay cover
        rows = s.nbase;
        cols = s.nc;
        row_begin = s.off;
        col_begin = s.off + s.nbase;
    } else if (u < n1 + n2) {
        kind = 2; // base cover x overlay end entries
        u -= n1;
        rows = s.nbase;
        cols = s.ecn;
        row_begin = s.off;
        col_begin = s.ec0;
    } else {
        kind = 3; // overlay cover x base end entries
        u -= n1 + n2;
        rows = s.nc;
        cols = s.ebn;
        row_begin = s.off + s.nbase;
        col_begin = s.eb0;
    }
    uint32_t r0, nrows, t0, t1;
    rect_unit(rows, cols, u, ftpu, r0, nrows, t0, t1);
    UnitDesc d;
    d.row0 = row_begin + r0;
    d.col0 = col_begin;
    d.cols = cols;
    d.t0 = t0;
    d.t1 = t1;
    d.nk = nrows | (kind << 8);
    d.blo = uint32_t(nl + s.lo);
    d.bhi = uint32_t(nl + s.hi);
    desc[w] = d;
}

constexpr int STG = 256; // hits staged per wave in LDS before a flush
constexpr uint32_t XCHUNK = 16; // consecutive blocks per XCD turn

struct StabArgs {
    const UnitDesc *desc;
    const uint32_t *cov_e, *lend_e;
    const YRangeF *cov_yr, *lend_yr;          // single-precision y-extents beside the entries
    const int *lend_bl, *etype, *eoff;
    const double *x0, *y0, *x1, *y1;          // endpoints (input coordinates)
    int nl, D, T;
    uint32_t nunits;
    uint32_t upw;                             // consecutive units per wave
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
__device__ inline uint32_t uni(uint32_t v) { return uint32_t(__builtin_amdgcn_readfirstlane(int(v))); }

__device__ inline void flush(const StabArgs &a, const uint64_t *stage, uint32_t cnt, int lane) {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(a.counter, (unsigned long long)cnt);
    base = __shfl(base, 0, WAVE);
    for (uint32_t i = lane; i < cnt; i += WAVE)
        if (base + i < a.cap) a.out[base + i] = stage[i];
}

// One wave per unit (or per run of upw consecutive units, a tuning setting):
// up to 64 rows against one or more 64-column tiles (one
// column per lane).  The rows' single-precision y-extents sit in LDS and are
// read as broadcasts, four rows at a time; a row that meets no column costs
// two compares.  Pairs that pass are queued per wave; a full queue (or the
// end of a tile) runs the exact LSMF box test and lineSegmentIntersection2's
// classification (exact division-free prefilter first) on 64 pairs at once,
// with the double endpoints fetched into LDS once per unit (rows) and per
// tile (columns).  Hits are staged in LDS and written with one atomic per
// flush.
__global__ __launch_bounds__(BLK) void k_stab(StabArgs a) {
    __shared__ uint64_t stage_all[BLK / WAVE][STG];
    __shared__ YRangeF ry_all[BLK / WAVE][RCH];
    __shared__ uint32_t queue_all[BLK / WAVE][WAVE]; // near pairs: row << 6 | column lane
    __shared__ double4 rxy_all[BLK / WAVE][WAVE];    // double endpoints of the rows
    __shared__ double4 cxy_all[BLK / WAVE][WAVE];    // double endpoints of the tile's columns
    const int lane = threadIdx.x & (WAVE - 1);
    const uint32_t wv = uni(threadIdx.x / WAVE);
    // blocks are dealt round-robin to the 8 XCDs: hand each XCD chunks of
    // XCHUNK consecutive logical blocks, so that neighbouring units (mostly of
    // one node) share an L2 while the chunks stay interleaved for balance.
    // The grid is a multiple of 8*XCHUNK blocks: this is a permutation.
    const uint32_t nbk = gridDim.x;
    const uint32_t r8 = blockIdx.x / 8, xcd = blockIdx.x % 8;
    const uint32_t lb = (r8 / XCHUNK) * (8 * XCHUNK) + xcd * XCHUNK + r8 % XCHUNK; // logical block
    // this wave runs units [u_first, u_last): a.upw consecutive units, hits
    // staged across them (one flush per full stage or per wave)
    const uint64_t u_first = uint64_t(lb * (BLK / WAVE) + wv) * a.upw;
    if (lb >= nbk || u_first >= a.nunits) return; // whole wave; no block barrier below
    const uint32_t u_last = uint32_t(min(u_first + a.upw, uint64_t(a.nunits)));
    uint64_t *stage = stage_all[wv];
    YRangeF *ry = ry_all[wv];
    uint32_t *queue = queue_all[wv];
    double4 *rxy = rxy_all[wv], *cxy = cxy_all[wv];
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (WAVE - lane));
    const int sb = a.bits_b + a.bits_c;
    uint32_t cnt = 0;
    for (uint32_t wg = uint32_t(u_first); wg < u_last; ++wg) {
        // the unit's record: a uniform load
        const uint4 *dp = reinterpret_cast<const uint4 *>(a.desc + wg);
        const uint4 d0 = dp[0], d1 = dp[1];
        const uint32_t row0 = uni(d0.x), col_begin = uni(d0.y), cols = uni(d0.z), t0 = uni(d0.w);
        const uint32_t t1 = uni(d1.x), nk = uni(d1.y), blo = uni(d1.z), bhi = uni(d1.w);
        const uint32_t nrows = nk & 0xffu;
        const int kind = int(nk >> 8);
        const uint32_t *colsrc = kind == 1 ? a.cov_e : a.lend_e;
        const YRangeF *colyr = kind == 1 ? a.cov_yr : a.lend_yr;
        // a tile's columns: entry, box and (end entries) the edge's "before" leaf
        auto load_tile = [&](uint32_t t, uint32_t &raw, YRangeF &bx, uint32_t &bl) {
            const uint32_t cb = t * WAVE;
            raw = 0xffffffffu;
            bl = 0;
            if (uint32_t(lane) < min(uint32_t(WAVE), cols - cb)) {
                const uint32_t ci = col_begin + cb + lane;
                raw = colsrc[ci];
                bx = colyr[ci];
                if (kind != 1) bl = uint32_t(a.lend_bl[ci]);
            }
        };
        uint32_t n_raw, n_bl;
        YRangeF n_box;
        load_tile(t0, n_raw, n_box, n_bl); // issued before the rows are waited for
        // this lane's row edge and box (the empty box past the last row)
        uint32_t re_l = 0;
        YRangeF rb = yrange_empty();
        if (uint32_t(lane) < nrows) {
            re_l = a.cov_e[row0 + lane];
            rb = a.cov_yr[row0 + lane];
        }
        ry[lane] = rb;
        __builtin_amdgcn_wave_barrier();
        bool rows_fetched = false;
        int rk = 0;
        const uint32_t nr4 = (nrows + 3u) & ~3u; // rows past nrows hold the empty box
        for (uint32_t t = t0; t < t1; ++t) {
            if (t != t0) load_tile(t, n_raw, n_box, n_bl);
            const uint32_t raw = n_raw, bl = n_bl;
            YRangeF cbf = n_box;
            bool valid = raw != 0xffffffffu;
            const uint32_t ce = raw & 0x7fffffffu;
            // "after" copy of an edge whose "before" leaf is under v as well
            if (valid && kind != 1 && (raw & 0x80000000u) && bl >= blo && bl < bhi) valid = false;
            if (!valid) cbf = yrange_empty();
            bool cols_fetched = false;
            int ck = 0;
            uint32_t qn = 0; // queued near pairs of this tile
            // LSMF on the doubles and the classification of the queued pairs, one per lane
            auto run_queue = [&]() {
                __builtin_amdgcn_wave_barrier();
                const bool act = uint32_t(lane) < qn;
                const uint32_t e = act ? queue[lane] : 0u;
                const int rr = int(e >> 6), cc = int(e & 63u);
                const double4 P = rxy[rr], Q = cxy[cc];
                const double px0 = P.x, py0 = P.y, px1 = P.z, py1 = P.w;
                const double qx0 = Q.x, qy0 = Q.y, qx1 = Q.z, qy1 = Q.w;
                const uint32_t re = uint32_t(__shfl(int(re_l), rr, WAVE)), qce = uint32_t(__shfl(int(ce), cc, WAVE));
                const int rkk = __shfl(rk, rr, WAVE), ckk = __shfl(ck, cc, WAVE);
                bool hit = false;
                uint64_t key = 0;
                if (act && box_overlap(box_of(px0, py0, px1, py1), box_of(qx0, qy0, qx1, qy1))) {
                    int tp;
                    uint32_t be, qe;
                    int kk;
                    if (kind == 3) { // row = overlay edge, column = base edge
                        tp = classify_pre(qx0, qy0, qx1, qy1, px0, py0, px1, py1);
                        be = qce;
                        qe = re;
                        kk = rkk;
                    } else {
                        tp = classify_pre(px0, py0, px1, py1, qx0, qy0, qx1, qy1);
                        be = re;
                        qe = qce;
                        kk = ckk;
                    }
                    if (tp != 0) {
                        hit = true;
                        const uint32_t local = qe - uint32_t(a.eoff[kk - 1]);
                        key = (uint64_t(kk) << sb) | (uint64_t(be) << a.bits_c) | uint64_t(local);
                    }
                }
                const uint64_t hm = __ballot(hit);
                if (hm) {
                    const uint32_t nh = uint32_t(__popcll(hm));
                    if (cnt + nh > uint32_t(STG)) {
                        flush(a, stage, cnt, lane);
                        cnt = 0;
                    }
                    if (hit) stage[cnt + uint32_t(__popcll(hm & lt))] = key;
                    cnt += nh;
                }
                qn = 0;
                __builtin_amdgcn_wave_barrier();
            };
            for (uint32_t r = 0; r < nr4; r += 4) {
                // single-precision screen (exact rejection) of four rows
                uint64_t m[4];
    #pragma unroll
                for (int k = 0; k < 4; ++k) {
                    // LSMF's y test, one ballot per comparison (the masks are ANDed in scalar
                    // registers).  Every tested pair overlaps in x by construction (a cover
                    // entry spans the node, an end entry has an endpoint in it), and the
                    // screen only has to keep every pair LSMF keeps.
                    const YRangeF q = ry[r + k];
                    m[k] = __builtin_amdgcn_ballot_w64(!(q.lo > cbf.hi)) & __builtin_amdgcn_ballot_w64(!(q.hi < cbf.lo));
                }
                if ((m[0] | m[1] | m[2] | m[3]) == 0) continue;
                if (!cols_fetched) {
                    if (valid) {
                        cxy[lane] = make_double4(a.x0[ce], a.y0[ce], a.x1[ce], a.y1[ce]);
                        ck = a.etype[ce];
                    }
                    cols_fetched = true;
                }
                if (!rows_fetched) {
                    if (uint32_t(lane) < nrows) {
                        rxy[lane] = make_double4(a.x0[re_l], a.y0[re_l], a.x1[re_l], a.y1[re_l]);
                        rk = a.etype[re_l];
                    }
                    rows_fetched = true;
                }
    #pragma unroll 1
                for (int k = 0; k < 4; ++k) {
                    const uint64_t mk = m[k];
                    if (mk == 0) continue;
                    const uint32_t nh = uint32_t(__popcll(mk));
                    if (qn + nh > uint32_t(WAVE)) run_queue();
                    if ((mk >> lane) & 1ull) queue[qn + uint32_t(__popcll(mk & lt))] = ((r + uint32_t(k)) << 6) | uint32_t(lane);
                    qn += nh;
                }
            }
            if (qn) run_queue();
        }
        __builtin_amdgcn_wave_barrier(); // the unit's LDS reads before the next unit's writes
    } // units of this wave
    if (cnt) flush(a, stage, cnt, lane);
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
// koff[k] = first pair of overlay k; the pair count is read on the device
__global__ void k_overlay_offsets(const uint64_t *__restrict__ pairs, const uint32_t *__restrict__ npairs, int K,
                                  int sh, uint32_t *off) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > K + 1) return;
    int64_t lo = 0, hi = int64_t(*npairs);
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
    SC_REQUIRE(c.stab_tiles >= 1 && uint32_t(c.stab_tiles) <= STAB_TILES_MAX, SC_EINVAL,
               "stab_tiles out of range");
    SC_REQUIRE(c.stab_upw >= 1 && c.stab_upw <= 32, SC_EINVAL, "stab_units_per_wave out of range");
    const uint32_t ftpu = uint32_t(c.stab_tiles);
    k_units<<<grid_for(T), BLK, 0, s>>>(c.cov_off, c.cov_split, c.lend_off, c.nleaves, c.depth, T, ftpu, units,
                                        tests);
    SC_HIP(hipGetLastError());
    {
        FillBatch fb(s);
        fb.add(units + T, 1, 0u);
        fb.add(tests + T, 2, 0u);
        fb.flush();
    }
    uint32_t *tot = c.d_u32_total.ensure(1);
    uint64_t *tot64 = c.d_u64_total.ensure(1);
    scan_exclusive_u32(units, units, int64_t(T) + 1, tot, s);
    scan_exclusive_u64(tests, tests, int64_t(T) + 1, tot64, s);
    uint32_t nunits = 0;
    uint64_t ntests = 0;
    {
        ReadBack rb(c); // one wait for both totals
        rb.add(&nunits, tot, 1);
        rb.add(&ntests, tot64, 1);
        rb.run();
    }
    c.n_tests = int64_t(ntests);
    c.n_units = nunits;
    uint32_t *unode = c.u32c.ensure(size_t(nunits) + 1);
    if (nunits > 0) {
        SC_HIP(hipMemsetAsync(unode, 0, sizeof(uint32_t) * nunits, s));
        k_unit_heads<<<grid_for(T), BLK, 0, s>>>(T, units, unode);
        SC_HIP(hipGetLastError());
        scan_inclusive_max_u32(unode, unode, nunits, s);
    }
    UnitDesc *desc = reinterpret_cast<UnitDesc *>(c.unit_desc.ensure(size_t(nunits > 0 ? nunits : 1) * 4));
    if (nunits > 0) {
        k_unit_desc<<<grid_for(nunits), BLK, 0, s>>>(nunits, unode, units, c.cov_off, c.cov_split, c.lend_off,
                                                     c.nleaves, c.depth, ftpu, desc);
        SC_HIP(hipGetLastError());
    }

    if (c.cand_cap == 0) c.cand_cap = std::max<int64_t>(1 << 20, 4 * int64_t(c.ne));
    unsigned long long *counter = reinterpret_cast<unsigned long long *>(c.d_u64_total.p);
    for (int attempt = 0; attempt < 2; ++attempt) {
        c.cand.ensure(size_t(c.cand_cap));
        c.cand_alt.ensure(size_t(c.cand_cap));
        SC_HIP(hipMemsetAsync(counter, 0, sizeof(uint64_t), s));
        StabArgs a;
        a.desc = desc; a.cov_e = c.cov_e; a.lend_e = c.lend_e;
        a.lend_bl = c.lend_bl; a.etype = c.etype; a.eoff = c.d_eoff;
        a.cov_yr = c.cov_yr; a.lend_yr = c.lend_yr;
        a.x0 = c.ex0; a.y0 = c.ey0; a.x1 = c.ex1; a.y1 = c.ey1;
        a.nl = c.nleaves; a.D = c.depth; a.T = T; a.nunits = nunits;
        a.upw = uint32_t(c.stab_upw);
        a.bits_b = c.bits_b; a.bits_c = c.bits_c;
        a.out = c.cand; a.cap = uint64_t(c.cand_cap); a.counter = counter;
        c.mark(8);
        if (nunits > 0) {
            // a multiple of 8*XCHUNK blocks: the XCD-aware block order is then a bijection
            const unsigned q = 8u * XCHUNK;
            const unsigned nblk = (grid_for((nunits + a.upw - 1u) / a.upw, BLK / WAVE) + q - 1u) / q * q;
            k_stab<<<nblk, BLK, 0, s>>>(a);
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
    radix_sort_u64_buf(c.cand, c.cand_alt, nullptr, nullptr, n, 0, kb, c.rs, s);
    uint32_t *flag = c.u32b.ensure(size_t(n) + 1);
    SC_HIP(hipMemsetAsync(flag, 0, sizeof(uint32_t) * (size_t(n) + 1), s));
    if (n > 0) {
        k_diff_flags64<<<grid_for(n), BLK, 0, s>>>(c.cand, n, flag);
        SC_HIP(hipGetLastError());
    }
    scan_exclusive_u32(flag, flag, n + 1, tot, s);
    // unique pairs compacted into a buffer sized for all candidates; their
    // count and the per-overlay offsets come back in one read-back
    c.pairs.ensure(size_t(n > 0 ? n : 1));
    uint32_t *koff = c.d_pair_off.ensure(size_t(c.K) + 2);
    if (n > 0) {
        k_compact_unique<<<grid_for(n), BLK, 0, s>>>(c.cand, flag, n, c.pairs);
        SC_HIP(hipGetLastError());
    }
    k_overlay_offsets<<<grid_for(int64_t(c.K) + 2), BLK, 0, s>>>(c.pairs, tot, c.K, c.bits_b + c.bits_c, koff);
    SC_HIP(hipGetLastError());
    // koff[k] = first pair of overlay k (k = 1..K+1): pairs of overlay k are [koff[k], koff[k+1])
    std::vector<uint32_t> h(size_t(c.K) + 2);
    uint32_t np = 0;
    {
        ReadBack rb(c);
        rb.add(h.data(), static_cast<const uint32_t *>(koff), h.size());
        rb.add(&np, static_cast<const uint32_t *>(tot), 1);
        rb.run();
    }
    c.n_pairs = np;
    c.pair_off.assign(size_t(c.K) + 1, 0);
    for (int k = 1; k <= c.K; ++k) c.pair_off[k] = h[k + 1];
}

} // namespace sc
