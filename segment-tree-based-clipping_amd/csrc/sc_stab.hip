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
// Each active node yields three rectangles (base cover x overlay cover, base
// cover x overlay end entries, overlay cover x base end entries), so the base
// and overlay classes are never tested against themselves.  Three kernels:
//   k_stab       rectangles cut into work units of (<=RCH rows, one or more
//                64-column tiles), one wave per run of units, columns in
//                lanes; per tile a 32-bin y-mask union skips the rows that
//                can meet no column, the rest get the exact single-precision
//                y-screen;
//   k_stab_rows  rectangles of at most SROWS rows and one tile, one lane per
//                row walking its columns;
//   k_near_classify  the pairs that pass the screen ("near" pairs), one per
//                lane: LSMF and lineSegmentIntersection2's classification on
//                the doubles, hits staged per block and appended.
// Near pairs are written to NEAR_SEGS segments, each with its own counter, so
// that the flush atomics do not meet on one address.  All overlays are handled
// in the same pass: a pair key carries the overlay.
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

// A rectangle of at most SROWS rows and one tile of columns is "small": its
// rows go to k_stab_rows, one lane per row, instead of work units, whose
// set-up would cost more than its few tests.
constexpr uint32_t SROWS = 8;
__host__ __device__ inline bool small_rect(uint32_t rows, uint32_t cols) {
    return rows > 0 && cols > 0 && rows <= SROWS && cols <= uint32_t(WAVE);
}
__host__ __device__ inline uint32_t tile_units(uint32_t rows, uint32_t cols, uint32_t ftpu) {
    return small_rect(rows, cols) ? 0u : rect_units(rows, cols, ftpu);
}

__global__ void k_units(const uint32_t *__restrict__ cov_off, const uint32_t *__restrict__ cov_split,
                        const uint32_t *__restrict__ lend_off, int nl, int D, int T, uint32_t ftpu,
                        uint32_t *units, uint32_t *srows, uint64_t *tests) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= T) return;
    uint32_t u = 0, sr = 0;
    uint64_t w = 0;
    if (v >= 1 && cov_off[v + 1] > cov_off[v]) {
        NodeShape s = node_shape(v, cov_off, cov_split, lend_off, nl, D);
        const uint32_t rr[3] = {s.nbase, s.nbase, s.nc}, cc[3] = {s.nc, s.ecn, s.ebn};
        for (int k = 0; k < 3; ++k) {
            if (small_rect(rr[k], cc[k])) sr += rr[k];
            else u += rect_units(rr[k], cc[k], ftpu);
        }
        w = uint64_t(s.nbase) * s.nc + uint64_t(s.nbase) * s.ecn + uint64_t(s.nc) * s.ebn;
    }
    units[v] = u;
    srows[v] = sr;
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
    const uint32_t n1 = tile_units(s.nbase, s.nc, ftpu), n2 = tile_units(s.nbase, s.ecn, ftpu);
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

// Near pairs go to NEAR_SEGS segments of the output, each with its own
// counter on its own 128-byte line: wave w appends to segment w % NEAR_SEGS,
// which spreads the flush atomics over many addresses.
constexpr int NEAR_SEGS = 512;
constexpr int NEAR_PAD = 16; // u64 per counter line

struct StabArgs {
    const UnitDesc *desc;
    const uint32_t *cov_e, *lend_e;
    const YRangeF *cov_yr, *lend_yr;          // single-precision y-extents beside the entries
    const int *lend_bl;
    uint32_t nunits;
    uint32_t upw;                             // consecutive units per wave
    uint2 *out;                               // near pairs: NEAR_SEGS segments of 'cap' entries
    uint64_t cap;
    unsigned long long *counter;              // NEAR_SEGS counters, NEAR_PAD apart
};

// value of lane 'src' broadcast to the wave (src is uniform)
__device__ inline double bcast(double v, int src) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane(int(b), src);
    int hi = __builtin_amdgcn_readlane(int(b >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ inline uint32_t uni(uint32_t v) { return uint32_t(__builtin_amdgcn_readfirstlane(int(v))); }

// Wave-wide reductions (every lane active): four DPP steps combine the 16
// lanes of each row (quad swaps, half-row and row mirrors), four reads of
// lanes 0/16/32/48 combine the rows.
template <class Op> __device__ inline uint32_t wave_reduce(uint32_t v, Op op) {
    v = op(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0xB1, 0xF, 0xF, false))); // quad_perm 1,0,3,2
    v = op(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x4E, 0xF, 0xF, false))); // quad_perm 2,3,0,1
    v = op(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x141, 0xF, 0xF, false))); // row_half_mirror
    v = op(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x140, 0xF, 0xF, false))); // row_mirror
    const uint32_t a = uint32_t(__builtin_amdgcn_readlane(int(v), 0)), b = uint32_t(__builtin_amdgcn_readlane(int(v), 16));
    const uint32_t c = uint32_t(__builtin_amdgcn_readlane(int(v), 32)), d = uint32_t(__builtin_amdgcn_readlane(int(v), 48));
    return op(op(a, b), op(c, d));
}
__device__ inline uint32_t wave_or32(uint32_t v) {
    return wave_reduce(v, [](uint32_t x, uint32_t y) { return x | y; });
}
// min / max of floats through an order-preserving integer key (NaN excluded:
// callers handle NaN extents before)
__device__ inline int fkey(float f) {
    const int b = __float_as_int(f);
    return b ^ ((b >> 31) & 0x7fffffff);
}
__device__ inline float funkey(int k) { return __int_as_float(k ^ ((k >> 31) & 0x7fffffff)); }
__device__ inline float wave_min(float v) {
    auto f = [](uint32_t x, uint32_t y) { return uint32_t(min(int(x), int(y))); };
    return funkey(int(wave_reduce(uint32_t(fkey(v)), f)));
}
__device__ inline float wave_max(float v) {
    auto f = [](uint32_t x, uint32_t y) { return uint32_t(max(int(x), int(y))); };
    return funkey(int(wave_reduce(uint32_t(fkey(v)), f)));
}

// y-bins of a unit: the span [ulo, uhi] of its rows' y-extents cut into 32
// bins.  bin_of is monotone in y (a subtraction and a multiplication by a
// non-negative constant, each correctly rounded, then a clamp), so two
// ranges that overlap have overlapping bin ranges: a row whose bins meet
// none of a tile's column bins meets no column of the tile, and skipping it
// drops no pair the exact screen would keep.
constexpr float NBINS = 32.0f;
__device__ inline uint32_t bin_of(float y, float ulo, float inv) {
    return uint32_t(__builtin_amdgcn_fmed3f((y - ulo) * inv, 0.0f, NBINS - 1.0f));
}
__device__ inline uint32_t bin_mask(float lo, float hi, float ulo, float inv) {
    const uint32_t b0 = bin_of(lo, ulo, inv), b1 = bin_of(hi, ulo, inv);
    return (0xffffffffu >> (31u - b1)) & (0xffffffffu << b0);
}

__device__ inline void flush_near(uint2 *out_all, uint64_t cap, unsigned long long *counter, uint32_t seg,
                                  const uint2 *stage, uint32_t cnt, int lane) {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(counter + size_t(seg) * NEAR_PAD, (unsigned long long)cnt);
    base = __shfl(base, 0, WAVE);
    uint2 *out = out_all + size_t(seg) * cap;
    for (uint32_t i = lane; i < cnt; i += WAVE)
        if (base + i < cap) out[base + i] = stage[i];
}
__device__ inline void flush(const StabArgs &a, uint32_t seg, const uint2 *stage, uint32_t cnt, int lane) {
    flush_near(a.out, a.cap, a.counter, seg, stage, cnt, lane);
}

// One wave per unit (or per run of upw consecutive units, a tuning setting):
// up to 64 rows against one or more 64-column tiles (one
// column per lane).  Each row and each column carries a 32-bit mask of the
// unit's y-bins it covers; per tile the columns' masks are ORed across the
// wave and only rows whose mask meets that union are screened, one at a
// time: the row's single-precision y-extent is read from its lane and
// compared with every column (two compares).  Pairs that pass ("near" pairs)
// are staged per wave in LDS as (row edge | kind-3 bit, column edge) and
// written with one atomic per flush; k_near_classify runs LSMF and the
// classification on them, so a unit with many near pairs does not hold its
// wave.
__global__ __launch_bounds__(BLK) void k_stab(StabArgs a) {
    __shared__ uint2 stage_all[BLK / WAVE][STG]; // near pairs: (row edge | kind 3 << 31, column edge)
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
    uint2 *stage = stage_all[wv];
    const uint32_t seg = uint32_t(lb * (BLK / WAVE) + wv) % uint32_t(NEAR_SEGS);
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (WAVE - lane));
    uint32_t cnt = 0; // near pairs staged
    // A unit's record (two 16-byte loads, made uniform only one unit later),
    // its rows, and tiles of its columns (entry, box and, for end entries,
    // the edge's "before" leaf).  At the start of unit u the rows and first
    // tile of unit u+1 and the record of unit u+2 are requested, and the next
    // tile of u before each tile is screened, so that load latency overlaps
    // the work of a whole unit.
    struct Unit {
        uint32_t row0, col_begin, cols, t0, t1, nrows, blo, bhi;
        int kind;
    };
    struct Tile {
        uint32_t raw, bl;
        YRangeF bx;
    };
    struct Rec {
        uint4 d0, d1;
    };
    auto load_rec = [&](uint32_t wg) {
        const uint4 *dp = reinterpret_cast<const uint4 *>(a.desc + wg);
        Rec r;
        r.d0 = dp[0];
        r.d1 = dp[1];
        return r;
    };
    auto decode = [&](const Rec &r) {
        Unit u;
        u.row0 = uni(r.d0.x);
        u.col_begin = uni(r.d0.y);
        u.cols = uni(r.d0.z);
        u.t0 = uni(r.d0.w);
        u.t1 = uni(r.d1.x);
        const uint32_t nk = uni(r.d1.y);
        u.blo = uni(r.d1.z);
        u.bhi = uni(r.d1.w);
        u.nrows = nk & 0xffu;
        u.kind = int(nk >> 8);
        return u;
    };
    auto load_tile = [&](const Unit &u, uint32_t t) {
        const uint32_t cb = t * WAVE;
        Tile c;
        c.raw = 0xffffffffu;
        c.bl = 0;
        c.bx = yrange_empty();
        if (uint32_t(lane) < min(uint32_t(WAVE), u.cols - cb)) {
            const uint32_t ci = u.col_begin + cb + lane;
            c.raw = (u.kind == 1 ? a.cov_e : a.lend_e)[ci];
            c.bx = (u.kind == 1 ? a.cov_yr : a.lend_yr)[ci];
            if (u.kind != 1) c.bl = uint32_t(a.lend_bl[ci]);
        }
        return c;
    };
    // this lane's row edge and box (the empty box past the last row)
    auto load_rows = [&](const Unit &u, uint32_t &re, YRangeF &yr) {
        re = 0;
        yr = yrange_empty();
        if (uint32_t(lane) < u.nrows) {
            re = a.cov_e[u.row0 + lane];
            yr = a.cov_yr[u.row0 + lane];
        }
    };
    Unit nu = decode(load_rec(uint32_t(u_first)));
    Tile nfirst = load_tile(nu, nu.t0);
    uint32_t n_re;
    YRangeF n_rb;
    load_rows(nu, n_re, n_rb);
    Rec nrec;
    if (u_first + 1 < u_last) nrec = load_rec(uint32_t(u_first) + 1);
    Tile ntile;
    for (uint32_t wg = uint32_t(u_first); wg < u_last; ++wg) {
        const Unit cu = nu;
        const uint32_t re_l = n_re;
        const YRangeF rb = n_rb;
        const Tile first = nfirst;
        if (wg + 1 < u_last) {
            nu = decode(nrec);
            nfirst = load_tile(nu, nu.t0);
            load_rows(nu, n_re, n_rb);
            if (wg + 2 < u_last) nrec = load_rec(wg + 2);
        }
        const uint32_t t0 = cu.t0, t1 = cu.t1, nrows = cu.nrows, blo = cu.blo, bhi = cu.bhi;
        const int kind = cu.kind;
        // the unit's y-bins: the span of its rows; a NaN row turns the bins off
        const bool row_nan = uint32_t(lane) < nrows && !(rb.lo <= rb.hi);
        const bool nofilter = __ballot(row_nan) != 0;
        // (rows past nrows hold the empty box, +inf / -inf, neutral here)
        const float ulo = wave_min(nofilter ? 0.0f : rb.lo), uhi = wave_max(nofilter ? 0.0f : rb.hi);
        // any non-negative finite scale keeps the bins monotone; a span that is
        // empty, infinite or too small for the reciprocal gives one bin
        float inv = uhi > ulo ? NBINS * __builtin_amdgcn_rcpf(uhi - ulo) : 0.0f;
        if (!(inv < INFINITY)) inv = 0.0f;
        const uint32_t rm = uint32_t(lane) < nrows ? bin_mask(rb.lo, rb.hi, ulo, inv) : 0u;
        const uint64_t all_rows = nrows >= 64u ? ~0ull : ((1ull << nrows) - 1ull);
        const uint32_t rtag = kind == 3 ? 0x80000000u : 0u;
        for (uint32_t t = t0; t < t1; ++t) {
            const Tile ct = t == t0 ? first : ntile;
            if (t + 1 < t1) ntile = load_tile(cu, t + 1);
            const uint32_t raw = ct.raw, bl = ct.bl;
            YRangeF cbf = ct.bx;
            bool valid = raw != 0xffffffffu;
            const uint32_t ce = raw & 0x7fffffffu;
            // "after" copy of an edge whose "before" leaf is under v as well
            if (valid && kind != 1 && (raw & 0x80000000u) && bl >= blo && bl < bhi) valid = false;
            if (!valid) cbf = yrange_empty();
            // the column's bins: none outside the rows' span, all for a NaN extent
            uint32_t cm = 0;
            if (valid && !(cbf.hi < ulo) && !(cbf.lo > uhi))
                cm = cbf.lo <= cbf.hi ? bin_mask(cbf.lo, cbf.hi, ulo, inv) : ~0u;
            const uint32_t tm = wave_or32(cm);
            uint64_t rset = nofilter ? all_rows : __ballot((rm & tm) != 0u);
            while (rset) {
                const int r = __builtin_ctzll(rset);
                rset &= rset - 1ull;
                // single-precision screen (exact rejection) of row r: LSMF's y test, one
                // ballot per comparison.  Every tested pair overlaps in x by construction
                // (a cover entry spans the node, an end entry has an endpoint in it), and
                // the screen only has to keep every pair LSMF keeps.
                const float qlo = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rb.lo), r));
                const float qhi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rb.hi), r));
                const uint64_t mk = __builtin_amdgcn_ballot_w64(!(qlo > cbf.hi)) & __builtin_amdgcn_ballot_w64(!(qhi < cbf.lo));
                if (mk == 0) continue;
                const uint32_t nh = uint32_t(__popcll(mk));
                if (cnt + nh > uint32_t(STG)) {
                    flush(a, seg, stage, cnt, lane);
                    cnt = 0;
                }
                const uint32_t rowe = uint32_t(__builtin_amdgcn_readlane(int(re_l), r)) | rtag;
                if ((mk >> lane) & 1ull) stage[cnt + uint32_t(__popcll(mk & lt))] = make_uint2(rowe, ce);
                cnt += nh;
            }
        }
    } // units of this wave
    if (cnt) flush(a, seg, stage, cnt, lane);
}

// The rows of small rectangles, one per lane: a lane walks its rectangle's
// columns (at most one tile) and applies the same validity rule and
// single-precision y-screen as k_stab, so the near pairs are the ones k_stab
// would have found.  A row is found from the row -> node map; within its
// node, small rectangles are taken in kind order.
struct RowArgs {
    const uint32_t *ioff, *inode;             // per-node row offsets; row -> node
    const uint32_t *cov_off, *cov_split, *lend_off;
    int nl, D;
    const uint32_t *cov_e, *lend_e;
    const YRangeF *cov_yr, *lend_yr;
    const int *lend_bl;
    uint32_t nitems;
    uint2 *out;
    uint64_t cap;
    unsigned long long *counter;
};
__global__ __launch_bounds__(BLK) void k_stab_rows(RowArgs a) {
    __shared__ uint2 stage_all[BLK / WAVE][STG];
    const int lane = threadIdx.x & (WAVE - 1);
    const uint32_t wv = uni(threadIdx.x / WAVE);
    uint2 *stage = stage_all[wv];
    const uint32_t seg = (blockIdx.x * (BLK / WAVE) + wv + uint32_t(NEAR_SEGS) / 2) % uint32_t(NEAR_SEGS);
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (WAVE - lane));
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    const bool act = i < a.nitems;
    uint32_t rowe = 0, col_begin = 0, cols = 0, blo = 0, bhi = 0;
    int kind = 1;
    YRangeF rb = yrange_empty();
    if (act) {
        const int v = int(a.inode[i]);
        uint32_t local = i - a.ioff[v];
        const NodeShape s = node_shape(v, a.cov_off, a.cov_split, a.lend_off, a.nl, a.D);
        const uint32_t rr[3] = {s.nbase, s.nbase, s.nc}, cc[3] = {s.nc, s.ecn, s.ebn};
        const uint32_t r0[3] = {s.off, s.off, s.off + s.nbase}, c0[3] = {s.off + s.nbase, s.ec0, s.eb0};
        uint32_t row = 0;
        for (int k = 0; k < 3; ++k) {
            if (!small_rect(rr[k], cc[k])) continue;
            if (local < rr[k]) {
                kind = k + 1;
                row = r0[k] + local;
                col_begin = c0[k];
                cols = cc[k];
                break;
            }
            local -= rr[k];
        }
        rowe = a.cov_e[row] | (kind == 3 ? 0x80000000u : 0u);
        rb = a.cov_yr[row];
        blo = uint32_t(a.nl + s.lo);
        bhi = uint32_t(a.nl + s.hi);
    }
    uint32_t cnt = 0;
    // columns four at a time: the loads of a step are issued together
    constexpr int CU = 4;
    for (uint32_t j = 0;; j += CU) {
        if (__ballot(act && j < cols) == 0) break;
        uint32_t raw[CU], bl[CU];
        YRangeF bx[CU];
#pragma unroll
        for (int q = 0; q < CU; ++q) {
            raw[q] = 0xffffffffu;
            bl[q] = 0;
            bx[q] = yrange_empty();
            if (act && j + q < cols) {
                const uint32_t ci = col_begin + j + q;
                raw[q] = (kind == 1 ? a.cov_e : a.lend_e)[ci];
                bx[q] = (kind == 1 ? a.cov_yr : a.lend_yr)[ci];
                if (kind != 1) bl[q] = uint32_t(a.lend_bl[ci]);
            }
        }
#pragma unroll
        for (int q = 0; q < CU; ++q) {
            // "after" copy of an edge whose "before" leaf is under the node as well
            bool valid = raw[q] != 0xffffffffu;
            if (valid && kind != 1 && (raw[q] & 0x80000000u) && bl[q] >= blo && bl[q] < bhi) valid = false;
            const bool hit = valid && !(rb.lo > bx[q].hi) && !(rb.hi < bx[q].lo);
            const uint64_t hm = __ballot(hit);
            if (hm) {
                const uint32_t nh = uint32_t(__popcll(hm));
                if (cnt + nh > uint32_t(STG)) {
                    flush_near(a.out, a.cap, a.counter, seg, stage, cnt, lane);
                    cnt = 0;
                }
                if (hit) stage[cnt + uint32_t(__popcll(hm & lt))] = make_uint2(rowe, raw[q] & 0x7fffffffu);
                cnt += nh;
            }
        }
    }
    if (cnt) flush_near(a.out, a.cap, a.counter, seg, stage, cnt, lane);
}

// LSMF on the doubles and lineSegmentIntersection2's classification (exact
// division-free prefilter first) of every near pair, one per lane; pairs with
// a non-zero type become candidate keys (overlay, base edge, overlay-local
// edge).  Every block scans the segment counts into LDS (block 0 also stores
// the stored total and the largest count for the host); a block then takes
// chunks of BLK*CL_PER pairs of the concatenated segments, stages its hits in
// LDS and writes them with one atomic per chunk.
constexpr int CL_PER = 4;
struct ClassArgs {
    const uint2 *near;
    const unsigned long long *seg_cnt;        // NEAR_SEGS counters, NEAR_PAD apart
    uint64_t seg_cap;
    unsigned long long *summary;              // [0] pairs stored, [1] largest segment count
    const double4 *exy;                       // (x0, y0, x1, y1) per edge
    const int *etype, *eoff;
    int bits_b, bits_c;
    uint64_t *out;
    uint64_t cap;
    unsigned long long *counter;
};
__global__ __launch_bounds__(BLK) void k_near_classify(ClassArgs a) {
    constexpr uint32_t HITS = 4 * BLK * CL_PER; // hits of several chunks, written when the next might not fit
    __shared__ uint64_t hits[HITS];
    __shared__ uint64_t soff[NEAR_SEGS + 1];
    __shared__ uint64_t wsum[BLK / WAVE], wmax[BLK / WAVE];
    __shared__ uint32_t nh;
    __shared__ unsigned long long base;
    static_assert(NEAR_SEGS == 2 * BLK, "two segments per thread");
    {
        const int t = threadIdx.x, lane = t & (WAVE - 1), w = t / WAVE;
        const uint64_t r0 = a.seg_cnt[size_t(2 * t) * NEAR_PAD], r1 = a.seg_cnt[size_t(2 * t + 1) * NEAR_PAD];
        const uint64_t c0 = min(r0, a.seg_cap), c1 = min(r1, a.seg_cap);
        uint64_t incl = c0 + c1, mx = max(r0, r1);
        for (int d = 1; d < WAVE; d <<= 1) {
            const uint64_t o = __shfl_up(incl, d, WAVE);
            if (lane >= d) incl += o;
            mx = max(mx, uint64_t(__shfl_xor(mx, d, WAVE)));
        }
        if (lane == WAVE - 1) wsum[w] = incl;
        if (lane == 0) wmax[w] = mx;
        __syncthreads();
        uint64_t pre = 0, tot = 0, gmx = 0;
        for (int i = 0; i < BLK / WAVE; ++i) {
            if (i < w) pre += wsum[i];
            tot += wsum[i];
            gmx = max(gmx, wmax[i]);
        }
        const uint64_t excl = pre + incl - c0 - c1;
        soff[2 * t] = excl;
        soff[2 * t + 1] = excl + c0;
        if (t == 0) soff[NEAR_SEGS] = tot;
        if (blockIdx.x == 0 && t == 0) {
            a.summary[0] = tot;
            a.summary[1] = gmx;
        }
        __syncthreads();
    }
    // chunks never straddle segments: segment s holds chunks [coff[s], coff[s+1])
    __shared__ uint32_t coff[NEAR_SEGS + 1];
    const uint64_t chunk = uint64_t(BLK) * CL_PER;
    {
        const int t = threadIdx.x, lane = t & (WAVE - 1), w = t / WAVE;
        const uint32_t k0 = uint32_t((soff[2 * t + 1] - soff[2 * t] + chunk - 1) / chunk);
        const uint32_t k1 = uint32_t((soff[2 * t + 2] - soff[2 * t + 1] + chunk - 1) / chunk);
        uint32_t incl = k0 + k1;
        for (int d = 1; d < WAVE; d <<= 1) {
            const uint32_t o = __shfl_up(incl, d, WAVE);
            if (lane >= d) incl += o;
        }
        __syncthreads(); // soff reads above before wsum is reused
        if (lane == WAVE - 1) wsum[w] = incl;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
        for (int i = 0; i < BLK / WAVE; ++i) {
            if (i < w) pre += uint32_t(wsum[i]);
            tot += uint32_t(wsum[i]);
        }
        const uint32_t excl = pre + incl - k0 - k1;
        coff[2 * t] = excl;
        coff[2 * t + 1] = excl + k0;
        if (t == 0) coff[NEAR_SEGS] = tot;
        __syncthreads();
    }
    const uint32_t nchunks = coff[NEAR_SEGS];
    if (threadIdx.x == 0) nh = 0;
    __syncthreads();
    // the staged hits to the candidate list, one atomic per call (block-uniform)
    auto flush_hits = [&]() {
        if (threadIdx.x == 0) base = nh ? atomicAdd(a.counter, (unsigned long long)nh) : 0ull;
        __syncthreads();
        const uint32_t n_h = nh;
        for (uint32_t i = threadIdx.x; i < n_h; i += BLK)
            if (base + i < a.cap) a.out[base + i] = hits[i];
        __syncthreads();
        if (threadIdx.x == 0) nh = 0;
        __syncthreads();
    };
    const int sb = a.bits_b + a.bits_c;
    for (uint32_t k = blockIdx.x; k < nchunks; k += gridDim.x) {
        int sg = 0, hi = NEAR_SEGS; // coff[sg] <= k < coff[sg + 1]
        while (hi - sg > 1) {
            const int mid = (sg + hi) >> 1;
            if (coff[mid] <= k) sg = mid;
            else hi = mid;
        }
        const uint64_t first = uint64_t(k - coff[sg]) * chunk;
        const uint64_t cnt = min(chunk, soff[sg + 1] - soff[sg] - first);
        const uint2 *src = a.near + size_t(sg) * a.seg_cap + first;
        // every load of the chunk's pairs is issued before the first test
        uint2 e[CL_PER];
#pragma unroll
        for (int j = 0; j < CL_PER; ++j) {
            const uint64_t i = uint64_t(j) * BLK + threadIdx.x;
            e[j] = i < cnt ? src[i] : make_uint2(0xffffffffu, 0u);
        }
        double4 B[CL_PER], C[CL_PER];
#pragma unroll
        for (int j = 0; j < CL_PER; ++j) {
            if (e[j].x != 0xffffffffu) {
                const bool swap = (e[j].x >> 31) != 0u; // row = overlay edge, column = base edge
                const uint32_t re = e[j].x & 0x7fffffffu;
                B[j] = a.exy[swap ? e[j].y : re];
                C[j] = a.exy[swap ? re : e[j].y];
            }
        }
#pragma unroll
        for (int j = 0; j < CL_PER; ++j) {
            if (e[j].x == 0xffffffffu) continue;
            const bool swap = (e[j].x >> 31) != 0u;
            const uint32_t re = e[j].x & 0x7fffffffu;
            const uint32_t be = swap ? e[j].y : re, qe = swap ? re : e[j].y;
            if (box_overlap(box_of(B[j].x, B[j].y, B[j].z, B[j].w), box_of(C[j].x, C[j].y, C[j].z, C[j].w)) &&
                classify_pre(B[j].x, B[j].y, B[j].z, B[j].w, C[j].x, C[j].y, C[j].z, C[j].w) != 0) {
                const int kk = a.etype[qe];
                const uint32_t local = qe - uint32_t(a.eoff[kk - 1]);
                hits[atomicAdd(&nh, 1u)] = (uint64_t(kk) << sb) | (uint64_t(be) << a.bits_c) | uint64_t(local);
            }
        }
        __syncthreads();
        if (nh + uint32_t(BLK * CL_PER) > HITS) flush_hits();
    }
    flush_hits();
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
    uint32_t *srows = c.srow_off.ensure(size_t(T) + 1);
    uint64_t *tests = c.u64b.ensure(size_t(T) + 1);
    SC_REQUIRE(c.stab_tiles >= 1 && uint32_t(c.stab_tiles) <= STAB_TILES_MAX, SC_EINVAL,
               "stab_tiles out of range");
    SC_REQUIRE(c.stab_upw >= 1 && c.stab_upw <= 32, SC_EINVAL, "stab_units_per_wave out of range");
    const uint32_t ftpu = uint32_t(c.stab_tiles);
    k_units<<<grid_for(T), BLK, 0, s>>>(c.cov_off, c.cov_split, c.lend_off, c.nleaves, c.depth, T, ftpu, units,
                                        srows, tests);
    SC_HIP(hipGetLastError());
    {
        FillBatch fb(s);
        fb.add(units + T, 1, 0u);
        fb.add(srows + T, 1, 0u);
        fb.add(tests + T, 2, 0u);
        fb.flush();
    }
    uint32_t *tot = c.d_u32_total.ensure(2);
    uint64_t *tot64 = c.d_u64_total.ensure(2); // later: the candidate counter
    scan_exclusive_u32_2(units, units, int64_t(T) + 1, tot, srows, srows, int64_t(T) + 1, tot + 1, s);
    sum_u64(tests, int64_t(T) + 1, tot64, s); // only the total (pair tests W) is used
    uint32_t tot2[2] = {0, 0};
    uint64_t ntests = 0;
    {
        ReadBack rb(c); // one wait for the three totals
        rb.add(tot2, tot, 2);
        rb.add(&ntests, tot64, 1);
        rb.run();
    }
    const uint32_t nunits = tot2[0], nitems = tot2[1];
    c.n_tests = int64_t(ntests);
    c.n_units = nunits;
    uint32_t *inode = c.srow_node.ensure(size_t(nitems) + 1);
    uint32_t *unode = c.u32c.ensure(size_t(nunits) + 1);
    unsigned long long *counter = reinterpret_cast<unsigned long long *>(c.d_u64_total.p);
    unsigned long long *ncnt =
        reinterpret_cast<unsigned long long *>(c.near_cnt.ensure(size_t(NEAR_SEGS) * NEAR_PAD + 2));
    unsigned long long *summary = ncnt + size_t(NEAR_SEGS) * NEAR_PAD;
    // the node maps and the stab's counters start at zero (one fill launch;
    // the totals above are already read back)
    {
        FillBatch fb(s);
        fb.add(inode, nitems, 0u);
        fb.add(unode, nunits, 0u);
        fb.add(counter, 2, 0u);
        fb.add(ncnt, 2 * (int64_t(NEAR_SEGS) * NEAR_PAD + 2), 0u);
        fb.flush();
    }
    if (nitems > 0) k_unit_heads<<<grid_for(T), BLK, 0, s>>>(T, srows, inode);
    if (nunits > 0) k_unit_heads<<<grid_for(T), BLK, 0, s>>>(T, units, unode);
    SC_HIP(hipGetLastError());
    scan_inclusive_max_u32_2(inode, inode, nitems, unode, unode, nunits, s);
    UnitDesc *desc = reinterpret_cast<UnitDesc *>(c.unit_desc.ensure(size_t(nunits > 0 ? nunits : 1) * 4));
    if (nunits > 0) {
        k_unit_desc<<<grid_for(nunits), BLK, 0, s>>>(nunits, unode, units, c.cov_off, c.cov_split, c.lend_off,
                                                     c.nleaves, c.depth, ftpu, desc);
        SC_HIP(hipGetLastError());
    }

    if (c.cand_cap == 0) c.cand_cap = std::max<int64_t>(1 << 20, 4 * int64_t(c.ne));
    if (c.near_seg_cap == 0) c.near_seg_cap = std::max<int64_t>(1 << 13, 16 * int64_t(c.ne) / NEAR_SEGS);
    int cus = 256, v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, c.device) == hipSuccess && v > 0) cus = v;
    // a buffer that overflows is counted in full: it is resized to the count
    // and the stab runs again (at most once per buffer)
    int64_t near_max = 0;
    for (int attempt = 0; attempt < 3; ++attempt) {
        c.near.ensure(size_t(c.near_seg_cap) * NEAR_SEGS);
        c.cand.ensure(size_t(c.cand_cap));
        c.cand_alt.ensure(size_t(c.cand_cap));
        if (attempt > 0) { // the counters of the run that overflowed
            FillBatch fb(s);
            fb.add(counter, 2, 0u);
            fb.add(ncnt, 2 * (int64_t(NEAR_SEGS) * NEAR_PAD + 2), 0u);
            fb.flush();
        }
        StabArgs a;
        a.desc = desc; a.cov_e = c.cov_e; a.lend_e = c.lend_e;
        a.lend_bl = c.lend_bl;
        a.cov_yr = c.cov_yr; a.lend_yr = c.lend_yr;
        a.nunits = nunits;
        a.upw = uint32_t(c.stab_upw);
        a.out = c.near; a.cap = uint64_t(c.near_seg_cap); a.counter = ncnt;
        c.mark(8);
        if (nunits > 0) {
            // a multiple of 8*XCHUNK blocks: the XCD-aware block order is then a bijection
            const unsigned q = 8u * XCHUNK;
            const unsigned nblk = (grid_for((nunits + a.upw - 1u) / a.upw, BLK / WAVE) + q - 1u) / q * q;
            k_stab<<<nblk, BLK, 0, s>>>(a);
            SC_HIP(hipGetLastError());
        }
        if (nitems > 0) {
            RowArgs r;
            r.ioff = srows; r.inode = inode;
            r.cov_off = c.cov_off; r.cov_split = c.cov_split; r.lend_off = c.lend_off;
            r.nl = c.nleaves; r.D = c.depth;
            r.cov_e = c.cov_e; r.lend_e = c.lend_e; r.cov_yr = c.cov_yr; r.lend_yr = c.lend_yr;
            r.lend_bl = c.lend_bl;
            r.nitems = nitems;
            r.out = c.near; r.cap = uint64_t(c.near_seg_cap); r.counter = ncnt;
            k_stab_rows<<<grid_for(nitems), BLK, 0, s>>>(r);
            SC_HIP(hipGetLastError());
        }
        if (nunits > 0 || nitems > 0) {
            ClassArgs b;
            b.near = c.near; b.seg_cnt = ncnt; b.seg_cap = uint64_t(c.near_seg_cap); b.summary = summary;
            b.exy = c.exy; b.etype = c.etype; b.eoff = c.d_eoff;
            b.bits_b = c.bits_b; b.bits_c = c.bits_c;
            b.out = c.cand; b.cap = uint64_t(c.cand_cap); b.counter = counter;
            const int64_t chunks = ceil_div(c.near_seg_cap * NEAR_SEGS, int64_t(BLK) * CL_PER);
            const unsigned ncl = unsigned(std::max<int64_t>(1, std::min<int64_t>(chunks, int64_t(cus) * 8)));
            k_near_classify<<<ncl, BLK, 0, s>>>(b);
            SC_HIP(hipGetLastError());
        }
        c.mark(9);
        uint64_t sm[2] = {0, 0}, nc = 0;
        if (nunits > 0 || nitems > 0) {
            ReadBack rb(c);
            rb.add(sm, reinterpret_cast<const uint64_t *>(summary), 2);
            rb.add(&nc, reinterpret_cast<const uint64_t *>(counter), 1);
            rb.run();
        }
        c.n_cand = int64_t(nc);
        c.n_near = int64_t(sm[0]);
        near_max = int64_t(sm[1]);
        if (near_max > c.near_seg_cap) {
            c.near_seg_cap = near_max + near_max / 4 + 1024;
            continue;
        }
        if (c.n_cand <= c.cand_cap) break;
        c.cand_cap = c.n_cand + c.n_cand / 4 + 1024;
    }
    SC_REQUIRE(near_max <= c.near_seg_cap && c.n_cand <= c.cand_cap, SC_EHIP, "stab buffer overflow after resize");

    // unique pairs sorted by (overlay, base edge, overlay-local edge)
    const int kb = c.bits_k + c.bits_b + c.bits_c;
    const int64_t n = c.n_cand;
    radix_sort_u64_buf(c.cand, c.cand_alt, nullptr, nullptr, n, 0, kb, c.rs, s);
    uint32_t *flag = c.u32b.ensure(size_t(n) + 1);
    {
        FillBatch fb(s);
        fb.add(flag, n + 1, 0u);
        fb.flush();
    }
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
