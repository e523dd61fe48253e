// sc_arrays.hip -- the reference's per-overlay array stages on caller arrays
// (host pointers in, host pointers out), run on the GPU.  These back the
// SegTree facade (segtree.h) for drivers that keep their own intermediate
// arrays between the stages, with the argument meaning of:
//   saveIntersectionsParallel         oacc_segment_tree.cpp:7257-7341
//   sortIntersectionsAtEdgesParallel  oacc:7436-7473
//   copyIntersectionsParallel         oacc:7508-7664
//   calculateInitLabelParallel        oacc:7925-7988
// Pair ids are the reference's (2 * vertex offset into bPoly / cPoly, both
// with their closing vertices), so the context's resident rings are used.
// Intended semantics where the reference reads memory it never wrote: all
// overlay source vertices are copied (the reference copies min(bSize,cSize)),
// and a pair never given a slot is reported (SEGCLIP_EUNDEF) instead of
// indexing with an uninitialised value.
#include "../../include/segclip.h"
#include "sc_ctx.hpp"
#include "sc_fp64.hpp"

#include <algorithm>

namespace sc {
namespace {

constexpr int BLK = 256;
inline unsigned grid_for(int64_t n) { return unsigned(ceil_div(n > 0 ? n : 1, BLK)); }

struct ArrScratch {
    DBuf<int> i1, i2, i3, i4, i5, i6, i7, i8, i9, flag;
    DBuf<double> d1, d2, d3;
    DBuf<uint64_t> keys;
    DBuf<uint32_t> vals;
    RadixScratch rs;
};
ArrScratch &ascr() {
    thread_local ArrScratch *p = new ArrScratch(); // kept for the thread's lifetime
    return *p;
}

template <class T> T *put(Ctx &c, DBuf<T> &d, const T *h, int64_t n) {
    T *p = d.ensure(size_t(n > 0 ? n : 1));
    if (n > 0) SC_HIP(hipMemcpyAsync(p, h, sizeof(T) * size_t(n), hipMemcpyHostToDevice, c.stream));
    return p;
}
template <class T> void get(Ctx &c, T *h, const T *d, int64_t n) {
    if (n > 0) SC_HIP(hipMemcpyAsync(h, d, sizeof(T) * size_t(n), hipMemcpyDeviceToHost, c.stream));
}

__global__ void k_save_ids(const double *bpoly, const double *cpoly, const int *bid, const int *cid, int n,
                           double *xy, int *type, int *qa, int *qb) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double *B = bpoly + bid[p], *C = cpoly + cid[p];
    double al, be;
    int t = classify(B[0], B[1], B[2], B[3], C[0], C[1], C[2], C[3], al, be);
    double x, y;
    intersection_point(t, al, B[0], B[1], B[2], B[3], C[0], C[1], x, y);
    xy[2 * p] = x;
    xy[2 * p + 1] = y;
    type[p] = t;
    qa[p] = alpha_to_int(al);
    qb[p] = alpha_to_int(be);
}

// per listed edge: its [s,e) range in the sorted order; sortable when it has
// more than one non-degenerate intersection
__global__ void k_edge_ranges(const int *uid, int nu, int c_start, const int *all, const int *nd, int ncnt, int n,
                              int *seg, int *sortable, int *bad) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nu) return;
    int eid = uid[id] / 2 - c_start;
    if (eid < 0 || eid + 1 >= ncnt) {
        atomicOr(bad, 1);
        return;
    }
    int s = all[eid], e = all[eid + 1];
    bool srt = nd[eid + 1] - nd[eid] > 1;
    if (s < 0 || e > n || s > e) {
        atomicOr(bad, 1);
        return;
    }
    for (int j = s; j < e; ++j) {
        seg[j] = s;
        sortable[j] = srt ? 1 : 0;
    }
}

__global__ void k_iota_seg(int n, int *seg, int *sortable) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    seg[i] = i;
    sortable[i] = 0;
}

// key: (range start, alpha modulo 1e9 -- the nine decimal digits the
// reference's LSD sort looks at) for sortable ranges, (position) otherwise
__global__ void k_sae_keys(int n, const int *seg, const int *sortable, const int *alpha, const int *sorted_index,
                           uint64_t *keys, uint32_t *vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = uint64_t(uint32_t(seg[i])) << 30;
    if (sortable[i]) k |= uint64_t(uint32_t(alpha[sorted_index[i]] % 1000000000));
    keys[i] = k;
    vals[i] = uint32_t(i);
}

// composite order: pair at sorted position i
__global__ void k_compose(int n, const int *sorted_index, const int *sorted, int *ord, int *bad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int a = sorted[i];
    if (a < 0 || a >= n) {
        atomicOr(bad, 1);
        ord[i] = 0;
        return;
    }
    int b = sorted_index[a];
    if (b < 0 || b >= n) {
        atomicOr(bad, 1);
        b = 0;
    }
    ord[i] = b;
}

__global__ void k_fill(int64_t n, int *a, int v, double *d) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    a[i] = v;
    if (d) {
        d[2 * i] = 0.0;
        d[2 * i + 1] = 0.0;
    }
}

__global__ void k_arr_sources(int size, const double *poly, int v0, const int *nd_ps, double *inter, int *alpha,
                              int *nbr) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= size) return;
    int pos = v + nd_ps[v];
    inter[2 * pos] = poly[2 * (int64_t(v0) + v)];
    inter[2 * pos + 1] = poly[2 * (int64_t(v0) + v) + 1];
    alpha[pos] = -100;
    nbr[pos] = -100;
}

// one thread per segment start of the sorted order, walking it like the
// reference's second loop (including its advance-only-before-the-last rule)
template <bool SIDE_B>
__global__ void k_arr_segments(int n, const int *ord, const int *ids, int c_start, const int *types,
                               const double *inters, const int *qalpha, const int *nd_ps, int size, int aug,
                               double *inter, int *alpha, int *slot_of, int *bad) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    int lid = ord[id];
    if (id > 0 && ids[ord[id - 1]] == ids[lid]) return;
    int pos = id;
    auto edge_of = [&](int l) { return ids[l] / 2 - c_start; };
    int t = types[lid];
    bool lead = SIDE_B ? degen_b_first(t) : degen_c_first(t);
    if (lead) {
        int e = edge_of(lid);
        if (e < 0 || e >= size) {
            atomicOr(bad, 1);
            return;
        }
        int src = e + nd_ps[e];
        alpha[src] = qalpha[lid];
        slot_of[lid] = src;
        if (pos < n - 1) lid = ord[++pos];
    }
    int e = edge_of(lid);
    if (e < 0 || e >= size) {
        atomicOr(bad, 1);
        return;
    }
    int start = nd_ps[e], end = nd_ps[e + 1];
    for (int i = 0, sl = start + e + 1; i < end - start; ++i, ++sl) {
        if (sl < 0 || sl >= aug) {
            atomicOr(bad, 1);
            return;
        }
        inter[2 * sl] = inters[2 * lid];
        inter[2 * sl + 1] = inters[2 * lid + 1];
        alpha[sl] = qalpha[lid];
        slot_of[lid] = sl;
        if (pos < n - 1) lid = ord[++pos];
    }
}

__global__ void k_arr_link(int n, const int *sb, const int *scn, int nba, int nca, int *nbr_b, int *nbr_c, int *bad) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int a = sb[p], b = scn[p];
    if (a < 0 || b < 0 || a >= nba || b >= nca) {
        atomicOr(bad, 1);
        return;
    }
    nbr_b[a] = b + 1;
    nbr_c[b] = a + 1;
}

__global__ void k_deinterleave(int64_t n, const double *xy, double *x, double *y) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = xy[2 * i];
    y[i] = xy[2 * i + 1];
}

int check_bad(Ctx &c, int *d_bad, const char *what) {
    int bad = 0;
    SC_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    SC_HIP(hipStreamSynchronize(c.stream));
    SC_REQUIRE(!bad, SC_EUNDEF, std::string(what) + ": inconsistent input arrays (an index outside its array)");
    return 0;
}

} // namespace

void arr_save_intersections(Ctx &c, const int *bid, const int *cid, int n, double *xy, int *type, int *qa,
                            int *qb) {
    SC_REQUIRE(c.ne > 0 && !c.paired, SC_ESTATE, "polygons not set (one base ring)");
    SC_REQUIRE(n >= 0 && (n == 0 || (bid && cid && xy && type && qa && qb)), SC_EINVAL, "null array");
    const int64_t nbd = 2 * int64_t(c.nb + 1), ncd = 2 * int64_t(c.nct + c.K);
    for (int p = 0; p < n; ++p)
        SC_REQUIRE(bid[p] >= 0 && (bid[p] & 1) == 0 && bid[p] + 3 < nbd && cid[p] >= 0 && (cid[p] & 1) == 0 &&
                       cid[p] + 3 < ncd,
                   SC_EINVAL, "pair id outside the polygons");
    if (n == 0) return;
    ArrScratch &a = ascr();
    int *db = put(c, a.i1, bid, n), *dc = put(c, a.i2, cid, n);
    double *dxy = a.d1.ensure(2 * size_t(n));
    int *dt = a.i3.ensure(n), *dqa = a.i4.ensure(n), *dqb = a.i5.ensure(n);
    k_save_ids<<<grid_for(n), BLK, 0, c.stream>>>(c.bpoly, c.cpoly, db, dc, n, dxy, dt, dqa, dqb);
    SC_HIP(hipGetLastError());
    get(c, xy, dxy, 2 * int64_t(n));
    get(c, type, dt, n);
    get(c, qa, dqa, n);
    get(c, qb, dqb, n);
    SC_HIP(hipStreamSynchronize(c.stream));
}

void arr_sort_at_edges(Ctx &c, const int *alpha, const int *sorted_index, int c_start, const int *all_counts,
                       const int *nd_counts, int ncounts, const int *unique_ids, int unique_count, int n, int *out) {
    SC_REQUIRE(n >= 0 && unique_count >= 0 && ncounts >= 1, SC_EINVAL, "bad sizes");
    if (n == 0) return;
    SC_REQUIRE(alpha && sorted_index && all_counts && nd_counts && out && (unique_count == 0 || unique_ids),
               SC_EINVAL, "null array");
    for (int i = 0; i < n; ++i)
        SC_REQUIRE(sorted_index[i] >= 0 && sorted_index[i] < n, SC_EINVAL, "sorted index outside the pair list");
    ArrScratch &a = ascr();
    hipStream_t s = c.stream;
    int *dal = put(c, a.i1, alpha, n), *dsi = put(c, a.i2, sorted_index, n);
    int *dall = put(c, a.i3, all_counts, ncounts), *dnd = put(c, a.i4, nd_counts, ncounts);
    int *duid = put(c, a.i5, unique_ids, unique_count);
    int *seg = a.i6.ensure(n), *srt = a.i7.ensure(n), *bad = a.flag.ensure(1);
    SC_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
    k_iota_seg<<<grid_for(n), BLK, 0, s>>>(n, seg, srt);
    k_edge_ranges<<<grid_for(unique_count), BLK, 0, s>>>(duid, unique_count, c_start, dall, dnd, ncounts, n, seg, srt,
                                                         bad);
    uint64_t *keys = a.keys.ensure(n);
    uint32_t *vals = a.vals.ensure(n);
    k_sae_keys<<<grid_for(n), BLK, 0, s>>>(n, seg, srt, dal, dsi, keys, vals);
    SC_HIP(hipGetLastError());
    check_bad(c, bad, "sortIntersectionsAtEdgesParallel");
    radix_sort_u64(keys, vals, n, 0, 30 + ilog2_exact(int64_t(n) + 1), a.rs, s);
    get(c, reinterpret_cast<uint32_t *>(out), vals, n);
    SC_HIP(hipStreamSynchronize(s));
}

void arr_copy_intersections(Ctx &c, int b_size, int c_size, int c_start, const int *b_ids, const int *c_ids, int n,
                            const int *b_sorted_index, const int *c_sorted_index, const int *sorted_b,
                            const int *sorted_c, const double *inters, const int *types, const int *b_nd_ps,
                            const int *c_nd_ps, const int *qa, const int *qb, double *inter_b, double *inter_c,
                            int *nbr_b, int *nbr_c, int *alpha_b, int *alpha_c) {
    SC_REQUIRE(c.ne > 0 && !c.paired, SC_ESTATE, "polygons not set (one base ring)");
    SC_REQUIRE(b_size == c.nb && c_size >= 3 && n >= 0, SC_EINVAL, "sizes do not match the resident polygons");
    SC_REQUIRE(c_start >= 0 && 2 * (int64_t(c_start) + c_size) < 2 * int64_t(c.nct + c.K), SC_EINVAL,
               "overlay outside the resident polygons");
    SC_REQUIRE(b_nd_ps && c_nd_ps, SC_EINVAL, "null array");
    for (int i = 0; i < b_size; ++i)
        SC_REQUIRE(b_nd_ps[i] >= 0 && b_nd_ps[i] <= b_nd_ps[i + 1], SC_EINVAL, "base counts are not a prefix sum");
    for (int i = 0; i < c_size; ++i)
        SC_REQUIRE(c_nd_ps[i] >= 0 && c_nd_ps[i] <= c_nd_ps[i + 1], SC_EINVAL, "overlay counts are not a prefix sum");
    const int nba = b_size + b_nd_ps[b_size], nca = c_size + c_nd_ps[c_size];
    ArrScratch &a = ascr();
    hipStream_t s = c.stream;
    int *bad = a.flag.ensure(1);
    SC_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
    int *dbnd = put(c, a.i1, b_nd_ps, b_size + 1), *dcnd = put(c, a.i2, c_nd_ps, c_size + 1);
    double *ib = a.d1.ensure(2 * size_t(nba)), *ic = a.d2.ensure(2 * size_t(nca));
    int *ab = a.i3.ensure(nba), *ac = a.i4.ensure(nca);
    int *nb_ = a.i5.ensure(nba), *nc_ = a.i6.ensure(nca);
    k_fill<<<grid_for(nba), BLK, 0, s>>>(nba, ab, -100, ib);
    k_fill<<<grid_for(nba), BLK, 0, s>>>(nba, nb_, -100, nullptr);
    k_fill<<<grid_for(nca), BLK, 0, s>>>(nca, ac, -100, ic);
    k_fill<<<grid_for(nca), BLK, 0, s>>>(nca, nc_, -100, nullptr);
    k_arr_sources<<<grid_for(b_size), BLK, 0, s>>>(b_size, c.bpoly, 0, dbnd, ib, ab, nb_);
    k_arr_sources<<<grid_for(c_size), BLK, 0, s>>>(c_size, c.cpoly, c_start, dcnd, ic, ac, nc_);
    SC_HIP(hipGetLastError());
    if (n > 0) {
        ArrScratch &t = a;
        DBuf<int> ordb_, ordc_, sb_, sc_, ids_b, ids_c, ty, qa_d, qb_d;
        DBuf<double> in_d;
        int *dbsi = put(c, ids_b, b_sorted_index, n), *dsb = put(c, sb_, sorted_b, n);
        int *ordb = ordb_.ensure(n), *ordc = ordc_.ensure(n);
        k_compose<<<grid_for(n), BLK, 0, s>>>(n, dbsi, dsb, ordb, bad);
        int *dcsi = put(c, ids_c, c_sorted_index, n), *dsc = put(c, sc_, sorted_c, n);
        k_compose<<<grid_for(n), BLK, 0, s>>>(n, dcsi, dsc, ordc, bad);
        int *dbid = put(c, t.i7, b_ids, n), *dcid = put(c, t.i8, c_ids, n), *dty = put(c, ty, types, n);
        double *dint = put(c, in_d, inters, 2 * int64_t(n));
        int *dqa = put(c, qa_d, qa, n), *dqb = put(c, qb_d, qb, n);
        int *slb = t.i9.ensure(2 * size_t(n)), *slc = slb + n;
        SC_HIP(hipMemsetAsync(slb, 0xff, sizeof(int) * 2 * size_t(n), s));
        k_arr_segments<true><<<grid_for(n), BLK, 0, s>>>(n, ordb, dbid, 0, dty, dint, dqa, dbnd, b_size, nba, ib, ab,
                                                        slb, bad);
        k_arr_segments<false><<<grid_for(n), BLK, 0, s>>>(n, ordc, dcid, c_start, dty, dint, dqb, dcnd, c_size, nca,
                                                         ic, ac, slc, bad);
        k_arr_link<<<grid_for(n), BLK, 0, s>>>(n, slb, slc, nba, nca, nb_, nc_, bad);
        SC_HIP(hipGetLastError());
        check_bad(c, bad, "copyIntersectionsParallel");
    }
    get(c, inter_b, ib, 2 * int64_t(nba));
    get(c, inter_c, ic, 2 * int64_t(nca));
    get(c, nbr_b, nb_, nba);
    get(c, nbr_c, nc_, nca);
    get(c, alpha_b, ab, nba);
    get(c, alpha_c, ac, nca);
    SC_HIP(hipStreamSynchronize(s));
}

void arr_init_labels(Ctx &c, int nba, int nca, const double *inter_b, const double *inter_c, const int *alpha_b,
                     const int *nbr_b, int *lab_b, int *lab_c) {
    SC_REQUIRE(nba >= 3 && nca >= 3, SC_EINVAL, "augmented rings need at least 3 vertices");
    ArrScratch &a = ascr();
    hipStream_t s = c.stream;
    double *ib = put(c, a.d1, inter_b, 2 * int64_t(nba)), *ic = put(c, a.d2, inter_c, 2 * int64_t(nca));
    double *xy = a.d3.ensure(2 * (size_t(nba) + nca));
    double *bx = xy, *by = xy + nba, *cx = xy + 2 * size_t(nba), *cy = cx + nca;
    k_deinterleave<<<grid_for(nba), BLK, 0, s>>>(nba, ib, bx, by);
    k_deinterleave<<<grid_for(nca), BLK, 0, s>>>(nca, ic, cx, cy);
    int *dab = put(c, a.i1, alpha_b, nba), *dnb = put(c, a.i2, nbr_b, nba);
    int *lb = a.i3.ensure(nba), *lc = a.i4.ensure(nca), *flags = a.flag.ensure(1);
    SC_HIP(hipMemsetAsync(flags, 0, sizeof(int), s));
    init_labels_one(c, nba, nca, bx, by, dab, dnb, cx, cy, lb, lc, flags);
    get(c, lab_b, lb, nba);
    get(c, lab_c, lc, nca);
    check_bad(c, flags, "calculateInitLabelParallel");
}

} // namespace sc
