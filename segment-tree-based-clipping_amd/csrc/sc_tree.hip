// sc_tree.hip -- input layout, elementary intervals, node intervals, leaf end
// entries and cover lists (rows A1-A7 of the hot path) for gfx950.
//
// Reference behaviour restated here (file:line in the reference repository):
//  - input layout and Edge records: src/clip.cpp:172-311, edge.h:190-201
//  - elementary intervals: oacc_segment_tree.cpp:4513-4692 (stable sort of
//    the 2N x-endpoints, keep the last of each run of equal values, the
//    vertical-edge duplicate rule with its own indexing, the special case
//    at position 2N-2, padding with the last kept value to 2^k+1 points)
//  - node intervals: oacc:4948-4994 (leaf l = [ea[l], ea[l+1]], parent =
//    [start(left), end(right)]), computed here in closed form per node
//  - leaf end entries: oacc:1970-2113 ("after" = leaf starting at e.end,
//    "before" = leaf ending at e.start, point-interval skip, degenerate
//    leaves excluded)
//  - cover lists: oacc:2496-2607.  e covers v iff Int(v) is inside
//    [e.start,e.end] and Int(parent) is not: the canonical decomposition of
//    the leaf range [L,R) with L = first point >= start, R = last point <= end.
#include "sc_ctx.hpp"

#include <algorithm>

namespace sc {

namespace {

constexpr int BLK = 256;

inline unsigned grid_for(int64_t n, int per = BLK) {
    int64_t g = ceil_div(n, per);
    if (g < 1) g = 1;
    if (g > (int64_t(1) << 30)) g = int64_t(1) << 30;
    return unsigned(g);
}

__device__ inline int find_overlay(const int *eoff, int K, int e) {
    // largest j with eoff[j] <= e, j in [0,K)
    int lo = 0, hi = K - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (eoff[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// overlay input vertex g (concatenated, no closing vertices) -> reference cPoly slot
__global__ void k_layout_overlays(const double *__restrict__ cxy, int nct, const int *__restrict__ eoff,
                                  const int *__restrict__ coff, int K, int nb, double *__restrict__ cpoly) {
    int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= nct) return;
    int j = find_overlay(eoff, K, int(g) + nb);
    int local = int(g) + nb - eoff[j];
    int dst = coff[j] + local;
    double x = cxy[2 * g], y = cxy[2 * g + 1];
    cpoly[2 * int64_t(dst)] = x;
    cpoly[2 * int64_t(dst) + 1] = y;
    if (local == 0) {
        int cl = coff[j] + (eoff[j + 1] - eoff[j]);
        cpoly[2 * int64_t(cl)] = x;
        cpoly[2 * int64_t(cl) + 1] = y;
    }
}


// smallest and largest x of a laid-out ring array (nv vertices), per block;
// min / max are exact, so the result does not depend on the reduction order
constexpr int XR_BLOCKS = 256;
__global__ void k_xrange_blocks(const double *__restrict__ poly, int64_t nv, double *bmin, double *bmax) {
    __shared__ double smin[BLK], smax[BLK];
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nv; i += int64_t(gridDim.x) * blockDim.x) {
        const double x = poly[2 * i];
        lo = fmin(lo, x);
        hi = fmax(hi, x);
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int d = BLK / 2; d > 0; d >>= 1) {
        if (int(threadIdx.x) < d) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + d]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + d]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bmin[blockIdx.x] = smin[0];
        bmax[blockIdx.x] = smax[0];
    }
}
__global__ void k_xrange_final(const double *__restrict__ bmin, const double *__restrict__ bmax, int n, double *out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double lo = bmin[0], hi = bmax[0];
    for (int i = 1; i < n; ++i) {
        lo = fmin(lo, bmin[i]);
        hi = fmax(hi, bmax[i]);
    }
    out[0] = lo;
    out[1] = hi;
}

// Edge records.  In paired mode the x-extents used as tree keys are moved by a
// per-pair shift so that different pairs occupy disjoint key ranges (no tree
// node is then covered by edges of two pairs); the endpoints used by every
// geometric predicate stay the input coordinates.
__global__ void k_edges(const double *__restrict__ bpoly, const double *__restrict__ cpoly, int nb, int ne,
                        const int *__restrict__ eoff, const int *__restrict__ coff, int K,
                        const int *__restrict__ beoff, const int *__restrict__ bvoff, int nbr,
                        const double *__restrict__ shift, double *x0, double *y0, double *x1, double *y1,
                        double *es, double *ee, double *eyl, double *eyh, int *ety, int *eid, double *uxs,
                        double *uxe, YRangeF *eyr, double4 *exy) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const double *P;
    int id, ty;
    double sh = 0.0;
    if (e < nb) {
        int j = find_overlay(beoff, nbr, e);
        P = bpoly;
        id = 2 * (bvoff[j] + (e - beoff[j]));
        ty = 0;
        if (shift) sh = shift[j];
    } else {
        int j = find_overlay(eoff, K, e);
        P = cpoly;
        id = 2 * (coff[j] + (e - eoff[j]));
        ty = j + 1;
        if (shift) sh = shift[j];
    }
    double a = P[id], b = P[id + 1], c = P[id + 2], d = P[id + 3];
    x0[e] = a;
    y0[e] = b;
    x1[e] = c;
    y1[e] = d;
    double ka = a + sh, kc = c + sh;
    if (ka < kc) {
        es[e] = ka;
        ee[e] = kc;
    } else {
        ee[e] = ka;
        es[e] = kc;
    }
    if (uxs) { // input x-extent for the bounding-box test
        uxs[e] = a < c ? a : c;
        uxe[e] = a < c ? c : a;
    }
    eyl[e] = b < d ? b : d;
    eyh[e] = b < d ? d : b;
    ety[e] = ty;
    eid[e] = id;
    eyr[e] = yrange_of(b, d);
    exy[e] = make_double4(a, b, c, d);
}

// ---- A3 ----
__global__ void k_endpoint_keys(const double *__restrict__ es, const double *__restrict__ ee, int ne, uint64_t *keys,
                                uint32_t *vals) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= 2 * int64_t(ne)) return;
    double v = (i & 1) ? ee[i >> 1] : es[i >> 1];
    keys[i] = f64_order_key(v);
    vals[i] = uint32_t(i);
}

__device__ inline double endpoint(const double *es, const double *ee, uint32_t s) { return (s & 1) ? ee[s >> 1] : es[s >> 1]; }

// verticalLines[i] = 1 for i = sortedIndex[j]/2, j < N, when edge i is vertical (oacc:4569-4577)
__global__ void k_vertical_flags(const uint32_t *__restrict__ sidx, const double *__restrict__ es,
                                 const double *__restrict__ ee, int ne, uint32_t *vl) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ne) return;
    int i = int(sidx[j] >> 1);
    if (es[i] == ee[i]) vl[i] = 1u;
}

__global__ void k_vertical_loc(const uint32_t *__restrict__ vl, int ne, uint32_t *vloc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne) return;
    if (vl[i] != vl[i + 1]) vloc[vl[i]] = uint32_t(i);
}

// unique flags with the reference's rules (oacc:4601-4635)
__global__ void k_unique_flags(const uint32_t *__restrict__ sidx, const double *__restrict__ es,
                               const double *__restrict__ ee, int ne, const uint32_t *__restrict__ vl,
                               const uint32_t *__restrict__ vloc, uint32_t *flag, double *val3) {
    int64_t m = 2 * int64_t(ne);
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= m - 1) return;
    uint32_t si = sidx[i], sn = sidx[i + 1];
    double a = endpoint(es, ee, si), b = endpoint(es, ee, sn);
    int intvi = int(si >> 1);
    uint32_t comid = vl[i / 2];
    uint32_t f = 0;
    double out = 0.0;
    if (a != b) {
        out = a;
        f = 1;
    } else if (es[intvi] == ee[intvi]) {
        bool ok = comid == 0;
        if (!ok) {
            int lk = int(vloc[comid - 1]);
            ok = es[lk] != es[intvi] && ee[lk] != es[intvi];
        }
        if (ok) {
            out = es[intvi];
            f = 1;
        }
    }
    flag[i] = f;
    val3[i] = out;
    if (i == m - 2) {
        double c = endpoint(es, ee, sidx[m - 2]), d = endpoint(es, ee, sidx[m - 1]);
        if (c != d) {
            flag[m - 1] = 1;
            val3[m - 1] = c;
        } else {
            flag[m - 1] = 0;
        }
        flag[m] = 0;
    }
}

__global__ void k_scatter_elementary(const uint32_t *__restrict__ pos, const double *__restrict__ val3, int64_t m,
                                     double *ea) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= m) return;
    if (pos[i] != pos[i + 1]) ea[pos[i]] = val3[i];
}

// Per endpoint (2e: the start of edge e, 2e + 1: its end) the elementary
// indices of its value: lo = the first index with ea[i] >= x, hi = the first
// with ea[i] > x, from its run of equal values among the sorted endpoints and
// the exclusive scan of the unique flags (pos).  Every run but the last two
// ends in a flag carrying its own value (the last flag carries the value
// before it, see k_unique_flags), so ea[lo] == x and nothing after hi equals
// x.  ER_NIL where the binary search must decide instead: the last two
// runs, NaN, runs longer than ER_WALK either side.
constexpr int ER_WALK = 8;
constexpr uint32_t ER_NIL = 0xffffffffu;
__global__ void k_endpoint_ranges(int64_t m, const uint64_t *__restrict__ keys, const uint32_t *__restrict__ sidx,
                                  const uint32_t *__restrict__ pos, uint2 *lohi) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const uint64_t k = keys[j], klast = keys[m - 1];
    // order keys of NaN: exponent all ones and a non-zero mantissa, either sign
    const uint64_t bits = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    bool ok = !((bits & 0x7ff0000000000000ull) == 0x7ff0000000000000ull && (bits & 0x000fffffffffffffull));
    int64_t j0 = j, j1 = j;
    for (int t = 0; t < ER_WALK && j0 > 0 && keys[j0 - 1] == k; ++t) --j0;
    for (int t = 0; t < ER_WALK && j1 + 1 < m && keys[j1 + 1] == k; ++t) ++j1;
    if (j0 > 0 && keys[j0 - 1] == k) ok = false;
    if (j1 + 1 < m && keys[j1 + 1] == k) ok = false;
    if (j1 + 1 >= m || keys[j1 + 1] == klast) ok = false; // the last run or the one before it
    lohi[sidx[j]] = ok ? make_uint2(pos[j0], pos[j1 + 1]) : make_uint2(ER_NIL, ER_NIL);
}

__global__ void k_pad_elementary(double *ea, int cnt, int round) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (cnt + i >= round) return;
    ea[cnt + i] = ea[cnt - 1];
}

// ---- A4 ----
__global__ void k_node_intervals(const double *__restrict__ ea, int T, int D, double *niv) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= T) return;
    if (v == 0) {
        niv[0] = 0.0;
        niv[1] = 0.0;
        return;
    }
    int d = 31 - __clz(v);          // depth of v
    int sh = D - 1 - d;             // log2 of leaves under v
    int first = (v - (1 << d)) << sh;
    niv[2 * v] = ea[first];
    niv[2 * v + 1] = ea[first + (1 << sh)];
}

// Two binary searches over a[lo0, n) run interleaved, so that their
// dependent loads overlap: r1 = first index with a[i] >= q1; r2 = first index
// with a[i] >= q2, or with a[i] > q2 when UPPER2.
template <bool UPPER2>
__device__ inline void search2(const double *a, int lo0, int n, double q1, double q2, int &r1, int &r2) {
    int l1 = lo0, h1 = n, l2 = lo0, h2 = n;
    while (l1 < h1 || l2 < h2) {
        const int m1 = (l1 + h1) >> 1, m2 = (l2 + h2) >> 1;
        const bool a1 = l1 < h1, a2 = l2 < h2;
        const double v1 = a1 ? a[m1] : 0.0, v2 = a2 ? a[m2] : 0.0;
        if (a1) {
            if (v1 < q1) l1 = m1 + 1;
            else h1 = m1;
        }
        if (a2) {
            if (UPPER2 ? v2 <= q2 : v2 < q2) l2 = m2 + 1;
            else h2 = m2;
        }
    }
    r1 = l1;
    r2 = l2;
}

// ---- A6 leaf end entries ----
// first i in [1, epc-1] with ea[i] == q, or -1 (intended findBoarder,
// oacc:420-460), for two queries at once
__device__ inline void first_point2(const double *ea, int epc, double q1, double q2, int &r1, int &r2) {
    int lo1, lo2;
    search2<false>(ea, 1, epc, q1, q2, lo1, lo2); // search [1, epc)
    r1 = lo1 < epc && ea[lo1] == q1 ? lo1 : -1;
    r2 = lo2 < epc && ea[lo2] == q2 ? lo2 : -1;
}


// per edge: "after"/"before" leaf node (or -1) and the two (class,leaf) keys of
// its leaf end entries; sentinel key 2*nleaves marks "not inserted"
__global__ void k_leaf_targets(const double *__restrict__ es, const double *__restrict__ ee,
                               const int *__restrict__ ety, int ne, const double *__restrict__ ea, int epc,
                               int nleaves, const uint2 *__restrict__ lohi, int *after, int *before,
                               uint64_t *keys, uint32_t *vals) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int half = nleaves, T = 2 * nleaves;
    int a = -1, b = -1;
    int ie, is;
    const uint2 rE = lohi[2 * int64_t(e) + 1], rS = lohi[2 * int64_t(e)];
    if (rE.x != ER_NIL && rS.x != ER_NIL) {
        // the search over [1, epc) finds lo when lo >= 1; for the first point
        // it looks at index 1 (epc >= 2)
        ie = rE.x >= 1u ? int(rE.x) : (ea[1] == ee[e] ? 1 : -1);
        is = rS.x >= 1u ? int(rS.x) : (ea[1] == es[e] ? 1 : -1);
    } else {
        first_point2(ea, epc, ee[e], es[e], ie, is);
    }
    if (ie >= 0) {
        int rl = half + ie;
        if (ie != 1 && ie + 1 <= epc - 1 && ea[ie] == ea[ie + 1]) rl = half + ie + 1;
        if (rl < T && ea[rl - half] != ea[rl - half + 1]) a = rl;
    }
    if (is >= 0) {
        int l = is - 1;
        if (ea[l] != ea[l + 1]) b = half + l;
    }
    after[e] = a;
    before[e] = b;
    uint64_t cls = ety[e] != 0 ? uint64_t(nleaves) : 0ull;
    const uint64_t none = 2ull * uint64_t(nleaves);
    keys[2 * int64_t(e)] = a >= 0 ? cls + uint64_t(a - half) : none;
    keys[2 * int64_t(e) + 1] = b >= 0 ? cls + uint64_t(b - half) : none;
    vals[2 * int64_t(e)] = uint32_t(e) | 0x80000000u; // "after" entry
    vals[2 * int64_t(e) + 1] = uint32_t(e);
}

// sorted keys: at the last entry of each run of key k < limit, end[k] = its
// position + 1 (an inclusive max-scan then fills the keys without entries)
__global__ void k_key_ends(const uint64_t *__restrict__ keys, int64_t n, uint64_t limit, uint32_t *end) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    if (k >= limit) return;
    if (i + 1 == n || keys[i + 1] != k) end[k] = uint32_t(i + 1);
}

// ---- A5 cover lists ----
// Cover list of an edge: the canonical decomposition of its leaf range
// [L,R) (L = first elementary point >= start, R = last point <= end), i.e. the
// nodes v with Int(v) inside [start,end] whose parent's interval is not
// (constructCoverListGPU, oacc:2958-3128).  EMIT = false counts them (and
// keeps L,R), EMIT = true writes (node, class) keys edge by edge.
template <bool EMIT>
__global__ void k_cover(const double *__restrict__ es, const double *__restrict__ ee, int ne, int nb,
                        const double *__restrict__ ea, int epc, int nleaves, const uint2 *__restrict__ lohi,
                        int ecount, uint32_t *__restrict__ cnt, const uint32_t *__restrict__ off,
                        uint64_t *__restrict__ keys, int2 *__restrict__ lr) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (!EMIT && e == 0) cnt[ne] = 0u; // the count scan's closing entry
    if (e >= ne) return;
    int L, R;
    if (!EMIT) {
        const uint2 rS = lohi[2 * int64_t(e)], rE = lohi[2 * int64_t(e) + 1];
        if (rS.x != ER_NIL && rE.x != ER_NIL) {
            // first index >= start; first index > end, the padding (copies
            // of the last point) counted when nothing real is larger
            L = int(rS.x);
            R = rE.y == uint32_t(ecount) ? epc : int(rE.y);
        } else {
            search2<true>(ea, 0, epc, es[e], ee[e], L, R);
        }
        R -= 1; // last point <= end
        lr[e] = make_int2(L, R);
    } else {
        int2 v = lr[e];
        L = v.x;
        R = v.y;
    }
    uint32_t n = 0, at = EMIT ? off[e] : 0u;
    const uint64_t cls = e < nb ? 0ull : 1ull;
    if (L < R) {
        int l = L + nleaves, r = R + nleaves;
        while (l < r) {
            if (l & 1) {
                if (EMIT) keys[at + n] = (((uint64_t(uint32_t(l)) << 1) | cls) << 32) | uint32_t(e);
                ++n;
                ++l;
            }
            if (r & 1) {
                --r;
                if (EMIT) keys[at + n] = (((uint64_t(uint32_t(r)) << 1) | cls) << 32) | uint32_t(e);
                ++n;
            }
            l >>= 1;
            r >>= 1;
        }
    }
    if (!EMIT) cnt[e] = n;
}

__global__ void k_cover_split(int T, const uint32_t *__restrict__ off, const uint32_t *__restrict__ bend,
                              uint32_t *split) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= T) return;
    split[v] = bend[v] ? bend[v] - off[v] : 0u;
}


// ---- materialised end lists (parity and drop-in API) ----
__global__ void k_mend_keys(const uint64_t *__restrict__ lkeys, const uint32_t *__restrict__ lvals, int64_t n,
                            int nleaves, int D, const int *__restrict__ eid, const int *__restrict__ ety, int idb,
                            int tb, uint64_t *out) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = lkeys[i];
    int leaf = int(k >= uint64_t(nleaves) ? k - uint64_t(nleaves) : k);
    uint32_t e = lvals[i] & 0x7fffffffu;
    uint64_t idt = (uint64_t(eid[e]) << tb) | uint64_t(ety[e]);
    int node = nleaves + leaf;
    for (int d = 0; d < D; ++d) {
        out[i * D + d] = (uint64_t(node) << (idb + tb)) | idt;
        node >>= 1;
    }
}

__global__ void k_diff_flags(const uint64_t *__restrict__ k, int64_t n, uint32_t *f) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f[i] = (i == 0 || k[i] != k[i - 1]) ? 1u : 0u;
    if (i == n - 1) f[n] = 0u;
}

// Keys are sorted by node, so the last key of each node's run writes where the
// node's distinct entries end (pos after it); a max-scan over nodes then gives
// every node's offset, empty nodes included. (A counter per node took one
// same-address atomic per entry: the root's entries all serialised.)
__global__ void k_mend_scatter(const uint64_t *__restrict__ k, const uint32_t *__restrict__ pos, int64_t n,
                               int idb, int tb, const int *__restrict__ eoff, const int *__restrict__ coff,
                               uint32_t *out_e, uint32_t *node_end) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t key = k[i];
    int node = int(key >> (idb + tb));
    if (i == n - 1 || int(k[i + 1] >> (idb + tb)) != node) node_end[node] = pos[i + 1];
    if (pos[i] == pos[i + 1]) return;
    int type = int(key & ((1ull << tb) - 1ull));
    int id = int((key >> tb) & ((1ull << idb) - 1ull));
    int e = type == 0 ? id / 2 : eoff[type - 1] + (id / 2 - coff[type - 1]);
    out_e[pos[i]] = uint32_t(e);
}

} // namespace

uint32_t read_u32(Ctx &c, const uint32_t *d) {
    uint32_t h = 0;
    ReadBack rb(c);
    rb.add(&h, d, 1);
    rb.run();
    return h;
}
uint64_t read_u64(Ctx &c, const uint64_t *d) {
    uint64_t h = 0;
    ReadBack rb(c);
    rb.add(&h, d, 1);
    rb.run();
    return h;
}

namespace {
// lay out a set of rings (concatenated, no closing vertices) with closing vertices
void layout_rings(Ctx &c, const double *src, int64_t nverts, const std::vector<int> &sizes,
                  const std::vector<int> &voff, DBuf<int> &d_voff, std::vector<int> &edge_off, int edge_base,
                  DBuf<int> &d_eoff, DBuf<double> &dst, bool device_ptrs) {
    hipStream_t s = c.stream;
    const int R = int(sizes.size());
    d_voff.ensure(R + 1);
    d_eoff.ensure(R + 1);
    SC_HIP(hipMemcpyAsync(d_voff.p, voff.data(), sizeof(int) * (R + 1), hipMemcpyHostToDevice, s));
    SC_HIP(hipMemcpyAsync(d_eoff.p, edge_off.data(), sizeof(int) * (R + 1), hipMemcpyHostToDevice, s));
    dst.ensure(2 * size_t(nverts + R));
    double *stage = reinterpret_cast<double *>(c.u64a.ensure(2 * size_t(nverts)));
    SC_HIP(hipMemcpyAsync(stage, src, sizeof(double) * 2 * size_t(nverts),
                          device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    k_layout_overlays<<<grid_for(nverts), BLK, 0, s>>>(stage, int(nverts), d_eoff.p, d_voff.p, R, edge_base, dst.p);
    SC_HIP(hipGetLastError());
    SC_HIP(hipStreamSynchronize(s)); // staging buffer is reused by the next call
}

void reset_derived(Ctx &c) {
    c.epc = c.nleaves = c.T = c.ea_count = 0;
    c.n_lend = c.n_cov = c.n_cand = c.n_pairs = c.n_mend = c.n_near = 0;
    c.cur_k = 0;
    c.pair_off.clear();
}

void load(Ctx &c, const double *bxy, const int *bsz, int nbr, const double *cxy, const int *csz, int K,
          bool paired, bool device_ptrs) {
    SC_REQUIRE(bxy && cxy && csz && bsz, SC_EINVAL, "null polygon pointer");
    pairs_graph_reset(c); // a captured forward refers to the old rings' sizes
    SC_REQUIRE(K >= 1 && nbr >= 1, SC_EINVAL, "at least one base and one overlay polygon are required");
    int64_t nb = 0, nct = 0;
    for (int i = 0; i < nbr; ++i) {
        SC_REQUIRE(bsz[i] >= 3, SC_EINVAL, "base polygon " + std::to_string(i) + " needs at least 3 vertices");
        nb += bsz[i];
    }
    for (int i = 0; i < K; ++i) {
        SC_REQUIRE(csz[i] >= 3, SC_EINVAL, "overlay polygon " + std::to_string(i + 1) + " needs at least 3 vertices");
        nct += csz[i];
    }
    SC_REQUIRE(2 * (nb + nct + K + nbr) < (int64_t(1) << 31), SC_ERANGE, "too many vertices for 32-bit ids");
    c.nb = int(nb);
    c.nbr = nbr;
    c.K = K;
    c.paired = paired;
    c.nct = int(nct);
    c.ne = int(nb + nct);
    c.bsz.assign(bsz, bsz + nbr);
    c.bvoff.assign(nbr + 1, 0);
    c.beoff.assign(nbr + 1, 0);
    c.maxbsz = 0;
    for (int i = 0; i < nbr; ++i) {
        c.bvoff[i + 1] = c.bvoff[i] + bsz[i] + 1;
        c.beoff[i + 1] = c.beoff[i] + bsz[i];
        c.maxbsz = std::max(c.maxbsz, bsz[i]);
    }
    c.csz.assign(csz, csz + K);
    c.coff.assign(K + 1, 0);
    c.eoff.assign(K + 1, 0);
    c.eoff[0] = c.nb;
    c.maxcsz = 0;
    for (int i = 0; i < K; ++i) {
        c.coff[i + 1] = c.coff[i] + csz[i] + 1;
        c.eoff[i + 1] = c.eoff[i] + csz[i];
        c.maxcsz = std::max(c.maxcsz, csz[i]);
    }
    layout_rings(c, bxy, nb, c.bsz, c.bvoff, c.d_bvoff, c.beoff, 0, c.d_beoff, c.bpoly, device_ptrs);
    layout_rings(c, cxy, nct, c.csz, c.coff, c.d_coff, c.eoff, c.nb, c.d_eoff, c.cpoly, device_ptrs);
    if (paired) {
        // per-pair key shift: pairs are laid side by side along x, each in a
        // window wider than the span of all rings
        // x-range of all rings, reduced on the device from the laid-out rings
        // (their closing vertices repeat first vertices: the same range)
        double lo = 0, hi = 0;
        {
            double *part = reinterpret_cast<double *>(c.u64b.ensure(4 * XR_BLOCKS + 2));
            k_xrange_blocks<<<XR_BLOCKS, BLK, 0, c.stream>>>(c.bpoly, nb + nbr, part, part + 2 * XR_BLOCKS);
            k_xrange_blocks<<<XR_BLOCKS, BLK, 0, c.stream>>>(c.cpoly, nct + K, part + XR_BLOCKS,
                                                             part + 3 * XR_BLOCKS);
            k_xrange_final<<<1, 1, 0, c.stream>>>(part, part + 2 * XR_BLOCKS, 2 * XR_BLOCKS, part + 4 * XR_BLOCKS);
            SC_HIP(hipGetLastError());
            double h[2];
            ReadBack rb(c);
            rb.add(h, static_cast<const double *>(part + 4 * XR_BLOCKS), 2);
            rb.run();
            lo = h[0];
            hi = h[1];
        }
        double span = (hi - lo) + 1.0;
        std::vector<double> sh(static_cast<size_t>(K));
        for (int k = 0; k < K; ++k) sh[k] = double(k) * span - lo;
        c.d_shift.ensure(K);
        SC_HIP(hipMemcpyAsync(c.d_shift.p, sh.data(), sizeof(double) * K, hipMemcpyHostToDevice, c.stream));
        SC_HIP(hipStreamSynchronize(c.stream));
    }
    reset_derived(c);
}
} // namespace

void ctx_set_polygons(Ctx &c, const double *bxy, int nb, const double *cxy, const int *csz, int K, bool device_ptrs) {
    SC_REQUIRE(bxy && cxy && csz, SC_EINVAL, "null polygon pointer");
    SC_REQUIRE(nb >= 3, SC_EINVAL, "base polygon needs at least 3 vertices");
    load(c, bxy, &nb, 1, cxy, csz, K, false, device_ptrs);
}

void ctx_set_pairs(Ctx &c, const double *bxy, const int *bsz, const double *cxy, const int *csz, int P,
                   bool device_ptrs) {
    load(c, bxy, bsz, P, cxy, csz, P, true, device_ptrs);
}

void build_edges(Ctx &c) {
    SC_REQUIRE(c.ne > 0, SC_ESTATE, "polygons not set");
    int ne = c.ne;
    c.ex0.ensure(ne); c.ey0.ensure(ne); c.ex1.ensure(ne); c.ey1.ensure(ne);
    c.estart.ensure(ne); c.eend.ensure(ne); c.etype.ensure(ne); c.eid.ensure(ne);
    c.eymin.ensure(ne); c.eymax.ensure(ne); c.eyr.ensure(ne); c.exy.ensure(ne);
    if (c.paired) {
        c.uxs.ensure(ne);
        c.uxe.ensure(ne);
    }
    k_edges<<<grid_for(ne), BLK, 0, c.stream>>>(c.bpoly, c.cpoly, c.nb, ne, c.d_eoff, c.d_coff, c.K, c.d_beoff,
                                                c.d_bvoff, c.nbr, c.paired ? c.d_shift.p : nullptr, c.ex0, c.ey0,
                                                c.ex1, c.ey1, c.estart, c.eend, c.eymin, c.eymax, c.etype, c.eid,
                                                c.paired ? c.uxs.p : nullptr, c.paired ? c.uxe.p : nullptr,
                                                c.eyr, c.exy);
    SC_HIP(hipGetLastError());
}

void build_elementary(Ctx &c) {
    SC_REQUIRE(c.ne > 0, SC_ESTATE, "edges not built");
    hipStream_t s = c.stream;
    const int ne = c.ne;
    const int64_t m = 2 * int64_t(ne);
    uint64_t *keys = c.k64a.ensure(m);
    uint32_t *sidx = c.u32a.ensure(m);
    k_endpoint_keys<<<grid_for(m), BLK, 0, s>>>(c.estart, c.eend, ne, keys, sidx);
    SC_HIP(hipGetLastError());
    radix_sort_u64(keys, sidx, m, 0, 64, c.rs, s);
    // vertical-edge bookkeeping
    uint32_t *vl = c.u32b.ensure(ne + 1);
    uint32_t *flag = c.u32d.ensure(m + 1);
    {
        FillBatch fb(s); // one launch for both (a memset of an odd size is two)
        fb.add(vl, ne + 1, 0u);
        fb.add(flag, m + 1, 0u);
        fb.flush();
    }
    k_vertical_flags<<<grid_for(ne), BLK, 0, s>>>(sidx, c.estart, c.eend, ne, vl);
    SC_HIP(hipGetLastError());
    scan_exclusive_u32(vl, vl, ne + 1, nullptr, s);
    uint32_t *vloc = c.u32c.ensure(ne + 1);
    k_vertical_loc<<<grid_for(ne), BLK, 0, s>>>(vl, ne, vloc);
    SC_HIP(hipGetLastError());
    // unique flags, compaction
    double *val3 = reinterpret_cast<double *>(c.u64b.ensure(m));
    k_unique_flags<<<grid_for(m), BLK, 0, s>>>(sidx, c.estart, c.eend, ne, vl, vloc, flag, val3);
    SC_HIP(hipGetLastError());
    uint32_t *tot = c.d_u32_total.ensure(1);
    scan_exclusive_u32(flag, flag, m + 1, tot, s);
    k_endpoint_ranges<<<grid_for(m), BLK, 0, s>>>(m, keys, sidx, flag, c.ep_lohi.ensure(size_t(m)));
    SC_HIP(hipGetLastError());
    int cnt = int(read_u32(c, tot));
    SC_REQUIRE(cnt >= 1, SC_EINVAL, "degenerate input: fewer than two distinct x values");
    int round = next_pow2_int(cnt) + 1;
    c.ea.ensure(round);
    k_scatter_elementary<<<grid_for(m), BLK, 0, s>>>(flag, val3, m, c.ea);
    SC_HIP(hipGetLastError());
    if (round > cnt) {
        k_pad_elementary<<<grid_for(round - cnt), BLK, 0, s>>>(c.ea, cnt, round);
        SC_HIP(hipGetLastError());
    }
    c.epc = round;
    c.ea_count = cnt;
    c.nleaves = round - 1;
    c.T = 2 * c.nleaves;
    c.depth = ilog2_exact(c.T);
}

void build_node_intervals(Ctx &c) {
    SC_REQUIRE(c.T > 0, SC_ESTATE, "elementary intervals not built");
    c.niv.ensure(2 * size_t(c.T));
    k_node_intervals<<<grid_for(c.T), BLK, 0, c.stream>>>(c.ea, c.T, c.depth, c.niv);
    SC_HIP(hipGetLastError());
}

namespace {
// beside every leaf end entry: its edge's "before" leaf and y-extent
__global__ void k_entry_before(const uint32_t *__restrict__ e, int64_t n, const int *__restrict__ before,
                               const YRangeF *__restrict__ eyr, int *bl, YRangeF *yr) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = e[i] & 0x7fffffffu;
    bl[i] = before[k];
    yr[i] = eyr[k];
}
// cover entries from their sorted packed keys: the edge (low word) and its
// y-extent; at the last entry of a (node, class) run, the node's end and (for
// the base class) its base end
__global__ void k_cover_entries(const uint64_t *__restrict__ keys, int64_t n, const YRangeF *__restrict__ eyr,
                                uint32_t *ce, YRangeF *yr, uint32_t *end, uint32_t *bend) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    const uint32_t e = uint32_t(k);
    ce[i] = e;
    yr[i] = eyr[e];
    const uint64_t hk = k >> 32;
    const uint64_t nk = i + 1 == n ? ~0ull : keys[i + 1] >> 32;
    if (nk == hk) return;
    const uint32_t v = uint32_t(hk >> 1);
    if (i + 1 == n || (nk >> 1) != v) end[v] = uint32_t(i + 1);
    if ((hk & 1ull) == 0) bend[v] = uint32_t(i + 1);
}
} // namespace

void build_leaf_end_entries(Ctx &c) {
    SC_REQUIRE(c.T > 0, SC_ESTATE, "elementary intervals not built");
    hipStream_t s = c.stream;
    const int ne = c.ne, nl = c.nleaves;
    const int64_t m = 2 * int64_t(ne);
    c.after_leaf.ensure(ne);
    c.before_leaf.ensure(ne);
    uint64_t *keys = c.lend_key.ensure(m);
    uint32_t *vals = c.lend_e.ensure(m);
    k_leaf_targets<<<grid_for(ne), BLK, 0, s>>>(c.estart, c.eend, c.etype, ne, c.ea, c.epc, nl, c.ep_lohi,
                                                c.after_leaf, c.before_leaf, keys, vals);
    SC_HIP(hipGetLastError());
    int kb = ilog2_exact(2 * int64_t(nl) + 1);
    radix_sort_u64_buf(c.lend_key, c.lend_key_alt, &c.lend_e, &c.lend_e_alt, m, 0, kb, c.rs, s);
    keys = c.lend_key.p;
    vals = c.lend_e.p;
    uint32_t *off = c.lend_off.ensure(2 * size_t(nl) + 1);
    // off[k] = entries with key < k: run ends of the sorted keys, then a running maximum
    {
        FillBatch fb(s);
        fb.add(off, 2 * int64_t(nl) + 1, 0u);
        fb.flush();
    }
    k_key_ends<<<grid_for(m), BLK, 0, s>>>(keys, m, 2ull * uint64_t(nl), off + 1);
    SC_HIP(hipGetLastError());
    scan_inclusive_max_u32(off + 1, off + 1, 2 * int64_t(nl), s);
    c.n_lend = read_u32(c, off + 2 * size_t(nl));
    // beside every entry, its edge's "before" leaf (the stab's dedupe test) and y-extent
    int *bl = c.lend_bl.ensure(size_t(c.n_lend > 0 ? c.n_lend : 1));
    YRangeF *bx = c.lend_yr.ensure(size_t(c.n_lend > 0 ? c.n_lend : 1));
    if (c.n_lend > 0) {
        k_entry_before<<<grid_for(c.n_lend), BLK, 0, s>>>(vals, c.n_lend, c.before_leaf, c.eyr, bl, bx);
        SC_HIP(hipGetLastError());
    }
}

void build_cover_lists(Ctx &c) {
    SC_REQUIRE(c.T > 0, SC_ESTATE, "elementary intervals not built");
    hipStream_t s = c.stream;
    const int T = c.T;
    const int ne = c.ne;
    // per-edge entry counts -> edge-major offsets
    uint32_t *cnt = c.u32b.ensure(size_t(ne) + 1);
    int2 *lr = reinterpret_cast<int2 *>(c.u64b.ensure(ne));
    k_cover<false><<<grid_for(ne), BLK, 0, s>>>(c.estart, c.eend, ne, c.nb, c.ea, c.epc, c.nleaves, c.ep_lohi,
                                                c.ea_count, cnt, nullptr, nullptr, lr);
    SC_HIP(hipGetLastError());
    uint32_t *tot = c.d_u32_total.ensure(1);
    scan_exclusive_u32(cnt, cnt, int64_t(ne) + 1, tot, s);
    c.n_cov = read_u32(c, tot);
    const int64_t L = c.n_cov;
    // (node, class) << 32 | edge keys, edge-major, then a stable radix sort on
    // the (node, class) bits: within a node base entries come first, each run
    // in edge order
    uint64_t *keys = c.cov_key.ensure(size_t(L > 0 ? L : 1));
    c.cov_e.ensure(size_t(L > 0 ? L : 1));
    if (L > 0) {
        k_cover<true><<<grid_for(ne), BLK, 0, s>>>(c.estart, c.eend, ne, c.nb, c.ea, c.epc, c.nleaves, c.ep_lohi,
                                                   c.ea_count, nullptr, cnt, keys, lr);
        SC_HIP(hipGetLastError());
        const int kb = ilog2_exact(T) + 1;
        SC_REQUIRE(kb <= 32, SC_ERANGE, "cover key too wide");
        radix_sort_u64_buf(c.cov_key, c.cov_key_alt, nullptr, nullptr, L, 32, 32 + kb, c.rs, s);
    }
    uint32_t *off = c.cov_off.ensure(size_t(T) + 1);
    uint32_t *split = c.cov_split.ensure(T);
    // off[v+1] = largest run end among nodes <= v; base runs end at bend
    uint32_t *bend = reinterpret_cast<uint32_t *>(c.u64b.ensure(size_t(T) / 2 + 1)); // T u32 (lr is done)
    {
        FillBatch fb(s);
        fb.add(off, int64_t(T) + 1, 0u);
        fb.add(bend, T, 0u);
        fb.flush();
    }
    if (L > 0) {
        k_cover_entries<<<grid_for(L), BLK, 0, s>>>(c.cov_key, L, c.eyr, c.cov_e, c.cov_yr.ensure(size_t(L)), off + 1,
                                                    bend);
        SC_HIP(hipGetLastError());
    }
    scan_inclusive_max_u32(off + 1, off + 1, T, s);
    k_cover_split<<<grid_for(T), BLK, 0, s>>>(T, off, bend, split);
    SC_HIP(hipGetLastError());
}

void materialize_end_lists(Ctx &c) {
    SC_REQUIRE(c.n_lend >= 0 && c.T > 0, SC_ESTATE, "leaf end entries not built");
    hipStream_t s = c.stream;
    const int D = c.depth;
    int64_t n = c.n_lend * D;
    int idb = ilog2_exact(2 * (int64_t(c.nct) + c.K + c.nb) + 2);
    int tb = ilog2_exact(int64_t(c.K) + 1);
    SC_REQUIRE(D + idb + tb <= 64, SC_ERANGE, "end-list key does not fit in 64 bits");
    uint64_t *keys = c.u64a.ensure(size_t(n > 0 ? n : 1));
    // leaf keys/vals are still sorted in lend_key / lend_e (first n_lend valid entries)
    k_mend_keys<<<grid_for(c.n_lend), BLK, 0, s>>>(c.lend_key, c.lend_e, c.n_lend, c.nleaves, D, c.eid, c.etype, idb, tb,
                                                   keys);
    SC_HIP(hipGetLastError());
    radix_sort_u64(keys, nullptr, n, 0, D + idb + tb, c.rs, s);
    uint32_t *flag = c.u32d.ensure(size_t(n) + 1);
    SC_HIP(hipMemsetAsync(flag, 0, sizeof(uint32_t) * (size_t(n) + 1), s));
    if (n > 0) {
        k_diff_flags<<<grid_for(n), BLK, 0, s>>>(keys, n, flag);
        SC_HIP(hipGetLastError());
    }
    uint32_t *tot = c.d_u32_total.ensure(1);
    scan_exclusive_u32(flag, flag, n + 1, tot, s);
    c.n_mend = read_u32(c, tot);
    c.mend_e.ensure(size_t(c.n_mend));
    // off[v + 1] = end of node v's entries; off[v] after the max-scan = entries of nodes < v
    uint32_t *off = c.mend_off.ensure(size_t(c.T) + 1);
    SC_HIP(hipMemsetAsync(off, 0, sizeof(uint32_t) * (size_t(c.T) + 1), s));
    if (n > 0) {
        k_mend_scatter<<<grid_for(n), BLK, 0, s>>>(keys, flag, n, idb, tb, c.d_eoff, c.d_coff, c.mend_e, off + 1);
        SC_HIP(hipGetLastError());
    }
    scan_inclusive_max_u32(off + 1, off + 1, c.T, s);
}

} // namespace sc
