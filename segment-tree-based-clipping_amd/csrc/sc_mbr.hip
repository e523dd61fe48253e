// sc_mbr.hip -- candidate polygon pairs of two layers by bounding-box overlap,
// the filter in front of many-to-many clipping.
//
// Reference: the data preparation joins two layers through an R-tree of
// polygon MBRs (lib/optimizedFostersAlgorithm/readShapefile.cpp:326-355
// getMBR, the librtree search in foster/Rtree-Topdown-version) and the
// clipper drops a pair whose MBRs are disjoint (polyclip_time.h:95-124,
// getCMBR: closed comparisons).  Here: one wave per ring computes its MBR,
// and the join is an interval join on x with a y test: with both layers
// sorted by min-x, an A box takes the B boxes whose min-x lies in
// [minx_A, maxx_A] and a B box the A boxes whose min-x lies in (minx_B,
// maxx_B]; the two sets are disjoint and hold every x-overlapping pair once.
// Work is proportional to the x-overlapping pairs, not to |A| x |B|.  The
// pairs come out sorted by (A index, B index).
#include "sc_ctx.hpp"
#include "../../include/segclip.h"

namespace sc {

namespace {

constexpr int BLK = 256;
inline unsigned grid_for(int64_t n, int per = BLK) {
    int64_t g = ceil_div(n, per);
    if (g < 1) g = 1;
    return unsigned(g);
}

// MBR of ring r (vertices off[r] .. off[r+1]-1 of xy, no closing vertex): one
// wave per ring; min / max are exact, so the lanes may combine in any order
__global__ void k_ring_mbr(const double *__restrict__ xy, const int64_t *__restrict__ off, int R, double *mbr) {
    const int r = int(blockIdx.x) * (BLK / WAVE) + int(threadIdx.x) / WAVE;
    const int lane = int(threadIdx.x) & (WAVE - 1);
    if (r >= R) return;
    double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
    for (int64_t i = off[r] + lane; i < off[r + 1]; i += WAVE) {
        const double x = xy[2 * i], y = xy[2 * i + 1];
        x0 = fmin(x0, x);
        x1 = fmax(x1, x);
        y0 = fmin(y0, y);
        y1 = fmax(y1, y);
    }
    for (int d = WAVE / 2; d > 0; d >>= 1) {
        x0 = fmin(x0, __shfl_xor(x0, d, WAVE));
        y0 = fmin(y0, __shfl_xor(y0, d, WAVE));
        x1 = fmax(x1, __shfl_xor(x1, d, WAVE));
        y1 = fmax(y1, __shfl_xor(y1, d, WAVE));
    }
    if (lane == 0) {
        mbr[4 * int64_t(r)] = x0;
        mbr[4 * int64_t(r) + 1] = y0;
        mbr[4 * int64_t(r) + 2] = x1;
        mbr[4 * int64_t(r) + 3] = y1;
    }
}

__global__ void k_minx_keys(const double *__restrict__ mbr, int n, uint64_t *key, uint32_t *idx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = f64_order_key(mbr[4 * int64_t(i)]);
    idx[i] = uint32_t(i);
}
__global__ void k_gather_minx(const double *__restrict__ mbr, const uint32_t *__restrict__ idx, int n, double *mx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) mx[i] = mbr[4 * int64_t(idx[i])];
}

// first i in [0, n) with v[i] > x (upper) or v[i] >= x (lower)
__device__ inline int bound(const double *v, int n, double x, bool upper) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (upper ? (v[mid] <= x) : (v[mid] < x)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// For every box of layer S: the boxes of layer O whose min-x lies in
// [minx_S, maxx_S] (lower_closed) or (minx_S, maxx_S], with their y-extents
// overlapping (closed).  count: per S box; emit: packed (A, B) keys at offsets.
template <bool EMIT>
__global__ void k_join(const double *__restrict__ smbr, int ns, const double *__restrict__ ombr,
                       const double *__restrict__ osorted_minx, const uint32_t *__restrict__ oidx, int no,
                       bool lower_closed, bool s_is_a, uint32_t *count, const uint32_t *__restrict__ off,
                       uint64_t *out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const double sx0 = smbr[4 * int64_t(s)], sy0 = smbr[4 * int64_t(s) + 1], sx1 = smbr[4 * int64_t(s) + 2],
                 sy1 = smbr[4 * int64_t(s) + 3];
    const int lo = bound(osorted_minx, no, sx0, !lower_closed), hi = bound(osorted_minx, no, sx1, true);
    uint32_t c = 0;
    uint64_t at = EMIT ? off[s] : 0;
    for (int k = lo; k < hi; ++k) {
        const int o = int(oidx[k]);
        const double oy0 = ombr[4 * int64_t(o) + 1], oy1 = ombr[4 * int64_t(o) + 3];
        if (sy0 > oy1 || sy1 < oy0) continue;
        if (EMIT) {
            const uint32_t a = s_is_a ? uint32_t(s) : uint32_t(o), b = s_is_a ? uint32_t(o) : uint32_t(s);
            out[at++] = (uint64_t(a) << 32) | b;
        } else {
            ++c;
        }
    }
    if (!EMIT) count[s] = c;
}

__global__ void k_unpack_pairs(const uint64_t *__restrict__ key, int64_t n, int32_t *ab) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ab[2 * i] = int32_t(key[i] >> 32);
    ab[2 * i + 1] = int32_t(key[i] & 0xffffffffu);
}

struct MbrScratch {
    DBuf<int64_t> offa, offb;
    DBuf<double> mbra, mbrb, sxa, sxb;
    DBuf<uint64_t> ka, kb, pairs, pairs_alt;
    DBuf<uint32_t> ia, ib, ca, cb, oa, ob, tot;
    DBuf<int32_t> ab;
    RadixScratch rs;
};
MbrScratch &mscratch() {
    thread_local MbrScratch *p = new MbrScratch();
    return *p;
}

// MBRs of a layer's rings on the device and the layer sorted by min-x
void layer_boxes(const double *xy, const int *sizes, int n, DBuf<int64_t> &doff, DBuf<double> &mbr,
                 DBuf<uint64_t> &key, DBuf<uint32_t> &idx, DBuf<double> &sx, RadixScratch &rs, hipStream_t s) {
    std::vector<int64_t> off(size_t(n) + 1, 0);
    for (int i = 0; i < n; ++i) {
        SC_REQUIRE(sizes[i] >= 1, SC_EINVAL, "a ring without vertices");
        off[i + 1] = off[i] + sizes[i];
    }
    doff.ensure(size_t(n) + 1);
    SC_HIP(hipMemcpyAsync(doff.p, off.data(), sizeof(int64_t) * off.size(), hipMemcpyHostToDevice, s));
    mbr.ensure(4 * size_t(n));
    k_ring_mbr<<<grid_for(n, BLK / WAVE), BLK, 0, s>>>(xy, doff.p, n, mbr.p);
    key.ensure(size_t(n));
    idx.ensure(size_t(n));
    k_minx_keys<<<grid_for(n), BLK, 0, s>>>(mbr.p, n, key.p, idx.p);
    SC_HIP(hipGetLastError());
    radix_sort_u64(key.p, idx.p, n, 0, 64, rs, s);
    sx.ensure(size_t(n));
    k_gather_minx<<<grid_for(n), BLK, 0, s>>>(mbr.p, idx.p, n, sx.p);
    SC_HIP(hipGetLastError());
    SC_HIP(hipStreamSynchronize(s)); // the host offsets array is released on return
}

} // namespace

// pairs (a, b) of overlapping boxes, sorted; returns the count and leaves the
// packed pairs in the scratch (device)
int64_t mbr_join(const double *axy, const int *asz, int na, const double *bxy, const int *bsz, int nb,
                 hipStream_t s, const int32_t **dev_pairs) {
    SC_REQUIRE(axy && bxy && asz && bsz && na >= 1 && nb >= 1, SC_EINVAL, "empty layer");
    MbrScratch &m = mscratch();
    layer_boxes(axy, asz, na, m.offa, m.mbra, m.ka, m.ia, m.sxa, m.rs, s);
    layer_boxes(bxy, bsz, nb, m.offb, m.mbrb, m.kb, m.ib, m.sxb, m.rs, s);
    uint32_t *ca = m.ca.ensure(size_t(na) + 1), *cb = m.cb.ensure(size_t(nb) + 1);
    {
        FillBatch fb(s);
        fb.add(ca + na, 1, 0u);
        fb.add(cb + nb, 1, 0u);
        fb.flush();
    }
    k_join<false><<<grid_for(na), BLK, 0, s>>>(m.mbra, na, m.mbrb, m.sxb, m.ib, nb, true, true, ca, nullptr, nullptr);
    k_join<false><<<grid_for(nb), BLK, 0, s>>>(m.mbrb, nb, m.mbra, m.sxa, m.ia, na, false, false, cb, nullptr,
                                               nullptr);
    SC_HIP(hipGetLastError());
    uint32_t *oa = m.oa.ensure(size_t(na) + 1), *ob = m.ob.ensure(size_t(nb) + 1), *tot = m.tot.ensure(2);
    scan_exclusive_u32_2(ca, oa, int64_t(na) + 1, tot, cb, ob, int64_t(nb) + 1, tot + 1, s);
    uint32_t h[2] = {0, 0};
    {
        // the totals: one pageable copy and a stream wait (a stateless call has no context staging)
        SC_HIP(hipMemcpyAsync(h, tot, sizeof(h), hipMemcpyDeviceToHost, s));
        SC_HIP(hipStreamSynchronize(s));
    }
    const int64_t n1 = h[0], n = int64_t(h[0]) + h[1];
    uint64_t *pairs = m.pairs.ensure(size_t(n > 0 ? n : 1));
    m.pairs_alt.ensure(size_t(n > 0 ? n : 1));
    if (h[0]) k_join<true><<<grid_for(na), BLK, 0, s>>>(m.mbra, na, m.mbrb, m.sxb, m.ib, nb, true, true, nullptr, oa,
                                                         pairs);
    if (h[1]) k_join<true><<<grid_for(nb), BLK, 0, s>>>(m.mbrb, nb, m.mbra, m.sxa, m.ia, na, false, false, nullptr, ob,
                                                         pairs + n1);
    SC_HIP(hipGetLastError());
    radix_sort_u64_buf(m.pairs, m.pairs_alt, nullptr, nullptr, n, 0, 64, m.rs, s);
    int32_t *ab = m.ab.ensure(2 * size_t(n > 0 ? n : 1));
    if (n > 0) {
        k_unpack_pairs<<<grid_for(n), BLK, 0, s>>>(m.pairs.p, n, ab);
        SC_HIP(hipGetLastError());
    }
    *dev_pairs = ab;
    return n;
}

} // namespace sc

extern "C" int segclip_mbr_join(const double *a_xy, const int *a_sizes, int na, const double *b_xy,
                                const int *b_sizes, int nb, int32_t *pairs, int64_t cap, int64_t *need,
                                void *stream) {
    try {
        hipStream_t s = static_cast<hipStream_t>(stream);
        const int32_t *dp = nullptr;
        const int64_t n = sc::mbr_join(a_xy, a_sizes, na, b_xy, b_sizes, nb, s, &dp);
        if (need) *need = n;
        if (n > cap) return SEGCLIP_ESPACE;
        if (n > 0) {
            if (!pairs) return SEGCLIP_EINVAL;
            SC_HIP(hipMemcpyAsync(pairs, dp, sizeof(int32_t) * 2 * size_t(n), hipMemcpyDeviceToHost, s));
        }
        SC_HIP(hipStreamSynchronize(s));
        return SEGCLIP_OK;
    } catch (const sc::Error &e) {
        return e.code;
    }
}
