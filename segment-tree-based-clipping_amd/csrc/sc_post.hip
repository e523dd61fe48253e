// sc_post.hip -- per-overlay intersection bookkeeping (rows A11-A15).
//
// Reference (src/clip.cpp / oacc_segment_tree.cpp):
//  - countEdgeIntersections + isNonDegenerateCase + exclusive scans
//    (clip:508-601, 1858-1861): per edge all / non-degenerate counts; a pair
//    is non-degenerate for the base edge if type in {1,3,5,7}, for the
//    overlay edge if type in {1,2,6}
//  - saveIntersectionsParallel (oacc:7257-7341): intersection point by type,
//    alpha/beta quantised with alphaToInt
//  - sortIntersectionsAtEdgesParallel (oacc:7436-7473): per edge, stable sort
//    by the quantised alpha (nine decimal digits, i.e. modulo 1e9)
//  - copyIntersectionsParallel (oacc:7508-7664): augmented rings, source
//    vertex i at i + (non-degenerate intersections before it); a leading
//    degenerate intersection marks the source vertex, the non-degenerate ones
//    follow it in alpha order; neighbour links are 1-based cross indices
//  - calculateInitLabelParallel (oacc:7925-7988): initial Foster labels
// Order: pairs come in (base edge, overlay edge) order and every edge's
// intersections are sorted by quantised alpha (ties keep that order).  On
// inputs where the reference is well defined this gives its result; the
// cases where it is not (an intersection never given a slot) are reported
// through the undefined-behaviour flags instead of being guessed.
#include "sc_ctx.hpp"
#include "sc_fp64.hpp"

namespace sc {

namespace {

constexpr int BLK = 256;
inline unsigned grid_for(int64_t n, int per = BLK) {
    int64_t g = ceil_div(n, per);
    if (g < 1) g = 1;
    return unsigned(g);
}

enum : int { UB_UNASSIGNED = 1, UB_WALK = 2, UB_TRAVERSAL = 4 };

__global__ void k_pair_decode(const uint64_t *__restrict__ pairs, int I, int bits_b, int bits_c, int cStart,
                              const double *__restrict__ bpoly, const double *__restrict__ cpoly, int *pb, int *pc,
                              int *pt, double *px, double *py, int *qa, int *qb, int *all_b, int *nd_b, int *all_c,
                              int *nd_c) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= I) return;
    uint64_t key = pairs[p];
    int b = int((key >> bits_c) & ((1ull << bits_b) - 1ull));
    int cl = int(key & ((1ull << bits_c) - 1ull));
    int bId = 2 * b, cId = 2 * (cStart + cl);
    double b1x = bpoly[bId], b1y = bpoly[bId + 1], b2x = bpoly[bId + 2], b2y = bpoly[bId + 3];
    double c1x = cpoly[cId], c1y = cpoly[cId + 1], c2x = cpoly[cId + 2], c2y = cpoly[cId + 3];
    double al, be;
    int t = classify(b1x, b1y, b2x, b2y, c1x, c1y, c2x, c2y, al, be);
    double x = 0, y = 0;
    if (t == 1) {
        double w = 1.0 - al;
        double ax = w * b1x, ay = w * b1y;
        double bx = al * b2x, by = al * b2y;
        x = ax + bx;
        y = ay + by;
    } else if (t == 3 || t == 4 || t == 5 || t == 7 || t == 8) {
        x = c1x;
        y = c1y;
    } else if (t == 2 || t == 6) {
        x = b1x;
        y = b1y;
    }
    pb[p] = b;
    pc[p] = cl;
    pt[p] = t;
    px[p] = x;
    py[p] = y;
    qa[p] = alpha_to_int(al);
    qb[p] = alpha_to_int(be);
    atomicAdd(&all_b[b], 1);
    atomicAdd(&all_c[cl], 1);
    if (nondeg_b(t)) atomicAdd(&nd_b[b], 1);
    if (nondeg_c(t)) atomicAdd(&nd_c[cl], 1);
}

// keys for the per-edge alpha order: quantised alpha modulo 1e9 (the nine
// decimal digits the reference's LSD sort looks at) in the low 30 bits
__global__ void k_alpha_keys(const int *__restrict__ q, const uint32_t *__restrict__ order, int I, uint64_t *keys,
                             uint32_t *vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= I) return;
    uint32_t p = order ? order[i] : uint32_t(i);
    keys[i] = uint64_t(uint32_t(q[p] % 1000000000));
    vals[i] = p;
}

// second key: the edge (stable on the alpha order from the first pass)
__global__ void k_edge_keys(const int *__restrict__ edge_of, const uint32_t *__restrict__ vals, int I,
                            uint64_t *keys) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= I) return;
    keys[i] = uint64_t(uint32_t(edge_of[vals[i]]));
}

__global__ void k_copy_order(const uint32_t *__restrict__ vals, int I, int *ord) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < I) ord[i] = int(vals[i]);
}

__global__ void k_ckeys(const int *__restrict__ pb, const int *__restrict__ pc, int I, int bits_b, uint64_t *keys,
                        uint32_t *vals) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= I) return;
    keys[p] = (uint64_t(pc[p]) << bits_b) | uint64_t(pb[p]);
    vals[p] = uint32_t(p);
}

__global__ void k_init_aug(int n, double *ax, double *ay, int *av, int *nbr, int *lab) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ax[i] = 0.0;
    ay[i] = 0.0;
    av[i] = -100;
    nbr[i] = -100;
    lab[i] = 0;
}

__global__ void k_aug_sources(int n, const double *__restrict__ poly, const int *__restrict__ nd_scan, double *ax,
                              double *ay) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int pos = v + nd_scan[v];
    ax[pos] = poly[2 * int64_t(v)];
    ay[pos] = poly[2 * int64_t(v) + 1];
}

// one thread per edge with intersections: leading degenerate marks the source
// vertex, then the non-degenerate ones fill the slots after it in alpha order
template <bool SIDE_B>
__global__ void k_aug_segments(const int *__restrict__ seg_edge_of, int I, const int *__restrict__ ord,
                               const int *__restrict__ all_scan, const int *__restrict__ nd_scan,
                               const int *__restrict__ pt, const double *__restrict__ px,
                               const double *__restrict__ py, const int *__restrict__ q, double *ax, double *ay,
                               int *av, int *slot, int *flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= I) return;
    int p0 = ord[i];
    int e = seg_edge_of[p0];
    int s = all_scan[e];
    if (i != s) return; // not the first position of this edge's segment
    int t = all_scan[e + 1];
    int nd = nd_scan[e + 1] - nd_scan[e];
    int src = e + nd_scan[e];
    int j = 0;
    int first = ord[s];
    bool lead_degen = SIDE_B ? degen_b_first(pt[first]) : degen_c_first(pt[first]);
    if (lead_degen) {
        av[src] = q[first];
        slot[first] = src;
        j = 1;
    }
    for (int k = 0; k < nd; ++k) {
        if (s + j + k >= t) {
            atomicOr(flags, UB_UNASSIGNED);
            break;
        }
        int p = ord[s + j + k];
        bool ok = SIDE_B ? nondeg_b(pt[p]) : nondeg_c(pt[p]);
        if (!ok) atomicOr(flags, UB_UNASSIGNED);
        int sl = src + 1 + k;
        ax[sl] = px[p];
        ay[sl] = py[p];
        av[sl] = q[p];
        slot[p] = sl;
    }
    if (j + nd != t - s) atomicOr(flags, UB_UNASSIGNED);
}

__global__ void k_link(const int *__restrict__ sb, const int *__restrict__ scn, int I, int *nbr_b, int *nbr_c,
                       int *flags) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= I) return;
    int a = sb[p], b = scn[p];
    if (a < 0 || b < 0) {
        atomicOr(flags, UB_UNASSIGNED);
        return;
    }
    nbr_b[a] = b + 1;
    nbr_c[b] = a + 1;
}

__device__ inline int rel_pos(int pMN, int pPN, int q, double qx, double qy, double pmx, double pmy, double cx,
                              double cy, double ppx, double ppy) {
    if (pMN != -100 && pMN == q) return 2;
    if (pPN != -100 && pPN == q) return 3;
    double s1 = areaA(qx, qy, pmx, pmy, cx, cy);
    double s2 = areaA(qx, qy, cx, cy, ppx, ppy);
    double s3 = areaA(pmx, pmy, cx, cy, ppx, ppy);
    if (s3 > 0) return (s1 > 0 && s2 > 0) ? 0 : 1;
    return (s1 < 0 && s2 < 0) ? 1 : 0;
}

__global__ void k_init_labels(int nbaug, int ncaug, const double *__restrict__ bx, const double *__restrict__ by,
                              const int *__restrict__ avb, const int *__restrict__ nbrb,
                              const double *__restrict__ cx, const double *__restrict__ cy, int *labb, int *labc,
                              int *flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nbaug) return;
    if (avb[i] == -100) return;
    int tm = i - 1 < 0 ? nbaug - 1 : i - 1;
    int tp = i + 1 == nbaug ? 0 : i + 1;
    int pMN = avb[tm] != -100 ? nbrb[tm] - 1 : -100;
    int pPN = avb[tp] != -100 ? nbrb[tp] - 1 : -100;
    int nId = nbrb[i] - 1;
    if (nId < 0) {
        atomicOr(flags, UB_UNASSIGNED);
        return;
    }
    int qm = nId - 1 < 0 ? ncaug - 1 : nId - 1;
    int qp = nId + 1 == ncaug ? 0 : nId + 1;
    double X = bx[i], Y = by[i];
    int a = rel_pos(pMN, pPN, qm, cx[qm], cy[qm], bx[tm], by[tm], X, Y, bx[tp], by[tp]);
    int b = rel_pos(pMN, pPN, qp, cx[qp], cy[qp], bx[tm], by[tm], X, Y, bx[tp], by[tp]);
    int lab;
    if ((a == 0 && b == 1) || (a == 1 && b == 0)) lab = 1;
    else if ((a == 0 && b == 0) || (a == 1 && b == 1)) lab = 2;
    else if ((b == 3 && a == 1) || (a == 3 && b == 1)) lab = 3;
    else if ((b == 3 && a == 0) || (a == 3 && b == 0)) lab = 4;
    else if ((b == 3 && a == 2) || (a == 3 && b == 2)) lab = 5;
    else if ((a == 2 && b == 1) || (b == 2 && a == 1)) lab = 6;
    else if ((a == 2 && b == 0) || (b == 2 && a == 0)) lab = 7;
    else lab = -102;
    labb[i] = lab;
    labc[nId] = lab;
}

void excl_scan_int(Ctx &c, int *a, int64_t n, int *d_total) {
    scan_exclusive_u32(reinterpret_cast<uint32_t *>(a), reinterpret_cast<uint32_t *>(a), n,
                       reinterpret_cast<uint32_t *>(d_total), c.stream);
}

} // namespace

void post_overlay(Ctx &c, int k) {
    SC_REQUIRE(k >= 1 && k <= c.K, SC_EINVAL, "overlay index out of range");
    SC_REQUIRE(int(c.pair_off.size()) == c.K + 1, SC_ESTATE, "candidates not computed");
    hipStream_t s = c.stream;
    const int nb = c.nb, cs = c.csz[k - 1], cStart = c.coff[k - 1];
    const int64_t pk0 = c.pair_off[k - 1];
    const int I = int(c.pair_off[k] - pk0);
    c.cur_k = k;
    c.I = I;
    const int Ia = I > 0 ? I : 1;
    c.p_b.ensure(Ia); c.p_c.ensure(Ia); c.p_type.ensure(Ia); c.p_x.ensure(Ia); c.p_y.ensure(Ia);
    c.p_qa.ensure(Ia); c.p_qb.ensure(Ia); c.ordB.ensure(Ia); c.ordC.ensure(Ia);
    c.slot_b.ensure(Ia); c.slot_c.ensure(Ia);
    int *all_b = c.all_b.ensure(nb + 1), *nd_b = c.nd_b.ensure(nb + 1);
    int *all_c = c.all_c.ensure(cs + 1), *nd_c = c.nd_c.ensure(cs + 1);
    int *flags = c.flags.ensure(4);
    SC_HIP(hipMemsetAsync(flags, 0, sizeof(int) * 4, s));
    SC_HIP(hipMemsetAsync(all_b, 0, sizeof(int) * (nb + 1), s));
    SC_HIP(hipMemsetAsync(nd_b, 0, sizeof(int) * (nb + 1), s));
    SC_HIP(hipMemsetAsync(all_c, 0, sizeof(int) * (cs + 1), s));
    SC_HIP(hipMemsetAsync(nd_c, 0, sizeof(int) * (cs + 1), s));
    SC_HIP(hipMemsetAsync(c.slot_b.p, 0xff, sizeof(int) * Ia, s));
    SC_HIP(hipMemsetAsync(c.slot_c.p, 0xff, sizeof(int) * Ia, s));
    if (I > 0) {
        k_pair_decode<<<grid_for(I), BLK, 0, s>>>(c.pairs.p + pk0, I, c.bits_b, c.bits_c, cStart, c.bpoly, c.cpoly,
                                                  c.p_b, c.p_c, c.p_type, c.p_x, c.p_y, c.p_qa, c.p_qb, all_b, nd_b,
                                                  all_c, nd_c);
        SC_HIP(hipGetLastError());
    }
    int *tot = reinterpret_cast<int *>(c.d_u32_total.ensure(4));
    excl_scan_int(c, all_b, nb + 1, nullptr);
    excl_scan_int(c, nd_b, nb + 1, tot);
    excl_scan_int(c, all_c, cs + 1, nullptr);
    excl_scan_int(c, nd_c, cs + 1, tot + 1);
    int htot[2];
    SC_HIP(hipMemcpyAsync(htot, tot, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
    SC_HIP(hipStreamSynchronize(s));
    c.nbaug = nb + htot[0];
    c.ncaug = cs + htot[1];

    if (I > 0) {
        // B side: pairs are in (base edge, overlay edge) order; stable sort by
        // alpha, then stably by base edge -> per edge alpha order, ties in pair order
        uint64_t *ak = c.k64a.ensure(Ia);
        uint32_t *av = c.u32a.ensure(Ia);
        k_alpha_keys<<<grid_for(I), BLK, 0, s>>>(c.p_qa, nullptr, I, ak, av);
        SC_HIP(hipGetLastError());
        radix_sort_u64(ak, av, I, 0, 30, c.rs, s);
        k_edge_keys<<<grid_for(I), BLK, 0, s>>>(c.p_b, av, I, ak);
        SC_HIP(hipGetLastError());
        radix_sort_u64(ak, av, I, 0, c.bits_b, c.rs, s);
        k_copy_order<<<grid_for(I), BLK, 0, s>>>(av, I, c.ordB);
        // C side: order by (overlay edge, base edge), then stably by alpha, then by overlay edge
        uint32_t *cv = c.u32b.ensure(Ia);
        k_ckeys<<<grid_for(I), BLK, 0, s>>>(c.p_b, c.p_c, I, c.bits_b, ak, cv);
        SC_HIP(hipGetLastError());
        radix_sort_u64(ak, cv, I, 0, c.bits_b + c.bits_c, c.rs, s);
        k_alpha_keys<<<grid_for(I), BLK, 0, s>>>(c.p_qb, cv, I, ak, av);
        SC_HIP(hipGetLastError());
        radix_sort_u64(ak, av, I, 0, 30, c.rs, s);
        k_edge_keys<<<grid_for(I), BLK, 0, s>>>(c.p_c, av, I, ak);
        SC_HIP(hipGetLastError());
        radix_sort_u64(ak, av, I, 0, c.bits_c, c.rs, s);
        k_copy_order<<<grid_for(I), BLK, 0, s>>>(av, I, c.ordC);
        SC_HIP(hipGetLastError());
    }
    const int nba = c.nbaug, nca = c.ncaug;
    c.aug_bx.ensure(nba); c.aug_by.ensure(nba); c.av_b.ensure(nba); c.nbr_b.ensure(nba); c.lab_b.ensure(nba);
    c.aug_cx.ensure(nca); c.aug_cy.ensure(nca); c.av_c.ensure(nca); c.nbr_c.ensure(nca); c.lab_c.ensure(nca);
    k_init_aug<<<grid_for(nba), BLK, 0, s>>>(nba, c.aug_bx, c.aug_by, c.av_b, c.nbr_b, c.lab_b);
    k_init_aug<<<grid_for(nca), BLK, 0, s>>>(nca, c.aug_cx, c.aug_cy, c.av_c, c.nbr_c, c.lab_c);
    SC_HIP(hipGetLastError());
    k_aug_sources<<<grid_for(nb), BLK, 0, s>>>(nb, c.bpoly, nd_b, c.aug_bx, c.aug_by);
    k_aug_sources<<<grid_for(cs), BLK, 0, s>>>(cs, c.cpoly.p + 2 * int64_t(cStart), nd_c, c.aug_cx, c.aug_cy);
    SC_HIP(hipGetLastError());
    if (I > 0) {
        k_aug_segments<true><<<grid_for(I), BLK, 0, s>>>(c.p_b, I, c.ordB, all_b, nd_b, c.p_type, c.p_x, c.p_y,
                                                         c.p_qa, c.aug_bx, c.aug_by, c.av_b, c.slot_b, flags);
        k_aug_segments<false><<<grid_for(I), BLK, 0, s>>>(c.p_c, I, c.ordC, all_c, nd_c, c.p_type, c.p_x, c.p_y,
                                                          c.p_qb, c.aug_cx, c.aug_cy, c.av_c, c.slot_c, flags);
        SC_HIP(hipGetLastError());
        k_link<<<grid_for(I), BLK, 0, s>>>(c.slot_b, c.slot_c, I, c.nbr_b, c.nbr_c, flags);
        SC_HIP(hipGetLastError());
        k_init_labels<<<grid_for(nba), BLK, 0, s>>>(nba, nca, c.aug_bx, c.aug_by, c.av_b, c.nbr_b, c.aug_cx,
                                                    c.aug_cy, c.lab_b, c.lab_c, flags);
        SC_HIP(hipGetLastError());
    }
    int hflags = 0;
    SC_HIP(hipMemcpyAsync(&hflags, flags, sizeof(int), hipMemcpyDeviceToHost, s));
    SC_HIP(hipStreamSynchronize(s));
    c.undefined = hflags;
}

} // namespace sc
