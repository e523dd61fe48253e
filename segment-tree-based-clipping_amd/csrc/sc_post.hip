// sc_post.hip -- intersection bookkeeping (rows A11-A15) for a batch of jobs.
//
// A job clips one base ring against one overlay ring (overlay k, base
// job_base(k)); all jobs of a batch run in the same launches, their arrays
// laid out job after job.
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
#include "sc_seg.hpp"
#include "sc_fp64.hpp"

namespace sc {

namespace {

constexpr int BLK = 256;
inline unsigned grid_for(int64_t n, int per = BLK) {
    int64_t g = ceil_div(n, per);
    if (g < 1) g = 1;
    return unsigned(g);
}

enum : int { UB_UNASSIGNED = 1, UB_ALPHA_RANGE = 16 };

struct JobArgs {
    int J;
    const int *p0;     // J+1 pair offsets (relative to the batch's first pair)
    const int *bvoff;  // base ring vertex offset in bpoly
    const int *beoff;  // base ring first global edge
    const int *nb, *nc;
    const int *cvoff;  // overlay vertex offset in cpoly
    const int *JB, *JC; // J+1 offsets of the per-edge count arrays (n+1 per job)
};

__global__ void k_pair_decode(JobArgs ja, const uint64_t *__restrict__ pairs, int I, int bits_b, int bits_c,
                              const double *__restrict__ bpoly, const double *__restrict__ cpoly, int *pj, int *pb,
                              int *pc, int *pt, double *px, double *py, int *qa, int *qb, int *all_b, int *nd_b,
                              int *all_c, int *nd_c, int *flags, int *arr_b, int *arr_c) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (WAVE - 1);
    bool counted = false; // a pair inside its job's rings
    int gb = -1, t = 0;
    if (p < I) {
        const int j = seg_of(ja.p0, ja.J, p);
        const uint64_t key = pairs[p];
        const int b = int((key >> bits_c) & ((1ull << bits_b) - 1ull)) - ja.beoff[j];
        const int cl = int(key & ((1ull << bits_c) - 1ull));
        pj[p] = j;
        if (b < 0 || b >= ja.nb[j] || cl >= ja.nc[j]) { // a pair outside the job's rings
            atomicOr(&flags[j], UB_UNASSIGNED);
            arr_c[p] = -1;
            if (arr_b) arr_b[p] = -1;
            pb[p] = 0;
            pc[p] = 0;
            pt[p] = 0;
            px[p] = 0.0;
            py[p] = 0.0;
            qa[p] = 0;
            qb[p] = 0;
        } else {
            const int64_t bId = 2 * int64_t(ja.bvoff[j] + b), cId = 2 * int64_t(ja.cvoff[j] + cl);
            const double b1x = bpoly[bId], b1y = bpoly[bId + 1], b2x = bpoly[bId + 2], b2y = bpoly[bId + 3];
            const double c1x = cpoly[cId], c1y = cpoly[cId + 1], c2x = cpoly[cId + 2], c2y = cpoly[cId + 3];
            double al, be;
            t = classify(b1x, b1y, b2x, b2y, c1x, c1y, c2x, c2y, al, be);
            double x, y;
            intersection_point(t, al, b1x, b1y, b2x, b2y, c1x, c1y, x, y);
            pb[p] = b;
            pc[p] = cl;
            pt[p] = t;
            px[p] = x;
            py[p] = y;
            qa[p] = alpha_to_int(al);
            qb[p] = alpha_to_int(be);
            if (!alpha_in_range(al) || !alpha_in_range(be)) atomicOr(&flags[j], UB_ALPHA_RANGE);
            arr_c[p] = atomicAdd(&all_c[ja.JC[j] + cl], 1); // arrival inside the overlay edge's group
            if (nondeg_c(t)) atomicAdd(&nd_c[ja.JC[j] + cl], 1);
            gb = ja.JB[j] + b;
            counted = true;
        }
    }
    if (!all_b) return; // uniform
    // pairs are sorted by (job, base edge): each run of one base edge inside
    // the wave adds its counts with one atomic per counter
    const int prev = __shfl_up(gb, 1, WAVE);
    const bool head = counted && (lane == 0 || prev != gb);
    const uint64_t hm = __ballot(head), cm = __ballot(counted), ndm = __ballot(counted && nondeg_b(t));
    int base = 0;
    if (head) {
        const uint64_t above = lane == WAVE - 1 ? 0ull : hm & (~0ull << (lane + 1));
        const int nxt = above ? __ffsll((unsigned long long)above) - 1 : WAVE;
        const uint64_t run = (nxt == WAVE ? ~0ull : ((1ull << nxt) - 1ull)) & (~0ull << lane) & cm;
        base = atomicAdd(&all_b[gb], __popcll(run));
        const int n_nd = __popcll(run & ndm);
        if (n_nd) atomicAdd(&nd_b[gb], n_nd);
    }
    // arrival inside the base edge's group: the run's base plus the counted lanes before this one
    const uint64_t upto = lane == WAVE - 1 ? ~0ull : ((2ull << lane) - 1ull);
    const int h = (hm & upto) ? 63 - __clzll(hm & upto) : 0; // this lane's run head
    const int hb = __shfl(base, h, WAVE);
    if (counted) arr_b[p] = hb + __popcll(cm & ((1ull << lane) - 1ull) & ~((1ull << h) - 1ull));
}

// Per-edge alpha order without a global sort.  Every counted pair has an
// arrival index inside its edge's group (taken from the counting atomics), so
// the groups are filled at their exclusive count offsets; inside a group the
// pairs are then ranked by (quantised alpha, pair index), keys that are all
// distinct.  The result equals a stable sort of the pairs (in pair order) by
// (edge slot, quantised alpha): ties keep pair order.
__global__ void k_seg_scatter(int I, const int *__restrict__ edge_of, const int *__restrict__ pj,
                              const int *__restrict__ JX, const int *__restrict__ scan, const int *__restrict__ arr,
                              const int *__restrict__ q, uint64_t *K) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= I) return;
    const int a = arr[p];
    if (a < 0) return; // not counted (flagged job)
    const int pos = scan[JX[pj[p]] + edge_of[p]] + a;
    K[pos] = (uint64_t(uint32_t(q[p] % 1000000000)) << 32) | uint32_t(p);
}

// keys for the per-edge alpha order: the edge's slot over the batch above
// the quantised alpha modulo 1e9 (the nine decimal digits the reference's
// LSD sort looks at, sortIntersectionsAtEdgesParallel oacc:7436-7473) in the
// low 30 bits; values are pair indices
__global__ void k_edge_alpha_keys(const int *__restrict__ edge_of, const int *__restrict__ pj,
                                  const int *__restrict__ off, const int *__restrict__ q, int I, uint64_t *keys,
                                  uint32_t *vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= I) return;
    keys[i] = (uint64_t(uint32_t(off[pj[i]] + edge_of[i])) << 30) | uint64_t(uint32_t(q[i] % 1000000000));
    vals[i] = uint32_t(i);
}

__global__ void k_copy_order(const uint32_t *__restrict__ vals, int I, int *ord) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < I) ord[i] = int(vals[i]);
}

// per job: augmented sizes nb + (non-degenerate intersections on the base), likewise overlay
__global__ void k_job_sizes(JobArgs ja, const int *__restrict__ nd_b, const int *__restrict__ nd_c, int *nba,
                            int *nca) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ja.J) return;
    if (nd_b) nba[j] = ja.nb[j] + (nd_b[ja.JB[j + 1]] - nd_b[ja.JB[j]]);
    nca[j] = ja.nc[j] + (nd_c[ja.JC[j + 1]] - nd_c[ja.JC[j]]);
}

__global__ void k_init_aug(int64_t n, double *ax, double *ay, int *av, int *nbr, int *lab, int *sv, int *sp) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ax[i] = 0.0;
    ay[i] = 0.0;
    av[i] = -100;
    nbr[i] = -100;
    lab[i] = 0;
    if (sv) {
        sv[i] = -1;
        sp[i] = -1;
    }
}

// source vertices: count slot i of job j is vertex i - JX[j] (the last slot of a job is padding)
__global__ void k_aug_sources(int J, const int *__restrict__ JX, const int *__restrict__ nverts,
                              const int *__restrict__ voff, const int *__restrict__ AX,
                              const double *__restrict__ poly, const int *__restrict__ nd_scan, double *ax,
                              double *ay, int *sv) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= JX[J]) return;
    int j = seg_of(JX, J, i);
    int v = int(i - JX[j]);
    if (v >= nverts[j]) return;
    int64_t pos = AX[j] + v + (nd_scan[i] - nd_scan[JX[j]]);
    int64_t g = 2 * int64_t(voff[j] + v);
    ax[pos] = poly[g];
    ay[pos] = poly[g + 1];
    if (sv) sv[pos] = v;
}

// one thread per position of the per-edge alpha order: a leading degenerate
// intersection marks its edge's source vertex, the non-degenerate ones fill
// the slots after it in order
template <bool SIDE_B>
__global__ void k_aug_segments(const int *__restrict__ edge_of, const int *__restrict__ pj, int I,
                               const int *__restrict__ ord, const int *__restrict__ JX, const int *__restrict__ AX,
                               const int *__restrict__ all_scan, const int *__restrict__ nd_scan,
                               const int *__restrict__ pt, const double *__restrict__ px,
                               const double *__restrict__ py, const int *__restrict__ q, double *ax, double *ay,
                               int *av, int *slot, int *flags, int *sp) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= I) return;
    const int p = ord[i];
    const int j = pj[p];
    const int e = edge_of[p];
    const int gi = JX[j] + e;
    const int s = all_scan[gi], t = all_scan[gi + 1];
    const int nd = nd_scan[gi + 1] - nd_scan[gi];
    const int src = AX[j] + e + (nd_scan[gi] - nd_scan[JX[j]]);
    const int first = i == s ? p : ord[s];
    const int k0 = (SIDE_B ? degen_b_first(pt[first]) : degen_c_first(pt[first])) ? 1 : 0;
    if (i == s && k0 + nd != t - s) atomicOr(&flags[j], UB_UNASSIGNED);
    if (i == s && k0) {
        av[src] = q[p];
        slot[p] = src;
        return;
    }
    const int k = i - s - k0;
    if (k >= nd) return; // more positions than non-degenerate slots: flagged above
    if (!(SIDE_B ? nondeg_b(pt[p]) : nondeg_c(pt[p]))) atomicOr(&flags[j], UB_UNASSIGNED);
    const int sl = src + 1 + k;
    ax[sl] = px[p];
    ay[sl] = py[p];
    av[sl] = q[p];
    slot[p] = sl;
    if (sp) sp[sl] = p;
}

// neighbour links, 1-based and local to the job's rings (as the reference's arrays)
__global__ void k_link(const int *__restrict__ sb, const int *__restrict__ scn, const int *__restrict__ pj, int I,
                       const int *__restrict__ AB, const int *__restrict__ AC, int *nbr_b, int *nbr_c, int *flags) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= I) return;
    int a = sb[p], b = scn[p], j = pj[p];
    if (a < 0 || b < 0) {
        atomicOr(&flags[j], UB_UNASSIGNED);
        return;
    }
    nbr_b[a] = b - AC[j] + 1;
    nbr_c[b] = a - AB[j] + 1;
}

__global__ void k_init_labels(int J, const int *__restrict__ AB, const int *__restrict__ AC,
                              const double *__restrict__ bx, const double *__restrict__ by,
                              const int *__restrict__ avb, const int *__restrict__ nbrb,
                              const double *__restrict__ cx, const double *__restrict__ cy, int *labb, int *labc,
                              int *flags) {
    int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= AB[J]) return;
    if (avb[g] == -100) return;
    int j = seg_of(AB, J, g);
    const int B0 = AB[j], nbaug = AB[j + 1] - B0, C0 = AC[j], ncaug = AC[j + 1] - C0;
    const int i = int(g - B0);
    int tm = i - 1 < 0 ? nbaug - 1 : i - 1;
    int tp = i + 1 == nbaug ? 0 : i + 1;
    int pMN = avb[B0 + tm] != -100 ? nbrb[B0 + tm] - 1 : -100;
    int pPN = avb[B0 + tp] != -100 ? nbrb[B0 + tp] - 1 : -100;
    int nId = nbrb[g] - 1;
    if (nId < 0 || nId >= ncaug) {
        atomicOr(&flags[j], UB_UNASSIGNED);
        return;
    }
    int qm = nId - 1 < 0 ? ncaug - 1 : nId - 1;
    int qp = nId + 1 == ncaug ? 0 : nId + 1;
    int lab = label_of(pMN, pPN, qm, qp, cx[C0 + qm], cy[C0 + qm], cx[C0 + qp], cy[C0 + qp], bx[B0 + tm],
                       by[B0 + tm], bx[g], by[g], bx[B0 + tp], by[B0 + tp]);
    labb[g] = lab;
    labc[C0 + nId] = lab;
}

// ---- sparse base ring ----

// per position: leading-degenerate / non-degenerate class (B side)
__global__ void k_sp_mark(int I, const int *__restrict__ ord, const int *__restrict__ pj, const int *__restrict__ pb,
                          const int *__restrict__ pt, const int *__restrict__ p0, int *kind, uint32_t *ndf,
                          int *flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= I) return;
    int p = ord[i], j = pj[p], e = pb[p], t = pt[p];
    bool start = i == p0[j] || pb[ord[i - 1]] != e;
    bool nd = nondeg_b(t), lead = start && degen_b_first(t);
    if (!nd && !lead) atomicOr(&flags[j], UB_UNASSIGNED);
    kind[i] = lead ? 1 : (nd ? 2 : 0);
    ndf[i] = nd ? 1u : 0u;
}

__global__ void k_sp_fill(int I, const int *__restrict__ ord, const int *__restrict__ pj, const int *__restrict__ pb,
                          const int *__restrict__ p0, const uint32_t *__restrict__ ndg, const int *__restrict__ bvoff,
                          const double *__restrict__ bpoly, const double *__restrict__ px,
                          const double *__restrict__ py, const int *__restrict__ qa, int *slot, int *ndx, int *edge,
                          int *av, int *nbr, int *lab, double *x, double *y, int *pos, const int *kind) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= I) return;
    int p = ord[i], j = pj[p], e = pb[p];
    int rel = int(ndg[i] - ndg[p0[j]]);
    int k = kind[i];
    slot[i] = k == 1 ? e + rel : (k == 2 ? e + 1 + rel : -1);
    ndx[i] = rel;
    edge[i] = e;
    av[i] = qa[p];
    nbr[i] = -100;
    lab[i] = 0;
    if (k == 2) {
        x[i] = px[p];
        y[i] = py[p];
    } else {
        const double *V = bpoly + 2 * (int64_t(bvoff[j]) + e);
        x[i] = V[0];
        y[i] = V[1];
    }
    pos[p] = i;
}

__global__ void k_sp_sizes(int J, const int *__restrict__ p0, const uint32_t *__restrict__ ndg,
                           const int *__restrict__ nb, int *nba) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= J) return;
    nba[j] = nb[j] + int(ndg[p0[j + 1]] - ndg[p0[j]]);
}

__global__ void k_sp_link(int I, const int *__restrict__ pos, const int *__restrict__ pj,
                          const int *__restrict__ slot, const int *__restrict__ scn, const int *__restrict__ AC,
                          int nca, int *nbr_p, int *nbr_c, int *flags) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= I) return;
    int i = pos[p], j = pj[p];
    int a = slot[i], b = scn[p];
    if (a < 0 || b < 0 || b >= nca) {
        atomicOr(&flags[j], UB_UNASSIGNED);
        return;
    }
    nbr_p[i] = b - AC[j] + 1;
    nbr_c[b] = a + 1;
}

__global__ void k_sp_init_labels(int I, int J, SparseP sp, const int *__restrict__ av, const int *__restrict__ nbr,
                                 const int *__restrict__ AC, const double *__restrict__ cx,
                                 const double *__restrict__ cy, int *lab, int *labc, int *flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= I) return;
    const int q = sp.slot[i];
    if (q < 0) return;
    const int j = seg_of(sp.P0, J, i);
    const int nbaug = sp.AB[j + 1] - sp.AB[j], C0 = AC[j], ncaug = AC[j + 1] - C0;
    const int tm = q - 1 < 0 ? nbaug - 1 : q - 1;
    const int tp = q + 1 == nbaug ? 0 : q + 1;
    const SlotInfo pm = sp_slot(sp, j, tm), pp = sp_slot(sp, j, tp);
    const int pMN = pm.pos >= 0 ? nbr[pm.pos] - 1 : -100;
    const int pPN = pp.pos >= 0 ? nbr[pp.pos] - 1 : -100;
    const int nId = nbr[i] - 1;
    if (nId < 0 || nId >= ncaug) {
        atomicOr(&flags[j], UB_UNASSIGNED);
        return;
    }
    const int qm = nId - 1 < 0 ? ncaug - 1 : nId - 1;
    const int qp = nId + 1 == ncaug ? 0 : nId + 1;
    const int l = label_of(pMN, pPN, qm, qp, cx[C0 + qm], cy[C0 + qm], cx[C0 + qp], cy[C0 + qp], pm.x, pm.y,
                           sp.x[i], sp.y[i], pp.x, pp.y);
    lab[i] = l;
    labc[C0 + nId] = l;
}

int bits_for(int64_t n) {
    int b = 0;
    while ((int64_t(1) << b) < n) b++;
    return b < 1 ? 1 : b;
}

} // namespace

void post_jobs(Ctx &c, int k0, int k1) {
    SC_REQUIRE(k0 >= 1 && k1 <= c.K && k0 <= k1, SC_EINVAL, "overlay index out of range");
    SC_REQUIRE(int(c.pair_off.size()) == c.K + 1, SC_ESTATE, "candidates not computed");
    hipStream_t s = c.stream;
    const int J = k1 - k0 + 1;
    JobSet &js = c.jobs;
    js.k0 = k0;
    js.J = J;
    js.p0.assign(J + 1, 0);
    js.bvoff.resize(J);
    js.beoff.resize(J);
    js.nb.resize(J);
    js.nc.resize(J);
    js.cvoff.resize(J);
    js.JB.assign(J + 1, 0);
    js.JC.assign(J + 1, 0);
    js.cin.resize(J + 1);
    const int64_t P0 = c.pair_off[k0 - 1];
    for (int j = 0; j < J; ++j) {
        int k = k0 + j, br = job_base(c, k);
        js.p0[j + 1] = int(c.pair_off[k] - P0);
        js.bvoff[j] = c.bvoff[br];
        js.beoff[j] = c.beoff[br];
        js.nb[j] = c.bsz[br];
        js.nc[j] = c.csz[k - 1];
        js.cvoff[j] = c.coff[k - 1];
        js.cin[j] = c.eoff[k - 1] - c.nb;
        js.JB[j + 1] = js.JB[j] + js.nb[j] + 1;
        js.JC[j + 1] = js.JC[j] + js.nc[j] + 1;
    }
    const int I = js.p0[J];
    c.cur_k = J == 1 ? k0 : 0;
    c.I = I;
    js.cin[J] = c.eoff[k1] - c.nb;
    {
        // the job arrays back to back, one copy to the device
        const std::vector<int> *parts[] = {&js.p0, &js.bvoff, &js.beoff, &js.nb, &js.nc,
                                           &js.cvoff, &js.JB, &js.JC, &js.cin};
        int **views[] = {&js.d_p0, &js.d_bvoff, &js.d_beoff, &js.d_nb, &js.d_nc,
                         &js.d_cvoff, &js.d_JB, &js.d_JC, &js.d_cin};
        size_t tot = 0;
        for (const std::vector<int> *v : parts) tot += v->size();
        js.h_pack.resize(tot);
        int *dp = js.d_pack.ensure(tot);
        size_t at = 0;
        for (int q = 0; q < 9; ++q) {
            std::copy(parts[q]->begin(), parts[q]->end(), js.h_pack.begin() + at);
            *views[q] = dp + at;
            at += parts[q]->size();
        }
        SC_HIP(hipMemcpyAsync(dp, js.h_pack.data(), sizeof(int) * tot, hipMemcpyHostToDevice, s));
    }
    JobArgs ja;
    ja.J = J;
    ja.p0 = js.d_p0;
    ja.bvoff = js.d_bvoff;
    ja.beoff = js.d_beoff;
    ja.nb = js.d_nb;
    ja.nc = js.d_nc;
    ja.cvoff = js.d_cvoff;
    ja.JB = js.d_JB;
    ja.JC = js.d_JC;
    const int nJB = js.JB[J], nJC = js.JC[J];
    const int Ia = I > 0 ? I : 1;
    // one-to-many batches keep the shared base ring sparse (no dense copy per job)
    const bool sparse = !c.paired && !c.prov && J >= 2;
    c.sparseP = sparse;
    c.p_j.ensure(Ia); c.p_b.ensure(Ia); c.p_c.ensure(Ia); c.p_type.ensure(Ia); c.p_x.ensure(Ia); c.p_y.ensure(Ia);
    c.p_qa.ensure(Ia); c.p_qb.ensure(Ia); c.ordB.ensure(Ia); c.ordC.ensure(Ia);
    c.slot_b.ensure(Ia); c.slot_c.ensure(Ia);
    FillBatch fb(s);
    int *all_b = nullptr, *nd_b = nullptr;
    if (!sparse) {
        all_b = c.all_b.ensure(nJB + 1);
        nd_b = c.nd_b.ensure(nJB + 1);
        fb.add(all_b, nJB + 1, 0u);
        fb.add(nd_b, nJB + 1, 0u);
    }
    int *all_c = c.all_c.ensure(nJC + 1), *nd_c = c.nd_c.ensure(nJC + 1);
    int *flags = c.flags.ensure(J);
    fb.add(flags, J, 0u);
    fb.add(all_c, nJC + 1, 0u);
    fb.add(nd_c, nJC + 1, 0u);
    fb.add(c.slot_b.p, Ia, ~0u);
    fb.add(c.slot_c.p, Ia, ~0u);
    fb.flush();
    if (I > 0) {
        k_pair_decode<<<grid_for(I), BLK, 0, s>>>(ja, c.pairs.p + P0, I, c.bits_b, c.bits_c, c.bpoly, c.cpoly, c.p_j,
                                                  c.p_b, c.p_c, c.p_type, c.p_x, c.p_y, c.p_qa, c.p_qb, all_b, nd_b,
                                                  all_c, nd_c, flags, sparse ? nullptr : c.arr_b.ensure(Ia),
                                                  c.arr_c.ensure(Ia));
        SC_HIP(hipGetLastError());
    }
    auto scan = [&](int *a, int64_t n) {
        scan_exclusive_u32(reinterpret_cast<uint32_t *>(a), reinterpret_cast<uint32_t *>(a), n, nullptr, s);
    };
    if (!sparse) {
        scan(all_b, nJB + 1);
        scan(nd_b, nJB + 1);
    }
    scan(all_c, nJC + 1);
    scan(nd_c, nJC + 1);

    // Per edge alpha order on each side (positions = the edge groups' count
    // offsets): by ranking inside the groups, or (a group beyond SEG_LDS_MAX,
    // or a sparse base side) by one stable radix sort of the pairs (in pair
    // order) by (edge slot, quantised alpha).  Pairs are in (job, base edge,
    // overlay edge) order, so ties keep the base side's pair order and, on the
    // overlay side, base-edge order.
    const int bits_eb = bits_for(int64_t(nJB) + 1), bits_ec = bits_for(int64_t(nJC) + 1);
    SC_REQUIRE(bits_eb + 30 <= 64 && bits_ec + 30 <= 64, SC_ERANGE, "edge order key too wide");
    auto radix_order = [&](bool side_b) {
        const int *edge_of = side_b ? c.p_b.p : c.p_c.p;
        const int *JX = side_b ? ja.JB : ja.JC;
        k_edge_alpha_keys<<<grid_for(I), BLK, 0, s>>>(edge_of, c.p_j, JX, side_b ? c.p_qa : c.p_qb, I,
                                                      c.ord_key.ensure(Ia), c.ord_val.ensure(Ia));
        SC_HIP(hipGetLastError());
        radix_sort_u64_buf(c.ord_key, c.ord_key_alt, &c.ord_val, &c.ord_val_alt, I, 0,
                           (side_b ? bits_eb : bits_ec) + 30, c.rs, s);
        k_copy_order<<<grid_for(I), BLK, 0, s>>>(c.ord_val, I, side_b ? c.ordB : c.ordC);
        SC_HIP(hipGetLastError());
    };
    uint32_t *segctl = c.seg_ctl.ensure(2 * SEG_CTL); // per side: list counts, too-large flag
    if (I > 0) {
        {
            FillBatch fb(s);
            fb.add(segctl, 2 * SEG_CTL, 0u);
            fb.flush();
        }
        auto seg_order = [&](const int *edge_of, const int *JX, const int *scan, const int *arr, const int *q,
                             int nseg, DBuf<uint64_t> &K, DBuf<uint32_t> &wl, DBuf<uint32_t> &bl, uint32_t *ctl,
                             int *ord) {
            k_seg_scatter<<<grid_for(I), BLK, 0, s>>>(I, edge_of, c.p_j, JX, scan, arr, q, K.ensure(Ia));
            // listed groups hold more than SEG_THREAD (wave list) or SEG_WAVE_MAX (block list) pairs each
            const int64_t nw = std::min<int64_t>(nseg, int64_t(I) / (SEG_THREAD + 1)) + 1;
            const int64_t nb = std::min<int64_t>(nseg, int64_t(I) / (SEG_WAVE_MAX + 1)) + 1;
            seg_sort<false>(nseg, scan, K, ord, wl.ensure(size_t(nw)), bl.ensure(size_t(nb)), ctl, s);
        };
        if (sparse) radix_order(true);
        else seg_order(c.p_b, ja.JB, all_b, c.arr_b, c.p_qa, nJB, c.seg_kb, c.seg_wb, c.seg_bb, segctl, c.ordB);
        seg_order(c.p_c, ja.JC, all_c, c.arr_c, c.p_qb, nJC, c.seg_kc, c.seg_wc, c.seg_bc, segctl + SEG_CTL, c.ordC);
    }
    // augmented ring sizes and offsets per job
    int *nba = js.d_AB.ensure(J + 1), *nca = js.d_AC.ensure(J + 1);
    if (sparse) {
        // virtual slot of every base-side intersection
        int *kind = reinterpret_cast<int *>(c.u32c.ensure(Ia));
        uint32_t *ndf = reinterpret_cast<uint32_t *>(c.sp_ndf.ensure(size_t(I) + 1));
        SC_HIP(hipMemsetAsync(ndf + I, 0, sizeof(uint32_t), s));
        if (I > 0) {
            k_sp_mark<<<grid_for(I), BLK, 0, s>>>(I, c.ordB, c.p_j, c.p_b, c.p_type, ja.p0, kind, ndf, flags);
            SC_HIP(hipGetLastError());
        }
        scan_exclusive_u32(ndf, ndf, int64_t(I) + 1, nullptr, s);
        int *slot = c.sp_slot.ensure(Ia), *ndx = c.sp_ndx.ensure(Ia), *edge = c.sp_edge.ensure(Ia);
        int *sav = c.sp_av.ensure(Ia), *snbr = c.sp_nbr.ensure(Ia), *pos = c.sp_pos.ensure(Ia);
        double *sx = c.sp_x.ensure(Ia), *sy = c.sp_y.ensure(Ia);
        c.sp_lab.ensure(Ia);
        if (I > 0) {
            k_sp_fill<<<grid_for(I), BLK, 0, s>>>(I, c.ordB, c.p_j, c.p_b, ja.p0, ndf, ja.bvoff, c.bpoly, c.p_x,
                                                  c.p_y, c.p_qa, slot, ndx, edge, sav, snbr, c.sp_lab, sx, sy, pos,
                                                  kind);
            SC_HIP(hipGetLastError());
        }
        k_sp_sizes<<<grid_for(J), BLK, 0, s>>>(J, ja.p0, ndf, ja.nb, nba);
        SC_HIP(hipGetLastError());
    }
    k_job_sizes<<<grid_for(J), BLK, 0, s>>>(ja, nd_b, nd_c, nba, nca);
    SC_HIP(hipGetLastError());
    {
        FillBatch fb(s);
        fb.add(nba + J, 1, 0u);
        fb.add(nca + J, 1, 0u);
        fb.flush();
    }
    uint32_t *tot = c.d_u32_total.ensure(2);
    scan_exclusive_u32_2(reinterpret_cast<uint32_t *>(nba), reinterpret_cast<uint32_t *>(nba), J + 1, tot,
                         reinterpret_cast<uint32_t *>(nca), reinterpret_cast<uint32_t *>(nca), J + 1, tot + 1, s);
    js.AB.resize(J + 1);
    js.AC.resize(J + 1);
    uint32_t hseg[2 * SEG_CTL] = {};
    {
        ReadBack rb(c);
        rb.add(js.AB.data(), static_cast<const int *>(nba), size_t(J) + 1);
        rb.add(js.AC.data(), static_cast<const int *>(nca), size_t(J) + 1);
        if (I > 0) rb.add(hseg, static_cast<const uint32_t *>(segctl), size_t(2 * SEG_CTL));
        rb.run();
    }
    // a group too large for LDS: that side's order by the radix sort instead
    if (hseg[2] && !sparse) radix_order(true);
    if (hseg[SEG_CTL + 2]) radix_order(false);
    const int64_t nba_tot = js.AB[J], nca_tot = js.AC[J];
    c.nbaug = int(nba_tot);
    c.ncaug = int(nca_tot);

    const int64_t nbaA = nba_tot > 0 ? nba_tot : 1, ncaA = nca_tot > 0 ? nca_tot : 1;
    c.aug_cx.ensure(ncaA); c.aug_cy.ensure(ncaA); c.av_c.ensure(ncaA); c.nbr_c.ensure(ncaA); c.lab_c.ensure(ncaA);
    int *svb = nullptr, *spb = nullptr, *svc = nullptr, *spc = nullptr;
    if (c.prov) {
        svb = c.sv_b.ensure(nbaA);
        spb = c.sp_b.ensure(nbaA);
        svc = c.sv_c.ensure(ncaA);
        spc = c.sp_c.ensure(ncaA);
    }
    if (!sparse) {
        c.aug_bx.ensure(nbaA); c.aug_by.ensure(nbaA); c.av_b.ensure(nbaA); c.nbr_b.ensure(nbaA); c.lab_b.ensure(nbaA);
        k_init_aug<<<grid_for(nba_tot), BLK, 0, s>>>(nba_tot, c.aug_bx, c.aug_by, c.av_b, c.nbr_b, c.lab_b, svb, spb);
        k_aug_sources<<<grid_for(nJB), BLK, 0, s>>>(J, ja.JB, ja.nb, ja.bvoff, nba, c.bpoly, nd_b, c.aug_bx,
                                                    c.aug_by, svb);
    }
    k_init_aug<<<grid_for(nca_tot), BLK, 0, s>>>(nca_tot, c.aug_cx, c.aug_cy, c.av_c, c.nbr_c, c.lab_c, svc, spc);
    k_aug_sources<<<grid_for(nJC), BLK, 0, s>>>(J, ja.JC, ja.nc, ja.cvoff, nca, c.cpoly, nd_c, c.aug_cx, c.aug_cy,
                                                svc);
    SC_HIP(hipGetLastError());
    if (I > 0) {
        if (!sparse)
            k_aug_segments<true><<<grid_for(I), BLK, 0, s>>>(c.p_b, c.p_j, I, c.ordB, ja.JB, nba, all_b, nd_b,
                                                             c.p_type, c.p_x, c.p_y, c.p_qa, c.aug_bx, c.aug_by,
                                                             c.av_b, c.slot_b, flags, spb);
        k_aug_segments<false><<<grid_for(I), BLK, 0, s>>>(c.p_c, c.p_j, I, c.ordC, ja.JC, nca, all_c, nd_c,
                                                          c.p_type, c.p_x, c.p_y, c.p_qb, c.aug_cx, c.aug_cy, c.av_c,
                                                          c.slot_c, flags, spc);
        SC_HIP(hipGetLastError());
        if (sparse) {
            k_sp_link<<<grid_for(I), BLK, 0, s>>>(I, c.sp_pos, c.p_j, c.sp_slot, c.slot_c, nca, int(nca_tot),
                                                  c.sp_nbr, c.nbr_c, flags);
            SC_HIP(hipGetLastError());
            k_sp_init_labels<<<grid_for(I), BLK, 0, s>>>(I, J, sparse_view(c), c.sp_av, c.sp_nbr, nca, c.aug_cx,
                                                         c.aug_cy, c.sp_lab, c.lab_c, flags);
            SC_HIP(hipGetLastError());
        } else {
            k_link<<<grid_for(I), BLK, 0, s>>>(c.slot_b, c.slot_c, c.p_j, I, nba, nca, c.nbr_b, c.nbr_c, flags);
            SC_HIP(hipGetLastError());
            k_init_labels<<<grid_for(nba_tot), BLK, 0, s>>>(J, nba, nca, c.aug_bx, c.aug_by, c.av_b, c.nbr_b,
                                                            c.aug_cx, c.aug_cy, c.lab_b, c.lab_c, flags);
            SC_HIP(hipGetLastError());
        }
    }
    js.flags.assign(size_t(J), 0);
    c.flags_pending = true; // read back later, with the next read-back that needs a wait anyway
}

void fetch_flags(Ctx &c) {
    if (!c.flags_pending) return;
    JobSet &js = c.jobs;
    ReadBack rb(c);
    rb.add(js.flags.data(), static_cast<const int *>(c.flags.p), size_t(js.J));
    rb.run();
    c.flags_pending = false;
    int any = 0;
    for (int f : js.flags) any |= f;
    c.undefined = any;
}

void post_overlay(Ctx &c, int k) { post_jobs(c, k, k); }

SparseP sparse_view(Ctx &c) {
    SparseP v;
    v.P0 = c.jobs.d_p0;
    v.AB = c.jobs.d_AB;
    v.nbv = c.jobs.d_nb;
    v.bvoff = c.jobs.d_bvoff;
    v.slot = c.sp_slot;
    v.ndx = c.sp_ndx;
    v.edge = c.sp_edge;
    v.x = c.sp_x;
    v.y = c.sp_y;
    v.bpoly = c.bpoly;
    return v;
}

void init_labels_one(Ctx &c, int nba, int nca, const double *bx, const double *by, const int *avb, const int *nbrb,
                     const double *cx, const double *cy, int *labb, int *labc, int *flags) {
    hipStream_t s = c.stream;
    std::vector<int> off = {0, nba, 0, nca};
    thread_local DBuf<int> *dd = new DBuf<int>(); // scratch, kept for the thread's lifetime
    int *d = dd->ensure(4);
    SC_HIP(hipMemcpyAsync(d, off.data(), sizeof(int) * 4, hipMemcpyHostToDevice, s));
    {
        FillBatch fb(s);
        fb.add(labb, nba, 0u);
        fb.add(labc, nca, 0u);
        fb.flush();
    }
    k_init_labels<<<grid_for(nba), BLK, 0, s>>>(1, d, d + 2, bx, by, avb, nbrb, cx, cy, labb, labc, flags);
    SC_HIP(hipGetLastError());
    SC_HIP(hipStreamSynchronize(s)); // off is a stack vector
}

} // namespace sc
