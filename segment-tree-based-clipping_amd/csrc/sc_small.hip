// sc_small.hip -- a batch of small independent pairs clipped by one persistent
// kernel: a device work queue of pairs, one wave per pair at a time, the
// whole clip of a pair in LDS (rows A8-A17 for rings of up to 64 vertices).
//
// A pair's candidate set is the brute-force set {(b, c): LSMF and a non-zero
// lineSegmentIntersection2 type}, which equals the set the reference's tree
// emits (saveIntersectionsIdsMulticore, oacc:6010-6244; the oracle's pair set
// is tested equal to brute force), in (base edge, overlay edge) order.  Then,
// as the staged path and the oracle do (clip:522-601, oacc:7257-7988):
// per-edge counts, intersection points and quantised alpha / beta, every
// edge's intersections ordered by (alpha modulo 1e9, pair), augmented rings
// with neighbour links, initial labels.  Labelling, traversal and clean-up
// follow labelIntersections2 / createResult2 / cleanUpResult2
// (foster/polyclip.h:337-787) on index-linked vertex records in LDS.
//
// The wave works in parallel where the result does not depend on order:
// pair tests (a lane per base edge), per-edge ordering (a lane per edge),
// slots, labels, point-in-polygon winding sums (integer sums over edges) and
// the entry/exit status (the parity of the crossings before a vertex in ring
// order).  Traversal and clean-up, whose cost is proportional to the output,
// run on one lane.
//
// Covered here: pairs whose intersection vertices are all labelled CROSSING
// or BOUNCING, whose rings keep a plain vertex, with no bouncing pair to
// split, every intersection given a slot, quantised alphas in range, and
// sizes within the LDS capacities below.  Any other pair sets its fallback
// word and the host reruns the whole batch on the staged path, so a batch's
// results never depend on which path ran.
#include "sc_ctx.hpp"
#include "sc_fp64.hpp"

namespace sc {

namespace {

constexpr int SP_NV = 64;                    // ring vertices per side
constexpr int SP_I = 128;                    // intersections per pair
constexpr int SP_AUG = SP_NV + SP_I;         // augmented ring slots per side
constexpr int SP_RES = 2 * SP_AUG;           // output ring vertex records
constexpr int SP_POOL = 2 * SP_AUG + SP_RES; // P slots, Q slots, output records
constexpr int SP_OUTR = 64;                  // output rings per pair
constexpr int SP_OUTV = SP_RES;              // output vertices per pair
constexpr int SP_NOV = -100;

enum : int { L_NONE = 0, L_CROSSING = 1, L_BOUNCING = 2 };
enum : int { EN_EXIT = 0, EN_ENTRY = 1, EN_NEITHER = 2 };
// fallback reasons, per pair
enum : int {
    F_SIZE = 1, F_PAIRS = 2, F_SLOT = 4, F_ALPHA = 8, F_WALK = 16, F_LABEL = 32, F_PLAIN = 64, F_SPLIT = 128,
    F_TRAV = 256, F_OUT = 512
};

struct SmallLDS {
    double bx[SP_NV + 1], by[SP_NV + 1], cx[SP_NV + 1], cy[SP_NV + 1]; // rings with the closing vertex
    double px[SP_I], py[SP_I];                                          // intersection points
    int qa[SP_I], qb[SP_I];                                             // quantised alpha / beta
    uint8_t pb[SP_I], pc[SP_I], pt[SP_I];                               // base edge, overlay edge, type
    uint8_t sB[SP_I], sC[SP_I];            // per side: position in the edge-group order -> pair
    int16_t niB[SP_I], niC[SP_I];          // per pair: its slot on each augmented ring
    int16_t cntB[SP_NV + 1], ndB[SP_NV + 1], cntC[SP_NV + 1], ndC[SP_NV + 1]; // exclusive scans
    int avB[SP_AUG], avC[SP_AUG];          // quantised alpha of a slot, -100 for a plain vertex
    int16_t nbB[SP_AUG], nbC[SP_AUG];      // 1-based partner slot, -100 for none
    // vertex records: P slots [0, nbAug), Q slots [SP_AUG, SP_AUG + ncAug),
    // output ring records from 2 * SP_AUG on
    double vx[SP_POOL], vy[SP_POOL];
    int16_t vprev[SP_POOL], vnext[SP_POOL], vnbr[SP_POOL];
    int8_t vlab[SP_POOL];
    uint8_t vfl[SP_POOL];                  // bit 0: intersection; bits 1-2: entry/exit status
    int16_t roots[SP_OUTR];
    int16_t ord[SP_OUTV];                  // output vertices in emission order (record indices)
    int16_t rlen[SP_OUTR];                 // output ring lengths
};

struct SmallArgs {
    int J;
    const int *bvoff, *nb, *cvoff, *nc; // per job: ring offsets (vertices, closing vertices included) and sizes
    const double *bpoly, *cpoly;
    uint32_t *queue;                    // next job to take (zero before the launch)
    int *fail;                          // per job: fallback reasons (0: clipped here)
    int *route;                         // per job: 0 crossing traversal, 1 no crossing vertex
    int *job_nv, *job_nr;               // per job: output vertices and rings
    double *out_x, *out_y;              // per job slot of SP_OUTV vertices
    int *out_len;                       // per job slot of SP_OUTR ring lengths
    double *out4;                       // optional: areas and IoU, 4 x J (the batch-area layout)
    uint32_t *anyfail;                  // optional: set when some job falls back
};

__device__ inline bool is_inter(const SmallLDS &L, int v) { return L.vfl[v] & 1; }
__device__ inline void set_inter(SmallLDS &L, int v, bool on) { L.vfl[v] = uint8_t((L.vfl[v] & ~1u) | (on ? 1u : 0u)); }
__device__ inline int enex_of(const SmallLDS &L, int v) { return (L.vfl[v] >> 1) & 3; }
__device__ inline void set_enex(SmallLDS &L, int v, int e) { L.vfl[v] = uint8_t((L.vfl[v] & 1u) | (unsigned(e) << 1)); }
__device__ inline int toggle(int s) { return s == EN_ENTRY ? EN_EXIT : (s == EN_EXIT ? EN_ENTRY : s); }

__device__ inline int wave_sum(int v) {
    for (int d = WAVE / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, WAVE);
    return v;
}
__device__ inline int wave_excl(int v, int lane, int &total) {
    int inc = v;
    for (int d = 1; d < WAVE; d <<= 1) {
        const int y = __shfl_up(inc, d, WAVE);
        if (lane >= d) inc += y;
    }
    total = __shfl(inc, WAVE - 1, WAVE);
    return inc - v;
}

// winding test of (rx, ry) against the ring of n records starting at 'base'
// in slot order (polygon::pointInPoly, polygon.h:435-457): an integer sum of
// per-edge terms, so the lanes may add them in any order
__device__ bool point_in_ring(const SmallLDS &L, int base, int n, double rx, double ry, int lane) {
    int w = 0;
    for (int k = lane; k < n; k += WAVE) {
        const int one = base + k, two = base + (k + 1 == n ? 0 : k + 1);
        const double p0x = L.vx[one], p0y = L.vy[one], p1x = L.vx[two], p1y = L.vy[two];
        if ((p0y < ry) != (p1y < ry)) {
            const int step = 2 * (p1y > p0y) - 1;
            if (p0x >= rx) {
                if (p1x > rx) w += step;
                else if ((areaA(p0x, p0y, p1x, p1y, rx, ry) > 0) == (p1y > p0y)) w += step;
            } else if (p1x > rx) {
                if ((areaA(p0x, p0y, p1x, p1y, rx, ry) > 0) == (p1y > p0y)) w += step;
            }
        }
    }
    return (wave_sum(w) % 2) != 0;
}

// first record without the intersection flag in ring order from the root, -1 if none
__device__ int first_plain(const SmallLDS &L, int base, int n, int lane) {
    int best = 0x7fffffff;
    for (int k = lane; k < n; k += WAVE)
        if (!is_inter(L, base + k) && k < best) best = k;
    for (int d = WAVE / 2; d > 0; d >>= 1) best = min(best, __shfl_xor(best, d, WAVE));
    return best == 0x7fffffff ? -1 : best;
}

// polygon::removeVertex (polygon.h:421-429) on an output ring with root *root
__device__ void ring_remove(SmallLDS &L, int &root, int V) {
    if (root == V) {
        root = L.vnext[V];
        if (L.vnext[root] == root) root = -1;
    }
    L.vnext[L.vprev[V]] = L.vnext[V];
    L.vprev[L.vnext[V]] = L.vprev[V];
}

__device__ inline double A3(const SmallLDS &L, int a, int b, int c) {
    return areaA(L.vx[a], L.vy[a], L.vx[b], L.vy[b], L.vx[c], L.vy[c]);
}

// Shoelace sums in the order of the batch-area kernels, so both give the same
// bits: a ring's terms x_i y_(i+1) - x_(i+1) y_i added per lane over
// i = lane, lane + 64, ... and then by a shuffle tree (32, 16, ..., 1); an input
// ring (at most 64 vertices) and the sum of a job's |S| over its rings (at most
// 64) as those kernels' 256-thread block sums, whose two upper tree steps add
// zeros.
__device__ inline double wave_tree(double s) {
    for (int d = WAVE / 2; d > 0; d >>= 1) s += __shfl_down(s, d, WAVE);
    return __shfl(s, 0, WAVE);
}
__device__ double input_ring_sum2(const double *x, const double *y, int n, int lane) {
    double s = 0.0;
    if (lane < n) s += x[lane] * y[lane + 1] - x[lane + 1] * y[lane];
    s = (s + 0.0) + 0.0;
    return wave_tree(s);
}
__device__ void pair_areas(const SmallArgs &a, const SmallLDS &L, int j, int nb, int nc, int nr, int lane) {
    double mine = 0.0; // lane r: |S| of output ring r
    int at = 0;
    for (int r = 0; r < nr; ++r) {
        const int len = L.rlen[r];
        double s = 0.0;
        for (int i = lane; i < len; i += WAVE) {
            const int v = L.ord[at + i], w = L.ord[at + (i + 1 == len ? 0 : i + 1)];
            s += L.vx[v] * L.vy[w] - L.vx[w] * L.vy[v];
        }
        const double rs = wave_tree(s);
        if (lane == r) mine += fabs(rs);
        at += len;
    }
    const double out2 = wave_tree((mine + 0.0) + 0.0);
    const double P2 = input_ring_sum2(L.bx, L.by, nb, lane), Q2 = input_ring_sum2(L.cx, L.cy, nc, lane);
    if (lane == 0) {
        const int J = a.J;
        const double I = 0.5 * out2, P = 0.5 * fabs(P2), Q = 0.5 * fabs(Q2);
        const double U = P + Q - I;
        a.out4[j] = I;
        a.out4[J + j] = P;
        a.out4[2 * J + j] = Q;
        a.out4[3 * J + j] = U != 0.0 ? I / U : 0.0;
    }
}

// one pair, the whole wave; returns the fallback reasons (0: clipped)
__device__ int clip_pair(const SmallArgs &a, SmallLDS &L, int j, int lane, int &route, int &nv_out, int &nr_out) {
    const int nb = a.nb[j], nc = a.nc[j];
    nv_out = nr_out = 0;
    route = 0;
    if (nb > SP_NV || nc > SP_NV || nb < 3 || nc < 3) return F_SIZE;
    {
        const double *B = a.bpoly + 2 * int64_t(a.bvoff[j]);
        const double *C = a.cpoly + 2 * int64_t(a.cvoff[j]);
        for (int i = lane; i <= nb; i += WAVE) {
            L.bx[i] = B[2 * i];
            L.by[i] = B[2 * i + 1];
        }
        for (int i = lane; i <= nc; i += WAVE) {
            L.cx[i] = C[2 * i];
            L.cy[i] = C[2 * i + 1];
        }
    }
    __builtin_amdgcn_wave_barrier();

    // ---- candidate pairs: lane = base edge, overlay edges in order ----
    uint64_t hits = 0;
    int nd_here = 0;
    if (lane < nb) {
        const double b1x = L.bx[lane], b1y = L.by[lane], b2x = L.bx[lane + 1], b2y = L.by[lane + 1];
        for (int c = 0; c < nc; ++c) {
            const double c1x = L.cx[c], c1y = L.cy[c], c2x = L.cx[c + 1], c2y = L.cy[c + 1];
            if (lsmf(b1x, b1y, b2x, b2y, c1x, c1y, c2x, c2y) &&
                classify_pre(b1x, b1y, b2x, b2y, c1x, c1y, c2x, c2y) != 0)
                hits |= 1ull << c;
        }
    }
    int I = 0;
    const int first = wave_excl(__popcll(hits), lane, I);
    if (I > SP_I) return F_PAIRS;
    bool alpha_bad = false;
    if (lane < nb) {
        const double b1x = L.bx[lane], b1y = L.by[lane], b2x = L.bx[lane + 1], b2y = L.by[lane + 1];
        int p = first;
        for (uint64_t m = hits; m; m &= m - 1, ++p) {
            const int c = __ffsll((unsigned long long)m) - 1;
            const double c1x = L.cx[c], c1y = L.cy[c];
            double al, be;
            const int t = classify(b1x, b1y, b2x, b2y, c1x, c1y, L.cx[c + 1], L.cy[c + 1], al, be);
            double x, y;
            intersection_point(t, al, b1x, b1y, b2x, b2y, c1x, c1y, x, y);
            L.pb[p] = uint8_t(lane);
            L.pc[p] = uint8_t(c);
            L.pt[p] = uint8_t(t);
            L.px[p] = x;
            L.py[p] = y;
            L.qa[p] = alpha_to_int(al);
            L.qb[p] = alpha_to_int(be);
            alpha_bad |= !alpha_in_range(al) || !alpha_in_range(be);
            nd_here += nondeg_b(t) ? 1 : 0;
        }
    }
    if (__ballot(alpha_bad)) return F_ALPHA;
    // per base edge: all / non-degenerate counts, exclusive (countEdgeIntersections, clip:522-601)
    {
        int tot;
        const int ex = wave_excl(nd_here, lane, tot);
        if (lane < nb) {
            L.cntB[lane] = int16_t(first);
            L.ndB[lane] = int16_t(ex);
        }
        if (lane == 0) {
            L.cntB[nb] = int16_t(I);
            L.ndB[nb] = int16_t(tot);
        }
    }
    __builtin_amdgcn_wave_barrier();
    // per overlay edge: its pairs in pair order, sorted by (beta mod 1e9, pair)
    int cn = 0, cnd = 0;
    if (lane < nc)
        for (int p = 0; p < I; ++p)
            if (L.pc[p] == lane) {
                ++cn;
                cnd += nondeg_c(L.pt[p]) ? 1 : 0;
            }
    {
        int tot, totd;
        const int ex = wave_excl(cn, lane, tot), exd = wave_excl(cnd, lane, totd);
        if (lane < nc) {
            L.cntC[lane] = int16_t(ex);
            L.ndC[lane] = int16_t(exd);
            int k = ex;
            for (int p = 0; p < I; ++p)
                if (L.pc[p] == lane) {
                    // insertion by (beta mod 1e9, pair): pairs arrive in pair order
                    const uint32_t key = uint32_t(L.qb[p] % 1000000000);
                    int q = k++;
                    while (q > ex && uint32_t(L.qb[L.sC[q - 1]] % 1000000000) > key) {
                        L.sC[q] = L.sC[q - 1];
                        --q;
                    }
                    L.sC[q] = uint8_t(p);
                }
        }
        if (lane == 0) {
            L.cntC[nc] = int16_t(tot);
            L.ndC[nc] = int16_t(totd);
        }
    }
    // per base edge: its pairs (contiguous in pair order) sorted by (alpha mod 1e9, pair)
    if (lane < nb) {
        const int g0 = L.cntB[lane], g1 = L.cntB[lane + 1];
        for (int p = g0; p < g1; ++p) {
            const uint32_t key = uint32_t(L.qa[p] % 1000000000);
            int q = p;
            while (q > g0 && uint32_t(L.qa[L.sB[q - 1]] % 1000000000) > key) {
                L.sB[q] = L.sB[q - 1];
                --q;
            }
            L.sB[q] = uint8_t(p);
        }
    }
    for (int p = lane; p < I; p += WAVE) L.niB[p] = L.niC[p] = -1;
    __builtin_amdgcn_wave_barrier();
    const int nbAug = nb + L.ndB[nb], ncAug = nc + L.ndC[nc];

    // ---- augmented rings (copyIntersectionsParallel, oacc:7508-7664): a lane per edge ----
    if (lane < nb) {
        const int g0 = L.cntB[lane], g1 = L.cntB[lane + 1], s0 = lane + L.ndB[lane], nd = L.ndB[lane + 1] - L.ndB[lane];
        L.vx[s0] = L.bx[lane];
        L.vy[s0] = L.by[lane];
        L.avB[s0] = SP_NOV;
        int k = g0;
        if (g0 < g1 && degen_b_first(L.pt[L.sB[g0]])) {
            const int p = L.sB[g0];
            L.avB[s0] = L.qa[p];
            L.niB[p] = int16_t(s0);
            ++k;
        }
        for (int i = 0; i < nd && k + i < g1; ++i) {
            const int p = L.sB[k + i], sl = s0 + 1 + i;
            L.vx[sl] = L.px[p];
            L.vy[sl] = L.py[p];
            L.avB[sl] = L.qa[p];
            L.niB[p] = int16_t(sl);
        }
    }
    if (lane < nc) {
        const int g0 = L.cntC[lane], g1 = L.cntC[lane + 1], s0 = lane + L.ndC[lane], nd = L.ndC[lane + 1] - L.ndC[lane];
        L.vx[SP_AUG + s0] = L.cx[lane];
        L.vy[SP_AUG + s0] = L.cy[lane];
        L.avC[s0] = SP_NOV;
        int k = g0;
        if (g0 < g1 && degen_c_first(L.pt[L.sC[g0]])) {
            const int p = L.sC[g0];
            L.avC[s0] = L.qb[p];
            L.niC[p] = int16_t(s0);
            ++k;
        }
        for (int i = 0; i < nd && k + i < g1; ++i) {
            const int p = L.sC[k + i], sl = s0 + 1 + i;
            L.vx[SP_AUG + sl] = L.px[p];
            L.vy[SP_AUG + sl] = L.py[p];
            L.avC[sl] = L.qb[p];
            L.niC[p] = int16_t(sl);
        }
    }
    for (int s = lane; s < nbAug; s += WAVE) L.nbB[s] = SP_NOV;
    for (int s = lane; s < ncAug; s += WAVE) {
        L.nbC[s] = SP_NOV;
        L.vlab[SP_AUG + s] = L_NONE;
    }
    __builtin_amdgcn_wave_barrier();
    // every pair holds a slot on both rings, else the reference leaves it out (undefined)
    bool bad = false;
    for (int p = lane; p < I; p += WAVE) {
        const int a0 = L.niB[p], c0 = L.niC[p];
        if (a0 < 0 || c0 < 0) bad = true;
        else {
            L.nbB[a0] = int16_t(c0 + 1);
            L.nbC[c0] = int16_t(a0 + 1);
        }
    }
    if (__ballot(bad)) return F_SLOT;
    __builtin_amdgcn_wave_barrier();

    // ---- initial labels (calculateInitLabelParallel, oacc:7925-7988), a lane per slot ----
    for (int i = lane; i < nbAug; i += WAVE) {
        int lab = L_NONE;
        if (L.avB[i] != SP_NOV && L.nbB[i] != SP_NOV) { // a slot without partner falls back below
            const int tm = i == 0 ? nbAug - 1 : i - 1, tp = i + 1 == nbAug ? 0 : i + 1;
            const int pMN = L.avB[tm] != SP_NOV ? L.nbB[tm] - 1 : SP_NOV;
            const int pPN = L.avB[tp] != SP_NOV ? L.nbB[tp] - 1 : SP_NOV;
            const int nId = L.nbB[i] - 1;
            const int qm = nId == 0 ? ncAug - 1 : nId - 1, qp = nId + 1 == ncAug ? 0 : nId + 1;
            lab = label_of(pMN, pPN, qm, qp, L.vx[SP_AUG + qm], L.vy[SP_AUG + qm], L.vx[SP_AUG + qp],
                           L.vy[SP_AUG + qp], L.vx[tm], L.vy[tm], L.vx[i], L.vy[i], L.vx[tp], L.vy[tp]);
            L.vlab[SP_AUG + nId] = int8_t(lab);
        }
        L.vlab[i] = int8_t(lab);
    }
    __builtin_amdgcn_wave_barrier();

    // ---- vertex records in ring order (copyToLinkedList, clip:606-695) ----
    // The reference matches the augmented arrays to the rings by coordinates;
    // that walk keeps slot order unless an inserted intersection equals the
    // next ring vertex's coordinates, which falls back.
    bool walk = false, unlinked = false, odd_label = false;
    for (int s = lane; s < nbAug; s += WAVE) {
        L.vprev[s] = int16_t(s == 0 ? nbAug - 1 : s - 1);
        L.vnext[s] = int16_t(s + 1 == nbAug ? 0 : s + 1);
        const bool in = L.avB[s] != SP_NOV;
        L.vfl[s] = uint8_t(in ? 1 : 0) | uint8_t(EN_NEITHER << 1);
        L.vnbr[s] = int16_t(L.nbB[s] != SP_NOV ? SP_AUG + L.nbB[s] - 1 : -1);
        if (in && L.nbB[s] == SP_NOV) unlinked = true;
        if (in && L.vlab[s] != L_CROSSING && L.vlab[s] != L_BOUNCING) odd_label = true;
    }
    for (int s = lane; s < ncAug; s += WAVE) {
        const int v = SP_AUG + s;
        L.vprev[v] = int16_t(SP_AUG + (s == 0 ? ncAug - 1 : s - 1));
        L.vnext[v] = int16_t(SP_AUG + (s + 1 == ncAug ? 0 : s + 1));
        const bool in = L.nbC[s] != SP_NOV;
        L.vfl[v] = uint8_t(in ? 1 : 0) | uint8_t(EN_NEITHER << 1);
        L.vnbr[v] = int16_t(in ? L.nbC[s] - 1 : -1);
        if (in && L.vlab[v] != L_CROSSING && L.vlab[v] != L_BOUNCING) odd_label = true;
    }
    // inserted intersections of edge e must differ from ring vertex e+1
    if (lane < nb) {
        const int s0 = lane + L.ndB[lane], s1 = lane + 1 + L.ndB[lane + 1];
        const int nx = lane + 1 == nb ? 0 : lane + 1;
        for (int s = s0 + 1; s < s1; ++s)
            if (L.vx[s] == L.bx[nx] && L.vy[s] == L.by[nx]) walk = true;
    }
    if (lane < nc) {
        const int s0 = lane + L.ndC[lane], s1 = lane + 1 + L.ndC[lane + 1];
        const int nx = lane + 1 == nc ? 0 : lane + 1;
        for (int s = s0 + 1; s < s1; ++s)
            if (L.vx[SP_AUG + s] == L.cx[nx] && L.vy[SP_AUG + s] == L.cy[nx]) walk = true;
    }
    if (__ballot(walk)) return F_WALK;
    if (__ballot(unlinked)) return F_SLOT;
    // labels other than CROSSING / BOUNCING (chains, ON labels, no rule) take the staged path
    if (__ballot(odd_label)) return F_LABEL;
    __builtin_amdgcn_wave_barrier();

    // ---- labelIntersections2: crossings, special cases, entry/exit ----
    bool cross_any = false;
    for (int s = lane; s < nbAug; s += WAVE)
        if (is_inter(L, s) && L.vlab[s] == L_CROSSING) cross_any = true;
    const bool crossP = __ballot(cross_any) != 0;
    cross_any = false;
    for (int s = lane; s < ncAug; s += WAVE)
        if (is_inter(L, SP_AUG + s) && L.vlab[SP_AUG + s] == L_CROSSING) cross_any = true;
    const bool crossQ = __ballot(cross_any) != 0;
    int nroots = 0;
    // 3.5) a ring without crossing vertices is an output ring when inside the other
    for (int side = 0; side < 2; ++side) {
        const bool has = side == 0 ? crossP : crossQ;
        if (has) continue;
        const int base = side == 0 ? 0 : SP_AUG, n = side == 0 ? nbAug : ncAug;
        const int obase = side == 0 ? SP_AUG : 0, on = side == 0 ? ncAug : nbAug;
        const int v = first_plain(L, base, n, lane);
        if (v < 0) return F_PLAIN;
        if (point_in_ring(L, obase, on, L.vx[base + v], L.vy[base + v], lane)) {
            if (lane == 0) L.roots[nroots] = int16_t(base);
            ++nroots;
        }
    }
    // 4) entry/exit: from the first plain vertex V, the k-th crossing after it
    // gets ENTRY/EXIT alternately starting from ENTRY (EXIT when V lies inside
    // the other ring); a bouncing vertex met in the EXIT state is a split candidate
    bool splitP[3] = {false, false, false};
    for (int side = 0; side < 2; ++side) {
        const bool has = side == 0 ? crossP : crossQ;
        if (!has) continue;
        const int base = side == 0 ? 0 : SP_AUG, n = side == 0 ? nbAug : ncAug;
        const int obase = side == 0 ? SP_AUG : 0, on = side == 0 ? ncAug : nbAug;
        const int v = first_plain(L, base, n, lane);
        if (v < 0) return F_PLAIN;
        int st0 = EN_ENTRY;
        if (point_in_ring(L, obase, on, L.vx[base + v], L.vy[base + v], lane)) st0 = EN_EXIT;
        int before = 0; // crossings in walk order before this chunk
        for (int w0 = 1; w0 < n; w0 += WAVE) {
            const int w = w0 + lane;
            const int r = w < n ? base + (v + w) % n : -1;
            const bool cr = r >= 0 && is_inter(L, r) && L.vlab[r] == L_CROSSING;
            const uint64_t m = __ballot(cr);
            const int k = before + __popcll(m & ((1ull << lane) - 1ull));
            const int st = (k & 1) ? toggle(st0) : st0;
            if (cr) set_enex(L, r, st);
            if (r >= 0 && is_inter(L, r) && L.vlab[r] == L_BOUNCING && st == EN_EXIT) {
                if (side == 0) L.vfl[r] |= 8u; // split candidate on P
                else L.vfl[r] |= 16u;          // split candidate on Q
            }
            before += __popcll(m);
        }
        splitP[side] = true;
    }
    __builtin_amdgcn_wave_barrier();
    // 5) a P candidate whose partner is a Q candidate is split: staged path
    bool split = false;
    if (splitP[0] && splitP[1])
        for (int s = lane; s < nbAug; s += WAVE)
            if ((L.vfl[s] & 8u) && L.vnbr[s] >= 0 && (L.vfl[L.vnbr[s]] & 16u)) split = true;
    if (__ballot(split)) return F_SPLIT;
    for (int s = lane; s < nbAug; s += WAVE) L.vfl[s] &= 7u;
    for (int s = lane; s < ncAug; s += WAVE) L.vfl[SP_AUG + s] &= 7u;
    __builtin_amdgcn_wave_barrier();
    route = (crossP || crossQ) ? 0 : 1;

    // ---- createResult2 and cleanUpResult2, one lane ----
    int fail = 0, nv = 0, nr = 0;
    if (lane == 0) {
        int top = 2 * SP_AUG; // next free output record
        const int root = 0;
        int I0 = root;
        bool have = is_inter(L, I0) && L.vlab[I0] == L_CROSSING;
        if (!have) {
            do { I0 = L.vnext[I0]; } while (!(is_inter(L, I0) && L.vlab[I0] == L_CROSSING) && I0 != root);
            have = I0 != root;
        }
        const int bound = 4 * (nbAug + ncAug) + 16;
        while (have && !fail) {
            int rroot = -1;
            int V = I0;
            set_inter(L, V, false);
            int steps = 0;
            do {
                const int st = toggle(enex_of(L, V));
                do {
                    V = st == EN_EXIT ? L.vnext[V] : L.vprev[V];
                    set_inter(L, V, false);
                    if (top >= SP_POOL || ++steps > bound) {
                        fail = top >= SP_POOL ? F_OUT : F_TRAV;
                        break;
                    }
                    const int t = top++;
                    L.vx[t] = L.vx[V];
                    L.vy[t] = L.vy[V];
                    if (rroot < 0) {
                        L.vnext[t] = L.vprev[t] = int16_t(t);
                        rroot = t;
                    } else {
                        const int last = L.vprev[rroot];
                        L.vprev[t] = int16_t(last);
                        L.vnext[t] = int16_t(rroot);
                        L.vnext[last] = int16_t(t);
                        L.vprev[rroot] = int16_t(t);
                    }
                } while (!(enex_of(L, V) == st) && V != I0);
                if (fail) break;
                if (V != I0) {
                    V = L.vnbr[V];
                    if (V < 0) {
                        fail = F_TRAV;
                        break;
                    }
                    set_inter(L, V, false);
                }
            } while (V != I0);
            if (fail) break;
            if (nroots >= SP_OUTR) {
                fail = F_OUT;
                break;
            }
            L.roots[nroots++] = int16_t(rroot);
            do { I0 = L.vnext[I0]; } while (!(is_inter(L, I0) && L.vlab[I0] == L_CROSSING) && I0 != root);
            have = I0 != root;
        }
        // clean-up: drop vertices with |A(prev, v, next)| < eps, root first
        for (int r = 0; r < nroots && !fail; ++r) {
            int R = L.roots[r];
            while (R >= 0 && fabs(A3(L, L.vprev[R], R, L.vnext[R])) < EPS) ring_remove(L, R, R);
            if (R >= 0) {
                const int it_root = R;
                int V = R;
                do {
                    const int nx = L.vnext[V];
                    if (fabs(A3(L, L.vprev[V], V, L.vnext[V])) < EPS) ring_remove(L, R, V);
                    V = nx;
                } while (V != it_root);
            }
            L.roots[r] = int16_t(R);
        }
        // emission into this job's output slot
        if (!fail) {
            double *ox = a.out_x + int64_t(j) * SP_OUTV, *oy = a.out_y + int64_t(j) * SP_OUTV;
            int *olen = a.out_len + int64_t(j) * SP_OUTR;
            for (int r = 0; r < nroots && !fail; ++r) {
                const int R = L.roots[r];
                if (R < 0) continue;
                int V = R, len = 0;
                do {
                    if (nv >= SP_OUTV) {
                        fail = F_OUT;
                        break;
                    }
                    ox[nv] = L.vx[V];
                    oy[nv] = L.vy[V];
                    L.ord[nv] = int16_t(V);
                    ++nv;
                    ++len;
                    V = L.vnext[V];
                } while (V != R);
                L.rlen[nr] = int16_t(len);
                olen[nr++] = len;
            }
        }
    }
    fail = __shfl(fail, 0, WAVE);
    nv_out = __shfl(nv, 0, WAVE);
    nr_out = __shfl(nr, 0, WAVE);
    __builtin_amdgcn_wave_barrier();
    if (!fail && a.out4) pair_areas(a, L, j, nb, nc, nr_out, lane);
    return fail;
}

// The persistent kernel: every wave takes the next pair from the queue until
// none is left (so waves of short pairs take more of them).
__global__ __launch_bounds__(WAVE) void k_small_pairs(SmallArgs a) {
    __shared__ SmallLDS L;
    const int lane = threadIdx.x;
    for (;;) {
        uint32_t j = 0;
        if (lane == 0) j = atomicAdd(a.queue, 1u);
        j = uint32_t(__builtin_amdgcn_readfirstlane(int(j)));
        if (j >= uint32_t(a.J)) break;
        int route = 0, nv = 0, nr = 0;
        const int f = clip_pair(a, L, int(j), lane, route, nv, nr);
        if (lane == 0) {
            a.fail[j] = f;
            a.route[j] = route;
            a.job_nv[j] = f ? 0 : nv;
            a.job_nr[j] = f ? 0 : nr;
            if (f && a.anyfail) atomicOr(a.anyfail, 1u);
        }
        __syncthreads(); // LDS of this pair read before the next pair's writes
    }
}

// output rings of every job at their final offsets: one block per job
__global__ void k_small_emit(int J, const int *__restrict__ voff, const int *__restrict__ roff,
                             const int *__restrict__ job_nr, const double *__restrict__ ox,
                             const double *__restrict__ oy, const int *__restrict__ olen, double *rx, double *ry,
                             int *ring_off) {
    const int j = blockIdx.x;
    const int v0 = voff[j], nv = voff[j + 1] - v0, r0 = roff[j], nr = job_nr[j];
    for (int i = threadIdx.x; i < nv; i += blockDim.x) {
        rx[v0 + i] = ox[int64_t(j) * SP_OUTV + i];
        ry[v0 + i] = oy[int64_t(j) * SP_OUTV + i];
    }
    if (threadIdx.x == 0) {
        int at = v0;
        for (int r = 0; r < nr; ++r) {
            ring_off[r0 + r] = at;
            at += olen[int64_t(j) * SP_OUTR + r];
        }
    }
}

// the totals after the exclusive scans of the per-job counts: ring_off[NR] = NV
__global__ void k_small_tail(int J, const int *__restrict__ voff, const int *__restrict__ roff, int *ring_off) {
    if (blockIdx.x == 0 && threadIdx.x == 0) ring_off[roff[J]] = voff[J];
}

// rings given back to back without closing vertices (the set_pairs layout)
// into the resident layout with the closing vertex; one thread per vertex
__global__ void k_pairs_layout(const double *__restrict__ xy, int64_t n, const int *__restrict__ in_off,
                               const int *__restrict__ out_off, int R, double *out) {
    const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int r = seg_of(in_off, R, g);
    const int local = int(g - in_off[r]);
    const double x = xy[2 * g], y = xy[2 * g + 1];
    const int64_t d = int64_t(out_off[r]) + local;
    out[2 * d] = x;
    out[2 * d + 1] = y;
    if (local == 0) {
        const int64_t cl = int64_t(out_off[r]) + (in_off[r + 1] - in_off[r]);
        out[2 * cl] = x;
        out[2 * cl + 1] = y;
    }
}

int small_grid(const Ctx &c, int J) {
    int cus = 256, v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, c.device) == hipSuccess && v > 0) cus = v;
    // LDS bounds the waves per CU; a few more blocks than fit keep the CUs fed
    return std::min(J, cus * 8);
}

struct SmallScratch {
    DBuf<int> jobs, fail, route, nv, nr, voff, roff, olen;
    DBuf<double> ox, oy;
    DBuf<uint32_t> queue, tot;
};

SmallScratch &sscratch() {
    thread_local SmallScratch *p = new SmallScratch();
    return *p;
}

} // namespace

bool small_batch_eligible(const Ctx &c, int k0, int k1) {
    if (!c.paired || c.prov || !c.small_fused || c.uni) return false;
    for (int k = k0; k <= k1; ++k)
        if (c.bsz[k - 1] > SP_NV || c.csz[k - 1] > SP_NV) return false;
    return true;
}

bool clip_small_batch(Ctx &c, int k0, int k1) {
    hipStream_t s = c.stream;
    const int J = k1 - k0 + 1;
    SmallScratch &S = sscratch();
    std::vector<int> h(4 * size_t(J));
    for (int j = 0; j < J; ++j) {
        const int k = k0 + j, br = job_base(c, k);
        h[j] = c.bvoff[br];
        h[J + j] = c.bsz[br];
        h[2 * J + j] = c.coff[k - 1];
        h[3 * J + j] = c.csz[k - 1];
    }
    int *dj = S.jobs.ensure(4 * size_t(J));
    SC_HIP(hipMemcpyAsync(dj, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, s));
    SmallArgs a;
    a.J = J;
    a.bvoff = dj;
    a.nb = dj + J;
    a.cvoff = dj + 2 * J;
    a.nc = dj + 3 * J;
    a.bpoly = c.bpoly;
    a.cpoly = c.cpoly;
    a.queue = S.queue.ensure(1);
    a.fail = S.fail.ensure(J);
    a.route = S.route.ensure(J);
    a.job_nv = S.nv.ensure(size_t(J) + 1);
    a.job_nr = S.nr.ensure(size_t(J) + 1);
    a.out_x = S.ox.ensure(size_t(J) * SP_OUTV);
    a.out_y = S.oy.ensure(size_t(J) * SP_OUTV);
    a.out_len = S.olen.ensure(size_t(J) * SP_OUTR);
    {
        FillBatch fb(s);
        fb.add(a.queue, 1, 0u);
        fb.add(a.job_nv + J, 1, 0u);
        fb.add(a.job_nr + J, 1, 0u);
        fb.flush();
    }
    a.out4 = nullptr;
    a.anyfail = nullptr;
    k_small_pairs<<<small_grid(c, J), WAVE, 0, s>>>(a);
    SC_HIP(hipGetLastError());
    int *voff = S.voff.ensure(size_t(J) + 1), *roff = S.roff.ensure(size_t(J) + 1);
    scan_exclusive_u32_2(reinterpret_cast<const uint32_t *>(a.job_nv), reinterpret_cast<uint32_t *>(voff), J + 1,
                         nullptr, reinterpret_cast<const uint32_t *>(a.job_nr), reinterpret_cast<uint32_t *>(roff),
                         J + 1, nullptr, s);
    // one read-back: ring offsets per job, vertex total, fallback reasons, routes
    JobSet &js = c.jobs;
    js.ring_off.assign(size_t(J) + 1, 0);
    std::vector<int> fails(static_cast<size_t>(J)), routes(static_cast<size_t>(J));
    int NV = 0;
    {
        ReadBack rb(c);
        rb.add(js.ring_off.data(), static_cast<const int *>(roff), size_t(J) + 1);
        rb.add(&NV, static_cast<const int *>(voff + J), 1);
        rb.add(fails.data(), static_cast<const int *>(a.fail), size_t(J));
        rb.add(routes.data(), static_cast<const int *>(a.route), size_t(J));
        rb.run();
    }
    int any = 0;
    for (int f : fails) any |= f;
    c.small_fail = any;
    if (any) return false; // the staged path reruns the batch
    const int NR = js.ring_off[J];
    c.d_ring_off.ensure(size_t(NR) + 1);
    c.d_ring_x.ensure(size_t(NV > 0 ? NV : 1));
    c.d_ring_y.ensure(size_t(NV > 0 ? NV : 1));
    k_small_emit<<<J, 256, 0, s>>>(J, voff, roff, a.job_nr, a.out_x, a.out_y, a.out_len, c.d_ring_x, c.d_ring_y,
                                   c.d_ring_off);
    k_small_tail<<<1, 1, 0, s>>>(J, voff, roff, c.d_ring_off);
    SC_HIP(hipGetLastError());
    c.nrings = NR;
    c.n_out = NV;
    c.undefined = 0;
    c.cur_k = J == 1 ? k0 : 0;
    c.flags_pending = false;
    js.k0 = k0;
    js.J = J;
    js.flags.assign(size_t(J), 0);
    js.route = routes;
    js.reason.assign(size_t(J), 0);
    return true;
}

void pairs_graph_reset(Ctx &c) {
    Ctx::PairsGraph &g = c.pg;
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g.exec = nullptr;
    g.graph = nullptr;
    g.bxy = g.cxy = nullptr;
    g.out4 = nullptr;
}

void pairs_area_graph(Ctx &c, const double *bxy, const double *cxy, double *out4) {
    SC_REQUIRE(c.paired && c.K >= 1, SC_ESTATE, "pairs not set (segclip_set_pairs)");
    SC_REQUIRE(!c.uni, SC_ESTATE, "the area / IoU graph is defined for the intersection");
    SC_REQUIRE(bxy && cxy && out4, SC_EINVAL, "null pointer");
    for (int k = 0; k < c.K; ++k)
        SC_REQUIRE(c.bsz[k] <= SP_NV && c.csz[k] <= SP_NV, SC_EINVAL,
                   "the graph forward takes rings of at most 64 vertices");
    hipStream_t s = c.stream;
    const int J = c.K;
    Ctx::PairsGraph &g = c.pg;
    if (!g.exec || g.bxy != bxy || g.cxy != cxy || g.out4 != out4) {
        pairs_graph_reset(c);
        // tables: job (bvoff, nb, cvoff, nc), base input / output offsets (J+1, J), overlay ones (J+1, J)
        std::vector<int> h(4 * size_t(J) + 2 * (2 * size_t(J) + 1));
        int *t = h.data();
        for (int j = 0; j < J; ++j) {
            t[j] = c.bvoff[j];
            t[J + j] = c.bsz[j];
            t[2 * J + j] = c.coff[j];
            t[3 * J + j] = c.csz[j];
        }
        int *bin = t + 4 * J, *bout = bin + J + 1, *cin = bout + J, *cout = cin + J + 1;
        for (int j = 0; j <= J; ++j) {
            bin[j] = c.beoff[j];
            cin[j] = c.eoff[j] - c.nb;
            if (j < J) {
                bout[j] = c.bvoff[j];
                cout[j] = c.coff[j];
            }
        }
        int *dt = g.tables.ensure(h.size());
        SC_HIP(hipMemcpyAsync(dt, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, s));
        uint32_t *words = g.words.ensure(2);
        SmallArgs a;
        a.J = J;
        a.bvoff = dt;
        a.nb = dt + J;
        a.cvoff = dt + 2 * J;
        a.nc = dt + 3 * J;
        a.bpoly = c.bpoly;
        a.cpoly = c.cpoly;
        a.queue = words;
        a.anyfail = words + 1;
        a.fail = g.fail.ensure(J);
        a.route = g.route.ensure(J);
        a.job_nv = g.nv.ensure(size_t(J) + 1);
        a.job_nr = g.nr.ensure(size_t(J) + 1);
        a.out_x = g.ox.ensure(size_t(J) * SP_OUTV);
        a.out_y = g.oy.ensure(size_t(J) * SP_OUTV);
        a.out_len = g.olen.ensure(size_t(J) * SP_OUTR);
        a.out4 = out4;
        SC_HIP(hipStreamSynchronize(s)); // the tables are on the device before the capture
        const int64_t nbt = c.nb, nct = c.nct;
        const int dbin = int(bin - t), dbout = int(bout - t), dcin = int(cin - t), dcout = int(cout - t);
        SC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        hipGraph_t graph = nullptr;
        try {
            // fill kernels, no memset nodes, inside the graph
            FillBatch fb(s);
            fb.add(words, 2, 0u);
            fb.flush();
            k_pairs_layout<<<unsigned(ceil_div(nbt, 256)), 256, 0, s>>>(bxy, nbt, dt + dbin, dt + dbout, J,
                                                                        c.bpoly);
            k_pairs_layout<<<unsigned(ceil_div(nct, 256)), 256, 0, s>>>(cxy, nct, dt + dcin, dt + dcout, J,
                                                                        c.cpoly);
            k_small_pairs<<<small_grid(c, J), WAVE, 0, s>>>(a);
            SC_HIP(hipGetLastError());
        } catch (...) {
            (void)hipStreamEndCapture(s, &graph); // the stream leaves capture mode either way
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        SC_HIP(hipStreamEndCapture(s, &graph));
        g.graph = graph;
        SC_HIP(hipGraphInstantiateWithFlags(&g.exec, g.graph, 0));
        g.bxy = bxy;
        g.cxy = cxy;
        g.out4 = out4;
    }
    SC_HIP(hipGraphLaunch(g.exec, s));
    uint32_t anyfail = 0;
    {
        ReadBack rb(c);
        rb.add(&anyfail, static_cast<const uint32_t *>(g.words.p + 1), 1);
        rb.run();
    }
    if (!anyfail) return;
    // some pair is not covered: the staged path on the same rings (the
    // context's rings now hold the caller's coordinates; the set-up is redone
    // for them, which drops the graph)
    std::vector<int> bsz = c.bsz, csz = c.csz;
    ctx_set_pairs(c, bxy, bsz.data(), cxy, csz.data(), J, true);
    const int keep = c.small_fused;
    c.small_fused = 0;
    try {
        clip_jobs(c, 1, J);
    } catch (...) {
        c.small_fused = keep;
        throw;
    }
    c.small_fused = keep;
    batch_area(c, out4);
    c.small_fail = 1;
}

} // namespace sc
