// sc_foster.hip -- Foster/Greiner-Hormann labelling, traversal and clean-up
// (rows A16-A17) on the GPU, for a batch of jobs in the same launches.
//
// Reference: copyToLinkedList (src/clip.cpp:606-695), labelIntersections2,
// createResult2, cleanUpResult2 (foster/polyclip.h:337-787), for the
// intersection (UNION = false) and, on the sequential route, the union.  The augmented rings built by post_jobs are the linked
// lists of copyToLinkedList laid out in ring order: slot i of a job's base
// ring P is followed by slot i+1 (mod its length), likewise for the overlay
// ring Q; a slot is an intersection vertex when its alpha (P) or neighbour
// (Q) is set.  Job j owns P slots [AB[j], AB[j+1]) and Q slots
// [AC[j], AC[j+1]); neighbour links are 1-based within the job.
//
// Every job takes one of three routes:
//  - crossing (every label CROSSING or BOUNCING, a plain vertex on both
//    rings, crossings alternate, no interior bouncing pair): the entry/exit
//    status of a crossing vertex is the parity of the crossings met since the
//    first plain vertex (one prefix count); a traversal piece runs from a
//    crossing vertex to the next crossing vertex in the walking direction,
//    then jumps to the neighbour of that vertex on the other ring.  Pieces
//    form cycles; a cycle's output component is started at its base-ring
//    vertex that comes first in base-ring order, and the pieces of a
//    component are laid out by their distance from that start (one doubling
//    pass finds both);
//  - no crossing (same conditions, no crossing vertex): each ring is output
//    when its first plain vertex lies inside the other ring;
//  - sequential: everything else -- intersection chains (LEFT_ON ...
//    ON_RIGHT), split bouncing pairs, rings made only of intersections,
//    inconsistent labels -- runs the reference's sequential algorithm, one
//    device wave per such job.
// The outputs of all jobs are laid out job after job, then cleaned up.
//
// Clean-up: the reference removes, in one pass from the root, every vertex
// nearly collinear (|A| < 1e-9) with the last kept vertex and its successor,
// after first dropping collinear roots.  The decisions are computed as the
// unique fixed point of keep[i] = |A(last kept before i, i, i+1)| >= eps,
// iterated in parallel from the guess "predecessor kept"; the fixed point
// equals the sequential result because it is determined left to right.
//
// Provenance: every output vertex carries the slot it came from (a P slot g
// as g, a Q slot g as ~g, none as SRC_NONE) for the area gradient.
#include "sc_ctx.hpp"
#include "sc_fp64.hpp"

#include <algorithm>
#include <climits>

namespace sc {

namespace {

constexpr int BLK = 256;
inline unsigned grid_for(int64_t n, int per = BLK) {
    int64_t g = ceil_div(n, per);
    if (g < 1) g = 1;
    return unsigned(g);
}

enum : int {
    L_NONE = 0, L_CROSSING = 1, L_BOUNCING = 2, L_LEFT_ON = 3, L_RIGHT_ON = 4, L_ON_ON = 5,
    L_ON_LEFT = 6, L_ON_RIGHT = 7, L_DCROSS = 8, L_DBOUNCE = 9
};
enum : int { EXIT = 0, ENTRY = 1, NEITHER = 2 };
enum : int { UB_TRAVERSAL = 4 };
constexpr int SRC_NONE = INT_MIN;

// reasons for taking the sequential route
enum : int { SLOW_LABEL = 1, SLOW_NOPLAIN = 2, SLOW_ODD = 4, SLOW_SPLIT = 8, SLOW_LINK = 16, SLOW_UNION = 32 };
// routes
enum : int { K_CROSS = 0, K_NOCROSS = 1, K_SEQ = 2 };

constexpr uint32_t NIL = 0xffffffffu;

// one side (all P rings or all Q rings) of a batch
struct Side {
    const int *off;  // J+1 slot offsets of this side's rings
    const int *ooff; // J+1 slot offsets of the other side's rings
    const double *x, *y;
    const int *lab, *nbr, *av; // av null: intersection iff nbr set
};

__device__ inline bool inter_at(const Side &s, int64_t i) { return s.av ? s.av[i] != -100 : s.nbr[i] != -100; }

// crossing flags; label and link checks
__global__ void k_jclassify(Side s, int J, int64_t n, uint32_t *cross, int *slow) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int j = seg_of(s.off, J, i);
    uint32_t c = 0;
    if (inter_at(s, i)) {
        int lab = s.lab[i];
        if (lab != L_CROSSING && lab != L_BOUNCING) atomicOr(&slow[j], SLOW_LABEL);
        int nb = s.nbr[i];
        if (nb < 1 || nb > s.ooff[j + 1] - s.ooff[j]) atomicOr(&slow[j], SLOW_LINK);
        c = lab == L_CROSSING ? 1u : 0u;
    }
    cross[i] = c;
}

// first plain vertex of each job's ring (local index, -1 if none): one wave per job
__global__ __launch_bounds__(BLK) void k_first_plain(Side s, int J, int *first) {
    const int j = int((int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / WAVE);
    const int lane = threadIdx.x & (WAVE - 1);
    if (j >= J) return; // whole wave
    const int64_t a = s.off[j], b = s.off[j + 1];
    int res = -1;
    for (int64_t base = a; base < b; base += WAVE) {
        int64_t i = base + lane;
        bool plain = i < b && !inter_at(s, i);
        uint64_t m = __ballot(plain);
        if (m) {
            res = int(base - a) + __ffsll((unsigned long long)m) - 1;
            break;
        }
    }
    if (lane == 0) first[j] = res;
}

// winding test of polygon::pointInPoly (polygon.h:435-457) of each job's
// query point (the other ring's first plain vertex) against this side's
// rings, summed per job
__global__ __launch_bounds__(BLK) void k_jwinding(Side r, int J, int64_t n, const double *qx, const double *qy,
                                                  const int *qok, int *wind) {
    // qx[2j], qy[2j]: the query point of job j (interleaved x,y)
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int w = 0, j = -1;
    if (i < n) {
        j = seg_of(r.off, J, i);
        if (qok[j]) {
            int64_t nx = i + 1 == r.off[j + 1] ? r.off[j] : i + 1;
            w = wind_edge(r.x[i], r.y[i], r.x[nx], r.y[nx], qx[2 * j], qy[2 * j]);
        }
    }
    // a wave inside one job adds once; a wave across jobs adds per lane
    const int j0 = __shfl(j, 0, WAVE), j1 = __shfl(j, WAVE - 1, WAVE);
    if (j0 == j1 && j0 >= 0) {
        for (int d = WAVE / 2; d > 0; d >>= 1) w += __shfl_down(w, d, WAVE);
        if ((threadIdx.x & (WAVE - 1)) == 0 && w) atomicAdd(&wind[j0], w);
    } else if (w && j >= 0) {
        atomicAdd(&wind[j], w);
    }
}

// query points: each ring's first plain vertex (dense rings, or the sparse base ring)
__global__ void k_query_pts(int J, Side P, Side Q, int sparse, SparseP sp, const int *firstP, const int *firstQ,
                            double *pq, int *pok, double *qq, int *qok) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= J) return;
    const int fp = firstP[j], fq = firstQ[j];
    pok[j] = fp >= 0;
    qok[j] = fq >= 0;
    double x = 0, y = 0;
    if (fp >= 0) {
        if (sparse) {
            SlotInfo si = sp_slot(sp, j, fp);
            x = si.x;
            y = si.y;
        } else {
            x = P.x[P.off[j] + fp];
            y = P.y[P.off[j] + fp];
        }
    }
    pq[2 * j] = x;
    pq[2 * j + 1] = y;
    x = y = 0;
    if (fq >= 0) {
        x = Q.x[Q.off[j] + fq];
        y = Q.y[Q.off[j] + fq];
    }
    qq[2 * j] = x;
    qq[2 * j + 1] = y;
}

// ---- sparse base ring (one-to-many batches) ----

// first plain slot of each job's virtual base ring: the first slot not held by
// an intersection; one wave per job over its ascending slot list
__global__ __launch_bounds__(BLK) void k_sp_first_plain(SparseP sp, int J, int *first) {
    const int j = int((int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / WAVE);
    const int lane = threadIdx.x & (WAVE - 1);
    if (j >= J) return; // whole wave
    const int a = sp.P0[j], b = sp.P0[j + 1];
    const int nslots = sp.AB[j + 1] - sp.AB[j];
    int res = b - a < nslots ? b - a : -1;
    for (int base = a; base < b; base += WAVE) {
        const int i = base + lane;
        const bool gap = i < b && sp.slot[i] != i - a;
        const uint64_t m = __ballot(gap);
        if (m) {
            res = base - a + __ffsll((unsigned long long)m) - 1;
            break;
        }
    }
    if (lane == 0) first[j] = res < nslots ? res : -1;
}

// winding of each job's query point against every base edge (all jobs share
// the base ring); corrected below on the edges that hold intersections
constexpr int WCH = 4096; // base edges per block
__global__ __launch_bounds__(BLK) void k_sp_wind_base(int J, int nb, const double *__restrict__ bpoly,
                                                      const double *__restrict__ qq, const int *__restrict__ qok,
                                                      int *wind) {
    __shared__ int sh[BLK / WAVE];
    const int chunks = (nb + WCH - 1) / WCH;
    const int j = blockIdx.x / chunks, ch = blockIdx.x % chunks;
    if (j >= J || !qok[j]) return; // whole block
    const double rx = qq[2 * j], ry = qq[2 * j + 1];
    const int e1 = min(nb, (ch + 1) * WCH);
    int w = 0;
    for (int e = ch * WCH + threadIdx.x; e < e1; e += BLK) {
        const double *P = bpoly + 2 * int64_t(e);
        w += wind_edge(P[0], P[1], P[2], P[3], rx, ry);
    }
    for (int d = WAVE / 2; d > 0; d >>= 1) w += __shfl_down(w, d, WAVE);
    if ((threadIdx.x & (WAVE - 1)) == 0) sh[threadIdx.x / WAVE] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < BLK / WAVE; ++k) t += sh[k];
        if (t) atomicAdd(&wind[j], t);
    }
}

// an edge split by non-degenerate intersections: replace its full-edge term
// by the terms of its pieces (source vertex -> X1 -> ... -> next vertex)
__global__ void k_sp_wind_fix(int I, int J, SparseP sp, const double *__restrict__ qq, const int *__restrict__ qok,
                              int *wind) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= I) return;
    const int j = seg_of(sp.P0, J, i);
    if (!qok[j]) return;
    const int e = sp.edge[i];
    auto is_nd = [&](int k) { return sp.slot[k] >= 0 && sp.slot[k] == sp.edge[k] + 1 + sp.ndx[k]; };
    if (!is_nd(i)) return;
    // Base edge e = (a, b) with the run of points p1..pm inserted on it counts
    // -w(a, b) + w(a, p1) + w(p1, p2) + ... + w(pm, b) instead of w(a, b):
    // each point adds the term of the piece that ends at it, the last point
    // also the closing piece, the first the removal of the whole edge.
    const double rx = qq[2 * j], ry = qq[2 * j + 1];
    const double *V = sp.bpoly + 2 * int64_t(sp.bvoff[j]);
    const double ax = V[2 * e], ay = V[2 * e + 1], bx = V[2 * e + 2], by = V[2 * e + 3]; // closing vertex repeated
    const bool first = !(i > sp.P0[j] && sp.edge[i - 1] == e && is_nd(i - 1));
    const bool last = !(i + 1 < sp.P0[j + 1] && sp.edge[i + 1] == e && is_nd(i + 1));
    const double px = first ? ax : sp.x[i - 1], py = first ? ay : sp.y[i - 1];
    const double x = sp.x[i], y = sp.y[i];
    int w = wind_edge(px, py, x, y, rx, ry);
    if (first) w -= wind_edge(ax, ay, bx, by, rx, ry);
    if (last) w += wind_edge(x, y, bx, by, rx, ry);
    if (w) atomicAdd(&wind[j], w);
}

// entry/exit status on the sparse base ring (positions = its intersections)
__global__ void k_sp_enex(int I, int J, SparseP sp, const int *lab, const uint32_t *cross, const uint32_t *cc,
                          const int *first, const int *wind, const int *kind, int *enex, uint8_t *split) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= I) return;
    const int j = seg_of(sp.P0, J, i);
    if (kind[j] != K_CROSS) {
        enex[i] = NEITHER;
        split[i] = 0;
        return;
    }
    const int V = first[j];
    const int kV = sp_lower(sp, j, V);
    const uint32_t m = cc[sp.P0[j + 1]] - cc[sp.P0[j]];
    const uint32_t c = cc[i] - cc[kV] + (sp.slot[i] < V ? m : 0u);
    const int st0 = (wind[j] % 2) != 0 ? EXIT : ENTRY;
    const int st = st0 ^ int(c & 1u);
    enex[i] = cross[i] ? st : NEITHER;
    split[i] = (lab[i] == L_BOUNCING && st == EXIT) ? 1 : 0;
}

struct JobState {
    int J;
    int uni; // the union: every job takes the sequential route
    const int *AB, *AC;
    int *slow, *firstP, *firstQ, *windP, *windQ, *kind, *mP, *mQ, *cP0, *cQ0;
};

// route of each job; first crossing index of each job on both sides (J+1 threads)
__global__ void k_jdecide(JobState s, const uint32_t *ccP, const uint32_t *ccQ) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > s.J) return;
    s.cP0[j] = int(ccP[s.AB[j]]);
    s.cQ0[j] = int(ccQ[s.AC[j]]);
    if (j == s.J) return;
    int mp = int(ccP[s.AB[j + 1]] - ccP[s.AB[j]]), mq = int(ccQ[s.AC[j + 1]] - ccQ[s.AC[j]]);
    s.mP[j] = mp;
    s.mQ[j] = mq;
    int sl = s.slow[j];
    if (s.firstP[j] < 0 || s.firstQ[j] < 0) sl |= SLOW_NOPLAIN;
    if (mp != mq || (mp & 1)) sl |= SLOW_ODD;
    if (s.uni) sl |= SLOW_UNION;
    s.slow[j] = sl;
    s.kind[j] = sl ? K_SEQ : (mp == 0 ? K_NOCROSS : K_CROSS);
}

// entry/exit status of every crossing vertex from the parity of crossings
// since the first plain vertex V; bouncing vertices met with status EXIT are
// split candidates (labelIntersections2 step 4, polyclip.h:551-635)
__global__ void k_jenex(Side s, int J, int64_t n, const uint32_t *cross, const uint32_t *cc, const int *first,
                        const int *wind, const int *kind, int *enex, uint8_t *split) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int j = seg_of(s.off, J, i);
    if (kind[j] != K_CROSS) {
        enex[i] = NEITHER;
        split[i] = 0;
        return;
    }
    int64_t V = s.off[j] + first[j];
    uint32_t m = cc[s.off[j + 1]] - cc[s.off[j]];
    uint32_t c = cc[i] - cc[V] + (i < V ? m : 0u);
    int st0 = (wind[j] % 2) != 0 ? EXIT : ENTRY;
    int st = st0 ^ int(c & 1u);
    enex[i] = cross[i] ? st : NEITHER;
    split[i] = (inter_at(s, i) && s.lab[i] == L_BOUNCING && st == EXIT) ? 1 : 0;
}

__global__ void k_jsplit_pairs(Side P, int J, int64_t nP, const uint8_t *splitP, const uint8_t *splitQ, int *slow) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nP || !splitP[i]) return;
    int j = seg_of(P.off, J, i);
    int64_t q = P.ooff[j] + P.nbr[i] - 1;
    if (splitQ[q]) atomicOr(&slow[j], SLOW_SPLIT);
}

__global__ void k_compact_cross(int64_t n, const uint32_t *cross, const uint32_t *ccount, int *cpos) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && cross[i]) cpos[ccount[i]] = int(i);
}

struct PieceArgs {
    int J;
    Side P, Q;
    const int *PAB;  // J+1 slot offsets of the base rings (virtual in the sparse layout)
    int sparse;      // base side given as positions of a sparse ring
    SparseP sp;
    const int *cposP, *cposQ;
    const uint32_t *crossP, *crossQ, *ccountP, *ccountQ;
    const int *enexP, *enexQ;
    const int *cP0, *cQ0, *kind;
    int mP, mQ; // crossing totals of the batch
    int *ppos, *plen, *pdir, *pj;
    uint32_t *succ, *val;
    int *slow;
};

// pieces: crossing x (P crossings first, then Q crossings) to the next
// crossing in its walking direction; ppos is local to the job's ring
__global__ void k_jpieces(PieceArgs a) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= a.mP + a.mQ) return;
    const bool onP = x < a.mP;
    const Side &R = onP ? a.P : a.Q;
    const int *cpos = onP ? a.cposP : a.cposQ;
    const int *c0 = onP ? a.cP0 : a.cQ0;
    const int *enex = onP ? a.enexP : a.enexQ;
    const int xl = onP ? x : x - a.mP;
    const int j = seg_of(c0, a.J, xl);
    a.pj[x] = j;
    if (a.kind[j] != K_CROSS) {
        a.ppos[x] = 0;
        a.plen[x] = 0;
        a.pdir[x] = 1;
        a.succ[x] = uint32_t(x);
        a.val[x] = NIL;
        return;
    }
    const int m = c0[j + 1] - c0[j];
    const int jl = xl - c0[j];
    const int pos = cpos[xl];
    const int e = enex[pos];
    const int d = e == ENTRY ? 1 : -1;
    int jw = jl + d;
    if (jw < 0) jw += m;
    if (jw >= m) jw -= m;
    const int posW = cpos[c0[j] + jw];
    if (enex[posW] != (e ^ 1)) atomicOr(&a.slow[j], SLOW_ODD);
    // local slots of the two crossings and the ring length
    int sl, slW, n;
    if (onP) {
        n = a.PAB[j + 1] - a.PAB[j];
        sl = a.sparse ? a.sp.slot[pos] : pos - R.off[j];
        slW = a.sparse ? a.sp.slot[posW] : posW - R.off[j];
    } else {
        n = R.off[j + 1] - R.off[j];
        sl = pos - R.off[j];
        slW = posW - R.off[j];
    }
    int len = d > 0 ? slW - sl : sl - slW;
    if (len <= 0) len += n;
    // the neighbour of the far crossing on the other ring, as a crossing index
    const int wl = R.nbr[posW] - 1; // local slot on the other ring
    uint32_t s = uint32_t(x);
    bool ok;
    uint32_t occw = 0;
    if (onP) {
        const int64_t w = int64_t(R.ooff[j]) + wl;
        ok = a.crossQ[w] != 0;
        occw = a.ccountQ[w];
    } else if (a.sparse) {
        const int k = sp_find(a.sp, j, wl);
        ok = k >= 0 && a.crossP[k] != 0;
        occw = k >= 0 ? a.ccountP[k] : 0u;
    } else {
        const int64_t w = int64_t(R.ooff[j]) + wl;
        ok = a.crossP[w] != 0;
        occw = a.ccountP[w];
    }
    if (!ok) atomicOr(&a.slow[j], SLOW_LINK);
    else s = (onP ? uint32_t(a.mP) : 0u) + occw;
    a.ppos[x] = sl;
    a.plen[x] = len;
    a.pdir[x] = d;
    a.succ[x] = s;
    a.val[x] = onP ? min(2u * uint32_t(sl), 2u * uint32_t(slW) + 1u) : NIL;
}

// jobs found inconsistent while building pieces take the sequential route
__global__ void k_jredecide(int J, const int *slow, int *kind) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < J && kind[j] == K_CROSS && slow[j]) kind[j] = K_SEQ;
}

__global__ void k_jmask(int M, const int *pj, const int *kind, uint32_t *succ, uint32_t *val) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= M || kind[pj[x]] == K_CROSS) return;
    succ[x] = uint32_t(x);
    val[x] = NIL;
}

// Cycles of the crossing graph by doubling.  On a crossing job successors
// alternate between P and Q crossings and only P crossings carry values, so
// the doubling runs over the P crossings alone, stepping two pieces at a time
// (a P piece and the Q piece after it).  After r rounds the record of P
// crossing x describes the window of 2^r such steps from x (wrapping around
// x's cycle): the window's minimum value, the summed piece lengths before
// that minimum's first occurrence, the summed lengths of the whole window and
// the P crossing after the window.  A window sum is read only when the window
// holds no occurrence of the cycle's minimum, so does not wrap and stays below
// the cycle's total; the 32-bit wrap-around of longer windows is never read.
__global__ void k_jump_init(int mP, int M, const uint32_t *val, const uint32_t *succ, const int *plen, uint4 *rec) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= mP) return;
    const uint32_t q = succ[x];
    uint32_t n2 = q < uint32_t(M) ? succ[q] : uint32_t(x);
    if (n2 >= uint32_t(mP)) n2 = uint32_t(x); // not a P-Q-P step: a crossing outside the traversed jobs
    const uint32_t w = uint32_t(plen[x]) + (q < uint32_t(M) && q != uint32_t(x) ? uint32_t(plen[q]) : 0u);
    rec[x] = make_uint4(val[x], 0u, w, n2);
}

// Round r doubles every window (2^r crossings to 2^(r+1)) and sets moved[r]
// when some window's minimum changed.  Values on a cycle are distinct, so a
// round that moves no minimum finds every cycle no longer than the windows
// it read: minima and distances are final in both buffers (the round copied
// them unchanged), and every later round returns at once.
__global__ void k_jump_rank(int mP, const uint4 *__restrict__ rec, uint4 *__restrict__ rec2, uint32_t *moved,
                            int r) {
    if (r > 0 && moved[r - 1] == 0u) return;
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= mP) return;
    const uint4 a = rec[x];
    const uint4 b = rec[a.w];
    uint4 o;
    o.x = a.x <= b.x ? a.x : b.x;
    o.y = a.x <= b.x ? a.y : a.z + b.y;
    o.z = a.z + b.z;
    o.w = b.w;
    rec2[x] = o;
    if (o.x != a.x) moved[r] = 1u;
}

// Per crossing (P and Q): its cycle's minimum, the list start of a traversed
// cycle (the P crossing whose even value 2*slot is the minimum; every other
// value on the cycle is larger) and the summed piece lengths from the
// crossing up to that start, the whole cycle's total at the start itself.
// Equivalent to the list ranking of the cycle broken before its start.
__global__ void k_rank_final(int M, int mP, const uint4 *__restrict__ rec, const uint32_t *succ, const int *pj,
                             const int *AB, int sparse, SparseP sp, const uint32_t *ccountP, const int *plen,
                             uint32_t *cmin, uint32_t *inode, uint32_t *dist) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= M) return;
    // the P crossing whose record describes x's cycle: x itself, or a Q crossing's successor
    const uint32_t p = x < mP ? uint32_t(x) : succ[x];
    uint32_t m = NIL, dp = 0u;
    if (p < uint32_t(mP)) {
        const uint4 r = rec[p];
        m = r.x;
        dp = r.y;
    }
    bool trav = m != NIL && (m & 1u) == 0u;
    uint32_t I = NIL;
    if (trav) {
        int k = sparse ? sp_find(sp, pj[x], int(m >> 1)) : int(AB[pj[x]] + int(m >> 1));
        I = k >= 0 ? ccountP[k] : NIL;
        trav = k >= 0;
    }
    cmin[x] = m;
    inode[x] = I;
    uint32_t d = 0u;
    if (trav) {
        if (x >= mP) {
            d = uint32_t(plen[x]) + (p == I ? 0u : dp);
        } else if (uint32_t(x) != I) {
            d = dp;
        } else { // the start: the whole cycle, its own piece then the Q piece after it and on
            const uint32_t q = succ[x], n2 = succ[q];
            d = uint32_t(plen[x]) + uint32_t(plen[q]) + (n2 == I || n2 >= uint32_t(mP) ? 0u : rec[n2].y);
        }
    }
    dist[x] = d;
}

// component totals at their starts (base-ring order)
__global__ void k_start_totals(int mP, const int *cposP, const int *pj, const int *AB, const int *pslot,
                               const uint32_t *cmin, const uint32_t *dist, uint32_t *tot, uint32_t *isstart) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= mP) return;
    uint32_t m = cmin[x];
    const int sl = pslot ? pslot[cposP[x]] : cposP[x] - AB[pj[x]];
    bool s = m != NIL && (m & 1u) == 0u && (m >> 1) == uint32_t(sl);
    tot[x] = s ? dist[x] : 0u;
    isstart[x] = s ? 1u : 0u;
}

struct CountArgs {
    int J;
    const int *AB, *AC, *kind, *windP, *windQ, *cP0, *sidx, *seq_res;
    const uint32_t *base, *ridx;
    uint32_t *jv, *jr;
};

// output vertices and rings of every job before clean-up
__global__ void k_jcounts(CountArgs a) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.J) return;
    uint32_t v = 0, r = 0;
    const int k = a.kind[j];
    if (k == K_CROSS) {
        v = a.base[a.cP0[j + 1]] - a.base[a.cP0[j]];
        r = a.ridx[a.cP0[j + 1]] - a.ridx[a.cP0[j]];
    } else if (k == K_NOCROSS) {
        bool ip = (a.windP[j] % 2) != 0, iq = (a.windQ[j] % 2) != 0;
        v = (ip ? uint32_t(a.AB[j + 1] - a.AB[j]) : 0u) + (iq ? uint32_t(a.AC[j + 1] - a.AC[j]) : 0u);
        r = uint32_t(ip) + uint32_t(iq);
    } else {
        int s = a.sidx[j];
        r = uint32_t(a.seq_res[3 * s]);
        v = uint32_t(a.seq_res[3 * s + 1]);
    }
    a.jv[j] = v;
    a.jr[j] = r;
}

__global__ void k_jring_offsets(int mP, const uint32_t *isstart, const uint32_t *ridx, const uint32_t *base,
                                const int *pj, const int *cP0, const uint32_t *JV, const uint32_t *JR, int *off) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= mP || !isstart[x]) return;
    int j = pj[x];
    off[JR[j] + ridx[x] - ridx[cP0[j]]] = int(JV[j] + base[x] - base[cP0[j]]);
}

__global__ void k_jheads(int M, const uint32_t *inode, const uint32_t *dist, const uint32_t *base, const int *pj,
                         const int *cP0, const uint32_t *JV, uint32_t *head, uint32_t *piece_at) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= M) return;
    uint32_t I = inode[x];
    if (I == NIL) return;
    int j = pj[x];
    uint32_t o = JV[j] + (base[I] - base[cP0[j]]) + (dist[I] - dist[x]);
    head[o] = o;
    piece_at[o] = uint32_t(x);
}

struct EmitArgs {
    int J, mP;
    const int *PAB; // base ring slot offsets (virtual in the sparse layout)
    int sparse;
    SparseP sp;
    const uint32_t *JV;
    const int *kind, *ppos, *pdir, *pj;
    const uint32_t *head, *piece_at;
    Side P, Q;
    double *ox, *oy;
    int *osrc;
};

__global__ void k_jemit(EmitArgs a, int64_t n_out) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_out) return;
    const int j = seg_of(reinterpret_cast<const int *>(a.JV), a.J, i);
    if (a.kind[j] != K_CROSS) return;
    uint32_t s = a.head[i];
    uint32_t x = a.piece_at[s];
    int k = int(i - s) + 1;
    const bool onP = int(x) < a.mP;
    if (onP) {
        const int n = a.PAB[j + 1] - a.PAB[j];
        int pos = (a.ppos[x] + a.pdir[x] * k) % n;
        if (pos < 0) pos += n;
        if (a.sparse) {
            const SlotInfo si = sp_slot(a.sp, j, pos);
            a.ox[i] = si.x;
            a.oy[i] = si.y;
            a.osrc[i] = SRC_NONE;
        } else {
            const int g = a.P.off[j] + pos;
            a.ox[i] = a.P.x[g];
            a.oy[i] = a.P.y[g];
            a.osrc[i] = g;
        }
    } else {
        const int n = a.Q.off[j + 1] - a.Q.off[j];
        int pos = (a.ppos[x] + a.pdir[x] * k) % n;
        if (pos < 0) pos += n;
        const int g = a.Q.off[j] + pos;
        a.ox[i] = a.Q.x[g];
        a.oy[i] = a.Q.y[g];
        a.osrc[i] = ~g;
    }
}

// the base rings of no-crossing jobs in the sparse layout, by output position
__global__ void k_sp_nocross_p(int64_t n_out, int J, const uint32_t *JV, const int *kind, const int *windP,
                               SparseP sp, double *ox, double *oy, int *osrc) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_out) return;
    const int j = seg_of(reinterpret_cast<const int *>(JV), J, i);
    if (kind[j] != K_NOCROSS || (windP[j] % 2) == 0) return;
    const int q = int(i - JV[j]);
    if (q >= sp.AB[j + 1] - sp.AB[j]) return; // the overlay part
    const SlotInfo si = sp_slot(sp, j, q);
    ox[i] = si.x;
    oy[i] = si.y;
    osrc[i] = SRC_NONE;
}

// dense copies of the sparse base rings of the sequential jobs
__global__ void k_sp_materialize(int q0, int nseq, const int *jobs_of, const int64_t *moff, SparseP sp,
                                 const int *av, const int *nbr, const int *lab, double *mx, double *my, int *mav,
                                 int *mnbr, int *mlab) {
    const int q = q0 + int(blockIdx.y);
    if (q >= nseq) return;
    const int j = jobs_of[q];
    const int n = sp.AB[j + 1] - sp.AB[j];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const SlotInfo si = sp_slot(sp, j, t);
        const int64_t o = moff[q] + t;
        mx[o] = si.x;
        my[o] = si.y;
        mav[o] = si.pos >= 0 ? av[si.pos] : -100;
        mnbr[o] = si.pos >= 0 ? nbr[si.pos] : -100;
        mlab[o] = si.pos >= 0 ? lab[si.pos] : 0;
    }
}

// rings of no-crossing jobs: P's ring when its plain vertex lies inside Q, then Q's
__global__ void k_jcopy_nocross(Side s, bool sideP, int J, int64_t n, const int *kind, const int *windP,
                                const int *windQ, const int *AB, const uint32_t *JV, double *ox, double *oy,
                                int *osrc) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int j = seg_of(s.off, J, i);
    if (kind[j] != K_NOCROSS) return;
    bool ip = (windP[j] % 2) != 0, iq = (windQ[j] % 2) != 0;
    if (sideP ? !ip : !iq) return;
    int64_t dst = JV[j] + (sideP || !ip ? 0 : (AB[j + 1] - AB[j])) + (i - s.off[j]);
    ox[dst] = s.x[i];
    oy[dst] = s.y[i];
    osrc[dst] = sideP ? int(i) : ~int(i);
}

__global__ void k_jnocross_off(int J, const int *kind, const int *windP, const int *windQ, const int *AB,
                               const uint32_t *JV, const uint32_t *JR, int *off) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= J || kind[j] != K_NOCROSS) return;
    bool ip = (windP[j] % 2) != 0, iq = (windQ[j] % 2) != 0;
    uint32_t r = JR[j];
    if (ip) off[r++] = int(JV[j]);
    if (iq) off[r] = int(JV[j] + (ip ? uint32_t(AB[j + 1] - AB[j]) : 0u));
}

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// clean-up (cleanUpResult2)
// ---------------------------------------------------------------------------
__device__ inline bool collinear(const double *x, const double *y, int a, int b, int c) {
    return fabs(areaA(x[a], y[a], x[b], y[b], x[c], y[c])) < EPS;
}

// drop collinear roots: the root's predecessor stays the ring's last vertex
__global__ void k_root_loop(int nr, const int *off, const double *x, const double *y, int *rstart) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    int s = off[r], last = off[r + 1] - 1;
    int root = s;
    while (root <= last) {
        int nx = root == last ? root : root + 1;
        int pv = root == last ? root : last;
        if (collinear(x, y, pv, root, nx)) ++root;
        else break;
    }
    rstart[r] = root;
}

// keep0 keeps the decisions (bytes) for the clean-up rounds: a vertex whose
// last kept predecessor is the vertex before it is decided exactly so again
__global__ void k_keep0(int64_t n, const int *ringid, const int *off, const int *rstart, const double *x,
                        const double *y, uint32_t *keep, uint8_t *keep0) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int r = ringid[i];
    int rs = rstart[r], last = off[r + 1] - 1;
    uint32_t k;
    if (i < rs) k = 0;
    else if (i == rs) k = 1;
    else {
        int nx = i == last ? rs : int(i) + 1;
        k = collinear(x, y, int(i) - 1, int(i), nx) ? 0u : 1u;
    }
    keep[i] = k;
    keep0[i] = uint8_t(k);
}

// One clean-up round in one launch.  A vertex after its ring's root stays
// unless it is collinear with the last vertex kept before it (under the
// previous round's flags) and its successor (the root, for the ring's last
// vertex); the rounds repeat until no flag changes.
// Tile t is block t.  A block waits only for tiles of lower index, and every
// XCD dispatches its blocks in index order, so the lowest block not yet
// dispatched waits for nothing that does not finish: every wait ends (and a
// wait is also bounded, reported as an error).  A ticket counter instead of
// the block index costs one contended atomic per tile.  A tile publishes the
// largest mark (kept index + 1) among its vertices under the previous round's
// flags; wave 0 then walks back over earlier tiles' words, 64 per round trip,
// until one carries an inclusive prefix.  Each vertex so learns the last
// vertex kept before it and its new flag is written in place: a tile reads
// only its own flags, other tiles' only through the published words.
constexpr int KEEP_ITEMS = 8;
constexpr int KEEP_TILE = BLK * KEEP_ITEMS;
constexpr int KEEP_ROUNDS = 48;                // rounds before the exact sequential pass
constexpr int KEEP_BATCH = 3;                  // rounds launched per convergence check
constexpr int KEEP_SPIN_LIMIT = 1 << 20;       // polls of a predecessor window before giving up
constexpr int KS_EPOCH_SHIFT = 56;             // status word: epoch | flag | value
constexpr unsigned long long KS_PRE = 1ull << 54, KS_VAL = 0xffffffffull;

struct KeepRound {
    int64_t n;
    const int *ringid, *off, *rstart;
    const double *x, *y;
    uint32_t *keep;
    const uint8_t *keep0;       // k_keep0's decisions (predecessor = the vertex before)
    unsigned long long *status; // per tile; no word carries this round's epoch before the launch
    unsigned long long epoch;   // 1..KEEP_ROUNDS, shifted into place
    uint32_t *changed;          // set when a flag changes
    uint32_t *err;              // set when a wait gives up (the result is then invalid)
    uint32_t *tstate;           // per tile: flags changed in its last round, the prefix that round saw, its largest mark
    int first;                  // the first round: every tile decides
};

__global__ __launch_bounds__(BLK) void k_keep_round(KeepRound a) {
    __shared__ uint32_t s_prefix, s_skip, s_ch, s_fast, s_agg;
    __shared__ uint32_t s_wmax[BLK / WAVE];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), w = tid / WAVE;
    const uint32_t tile = blockIdx.x;
    const int64_t i0 = int64_t(tile) * KEEP_TILE + int64_t(tid) * KEEP_ITEMS;
    // a tile whose flags did not change in its last round still has the
    // largest mark it computed then: it publishes that without reading its
    // flags, and reads them only if its incoming prefix changed
    if (tid == 0) {
        const bool f = !a.first && a.tstate[3 * tile] == 0u;
        s_fast = f ? 1u : 0u;
        s_agg = f ? a.tstate[3 * tile + 2] : 0u;
    }
    __syncthreads();
    const bool fast = s_fast != 0u;
    uint32_t kp[KEEP_ITEMS];
    uint32_t ex = 0, agg = 0;
    // the tile's flags, the largest mark before each thread's vertices and
    // the tile's largest (every thread of the block calls it)
    auto load_scan = [&]() {
        uint32_t mine = 0; // marks grow with the index: the thread's largest is its last kept vertex
#pragma unroll
        for (int it = 0; it < KEEP_ITEMS; ++it) {
            const int64_t i = i0 + it;
            kp[it] = i < a.n ? a.keep[i] : 0u;
            if (kp[it]) mine = uint32_t(i) + 1u;
        }
        uint32_t inc = mine;
#pragma unroll
        for (int d = 1; d < WAVE; d <<= 1) {
            const uint32_t y = __shfl_up(inc, d, WAVE);
            if (lane >= d) inc = inc > y ? inc : y;
        }
        ex = __shfl_up(inc, 1, WAVE);
        if (lane == 0) ex = 0;
        if (lane == WAVE - 1) s_wmax[w] = inc;
        __syncthreads();
        agg = 0;
#pragma unroll
        for (int q = 0; q < BLK / WAVE; ++q) {
            const uint32_t m = s_wmax[q];
            if (q < w) ex = ex > m ? ex : m;
            agg = agg > m ? agg : m;
        }
    };
    if (fast) agg = s_agg;
    else load_scan();
    // marks grow with the index, so a tile holding a kept vertex knows its
    // inclusive prefix already (its own largest mark) and publishes it as one;
    // a look-back then normally stops at the nearest predecessor
    unsigned long long *my = a.status + tile;
    if (tid == 0)
        __hip_atomic_store(my, a.epoch | (tile == 0 || agg != 0 ? KS_PRE : 0ull) | agg, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    if (w == 0) {
        uint32_t prefix = 0;
        if (tile > 0) {
            const unsigned long long EM = 0xffull << KS_EPOCH_SHIFT;
            int64_t t = int64_t(tile) - 1; // nearest predecessor not yet accounted for
            int spins = 0;
            while (true) {
                const int64_t q = t - lane;
                const unsigned long long word =
                    q >= 0 ? __hip_atomic_load(a.status + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                           : (a.epoch | KS_PRE); // before tile 0: an empty prefix
                const bool ready = (word & EM) == a.epoch;
                const uint64_t notready = __ballot(!ready);
                const uint64_t pre = __ballot(ready && (word & KS_PRE));
                int upto = -1; // lanes 0..upto are summed
                bool finished = false;
                if (pre) {
                    const int p = __builtin_ctzll(pre);
                    const uint64_t need = p == WAVE - 1 ? ~0ull : ((2ull << p) - 1ull);
                    if ((notready & need) == 0) {
                        upto = p;
                        finished = true;
                    }
                } else if (notready == 0) {
                    upto = WAVE - 1;
                }
                if (upto >= 0) {
                    uint32_t v = lane <= upto ? uint32_t(word & KS_VAL) : 0u;
#pragma unroll
                    for (int d = 1; d < WAVE; d <<= 1) {
                        const uint32_t y = __shfl_xor(v, d, WAVE);
                        v = v > y ? v : y;
                    }
                    prefix = prefix > v ? prefix : v;
                    if (finished) break;
                    t -= WAVE;
                    spins = 0;
                } else {
                    if (++spins > KEEP_SPIN_LIMIT) {
                        if (lane == 0) atomicOr(a.err, 1u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            if (lane == 0 && agg == 0)
                __hip_atomic_store(my, a.epoch | KS_PRE | prefix, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) {
            s_prefix = prefix;
            // a tile whose flags did not change in its last round and whose
            // incoming prefix is the same sees the same last-kept vertices:
            // its decisions, and so its flags, stay as they are
            s_skip = fast && a.tstate[3 * tile + 1] == prefix ? 1u : 0u;
            s_ch = 0u;
        }
    }
    __syncthreads();
    if (s_skip) {
        if (tid == 0) a.tstate[3 * tile] = 0u;
        return;
    }
    if (fast) load_scan(); // its maximum is the one published
    uint32_t cur = s_prefix > ex ? s_prefix : ex; // mark of the last vertex kept before i
    // the thread's initial decisions, one byte per vertex (i0 is a multiple of 8)
    uint64_t k0w = 0;
    if (i0 + KEEP_ITEMS <= a.n) {
        k0w = *reinterpret_cast<const uint64_t *>(a.keep0 + i0);
    } else {
        for (int it = 0; it < KEEP_ITEMS; ++it)
            if (i0 + it < a.n) k0w |= uint64_t(a.keep0[i0 + it]) << (8 * it);
    }
    bool ch = false;
#pragma unroll
    for (int it = 0; it < KEEP_ITEMS; ++it) {
        const int64_t i = i0 + it;
        if (i < a.n) {
            uint32_t k = kp[it];
            if (cur == uint32_t(i)) {
                // the last vertex kept before i is i - 1: the initial decision's
                // inputs (for a vertex up to its ring's root that decision is
                // its flag, which no round changes)
                k = uint32_t(k0w >> (8 * it)) & 0xffu;
            } else {
                const int r = a.ringid[i];
                const int rs = a.rstart[r], last = a.off[r + 1] - 1;
                if (i > rs) {
                    const int pv = int(cur) - 1; // >= rs: the ring's root is always kept
                    const int nx = i == last ? rs : int(i) + 1;
                    k = collinear(a.x, a.y, pv, int(i), nx) ? 0u : 1u;
                }
            }
            if (k != kp[it]) {
                a.keep[i] = k;
                ch = true;
            }
        }
        if (kp[it]) cur = uint32_t(i) + 1u;
    }
    if (__ballot(ch) != 0ull && lane == 0) {
        atomicOr(a.changed, 1u);
        s_ch = 1u;
    }
    __syncthreads();
    if (tid == 0) {
        a.tstate[3 * tile] = s_ch;
        a.tstate[3 * tile + 1] = s_prefix;
        a.tstate[3 * tile + 2] = agg;
    }
}

// exact sequential pass per ring (used when the fixed point needs too many rounds)
__global__ void k_cleanup_seq(int nr, const int *off, const int *rstart, const double *x, const double *y,
                              uint32_t *keep) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    int s = off[r], last = off[r + 1] - 1, rs = rstart[r];
    for (int i = s; i < rs && i <= last; ++i) keep[i] = 0;
    if (rs > last) return;
    keep[rs] = 1;
    int lk = rs;
    for (int i = rs + 1; i <= last; ++i) {
        int nx = i == last ? rs : i + 1;
        if (collinear(x, y, lk, i, nx)) keep[i] = 0;
        else {
            keep[i] = 1;
            lk = i;
        }
    }
}

// ring heads carry their ring index; an inclusive max-scan spreads it
__global__ void k_ring_heads(int nr, const int *off, uint32_t *head) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    if (off[r + 1] > off[r]) head[off[r]] = uint32_t(r);
}

__global__ void k_ring_counts(int nr, const int *off, const uint32_t *kscan, uint32_t *cnt) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    uint32_t c = kscan[off[r + 1]] - kscan[off[r]];
    cnt[r] = c;
}

__global__ void k_compact_out(int64_t n, const uint32_t *keep, const uint32_t *kscan, const double *x,
                              const double *y, const int *src, double *x2, double *y2, int *src2) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    x2[kscan[i]] = x[i];
    y2[kscan[i]] = y[i];
    src2[kscan[i]] = src[i];
}

__global__ void k_ring_off_final(int nr, const uint32_t *cnt, const uint32_t *rscan, const uint32_t *nonempty,
                                 const int *off, const uint32_t *kscan, int *off2) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    if (cnt[r]) off2[nonempty[r]] = int(kscan[off[r]]);
}

__global__ void k_nonempty(int nr, const uint32_t *cnt, uint32_t *flag) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nr) flag[r] = cnt[r] ? 1u : 0u;
}

// per job: its first ring after clean-up (jr = first ring before it)
__global__ void k_job_rings(int J, const uint32_t *jr, const uint32_t *nonempty, uint32_t *out) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j <= J) out[j] = nonempty[jr[j]];
}

// ---------------------------------------------------------------------------
// sequential path: the reference's algorithm on index-linked vertices, one thread
// ---------------------------------------------------------------------------
struct SeqArgs {
    int nP, nQ, cap; // ring nodes of P and Q (an intersection vertex or a run of plain vertices)
    // node records: P nodes [0,nP), Q nodes [nP,nP+nQ)
    const double *nx0, *ny0, *nx1, *ny1; // first / last vertex of the node
    const int *nlen, *nslot, *ninter, *nlab, *nnb;
    const double *px, *py, *qx, *qy; // ring vertices by slot (run expansion)
    int windP, windQ;                // winding of each ring's first plain vertex w.r.t. the other ring
    double *vx, *vy, *vx1, *vy1;
    int *vlen, *vslot, *nx, *pv, *nb, *lab, *enex;
    uint8_t *inter;
    int *cand; // scratch list, cap entries
    int *seen; // 2 per vertex: the traversal component that last walked it in each direction
    int4 *rec; // per node for createResult2's walks: nx, pv, nb, enex | vlen << 2
    int *pstep; // per walk candidate: its walk's node count, then its first step record
    int2 *emit; // 2 * cap step records (node << 1 | forward, output place) of one round of walks
    double *ox, *oy;
    int *vsrc, *osrc; // provenance of inserted vertices / output vertices (slot reference)
    int P0, Q0;       // first slot of the job's rings in the batch arrays
    int uni;          // 1: the union (UNION of foster/polyclip.h:39 and its XORs)
    int par;          // 1: createResult2's rings walked one per lane first
    int out_cap;
    int *ring_off;
    int ring_cap;
    int *result; // [0] rings, [1] vertices, [2] status bits
};

// vertex t of node V in ring order
__device__ inline void seq_vertex(const SeqArgs &a, int V, int t, double &x, double &y) {
    if (a.vslot[V] >= 0 && a.vlen[V] > 1) {
        const bool onP = V < a.nP;
        const int sl = a.vslot[V] + t;
        x = onP ? a.px[sl] : a.qx[sl];
        y = onP ? a.py[sl] : a.qy[sl];
    } else {
        x = a.vx[V];
        y = a.vy[V];
    }
}

// point-in-polygon over the ring from root, every vertex of every node
__device__ __attribute__((always_inline)) bool seq_point_in(SeqArgs &a, int root, double rx, double ry) {
    int w = 0, one = root;
    bool have = false;
    double fx = 0, fy = 0, px = 0, py = 0;
    do {
        for (int t = 0; t < a.vlen[one]; ++t) {
            double cx, cy;
            seq_vertex(a, one, t, cx, cy);
            if (!have) {
                fx = cx;
                fy = cy;
                have = true;
            } else {
                w += wind_edge(px, py, cx, cy, rx, ry);
            }
            px = cx;
            py = cy;
        }
        one = a.nx[one];
    } while (one != root);
    w += wind_edge(px, py, fx, fy, rx, ry);
    return (w % 2) != 0;
}

// provenance of a single-vertex node
__device__ inline int seq_src(const SeqArgs &a, int V) {
    if (a.vslot[V] < 0) return a.vsrc[V];
    return V < a.nP ? a.P0 + a.vslot[V] : ~(a.Q0 + a.vslot[V]);
}

__device__ int seq_toggle(int s) { return s == ENTRY ? EXIT : (s == EXIT ? ENTRY : s); }

// node records -> the job's linked rings (all threads of the block)
// createResult2's walks, one per lane (see seq_run).  Returns PT_DONE with
// rings/outn advanced and the rings written, PT_FAIL when the single walk
// would fail (a ring that revisits a (node, direction) state, a jump to no
// neighbour, or the output capacities), or PT_FALLBACK after clearing what it
// wrote.  Candidate ranks live in seen[2v] (rank + 1) and the first ring that
// visits a candidate in seen[2v + 1] (as SEEN_FIRST - ring, kept by atomicMax),
// both zero before and after.
enum : int { PT_DONE = 0, PT_FAIL = 1, PT_FALLBACK = 2 };
constexpr int PT_MAX_STEPS = 1 << 15; // a longer ring walks in the single walk
constexpr int SEEN_FIRST = 1 << 30;
constexpr int PT_CHUNK = 2048; // candidates per round of walks

// one ring's walk from node I by one lane, as createResult2 walks it
// (foster/polyclip.h:717-759): returns 0 when it closes, 1 when it fails, 2
// when it is longer than PT_MAX_STEPS, 3 when 'visit' asks it to stop.
// 'visit' sees every node the walk marks (the start, every node stepped to,
// every jump target); 'step' sees every node stepped to with its direction
// and vertex count (the emitted ones).  A node's links, label and length come
// in one 16-byte record, so a step waits for one load.
template <class Visit, class Step>
__device__ __attribute__((always_inline)) int pt_walk(const SeqArgs &a, int I, Visit visit, Step step) {
    const bool uni = a.uni != 0;
    int V = I;
    if (visit(V)) return 3;
    int4 r = a.rec[V];
    int steps = 0, lam = 0, power = 1, sv = -1, sd = -1; // Brent's cycle test on (node, direction)
    do {
        const int e = r.w & 3;
        const int st = e == ENTRY ? EXIT : (e == EXIT ? ENTRY : e);
        do {
            const bool fwd = (st == EXIT) != uni;
            V = fwd ? r.x : r.y;
            const int dir = fwd ? 1 : 0;
            if (V == sv && dir == sd) return 1; // a state seen again: the walk never closes
            if (++lam == power) {
                sv = V;
                sd = dir;
                power <<= 1;
                lam = 0;
            }
            if (++steps > PT_MAX_STEPS) return 2;
            r = a.rec[V];
            if (visit(V)) return 3;
            step(V, fwd, r.w >> 2);
        } while (!((r.w & 3) == st) && V != I);
        if (V != I) {
            V = r.z;
            if (V < 0) return 1;
            if (visit(V)) return 3;
            r = a.rec[V];
        }
    } while (V != I);
    return 0;
}

// what the walks need from the wave that ran the steps before them, and
// what they give back (in LDS, one per block)
struct SeqShared {
    int go;                // 2: the entry/exit walks run, 1: createResult2's walks run, 0: result written
    int n, status;         // nodes (inserted ones included), status bits so far
    int rings, outn;       // rings and output vertices written so far
    int ptr;               // the walks' outcome (PT_*)
    int scan[2], sV[2];    // ring s: its entry/exit walk runs on every wave, from node sV[s]
    int sst[2];            //   with this status at sV[s]
    int nsplit[2], ncross[2]; // the split / crossing candidate lists' lengths
};

constexpr int SEQ_BLK = 1024; // threads of a sequential job's block

// exclusive scan of one value per thread over the block, in thread order;
// tot = the block's sum.  Every thread of the block calls it.
template <class T> __device__ inline T seq_block_scan(T c, T &tot, T *s_w) {
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    T incl = c;
    for (int d = 1; d < WAVE; d <<= 1) {
        const T o = __shfl_up(incl, d, WAVE);
        if (lane >= d) incl += o;
    }
    if (lane == WAVE - 1) s_w[wid] = incl;
    __syncthreads();
    T before = 0, t = 0;
    for (int w = 0; w < nw; ++w) {
        const T v = s_w[w];
        if (w < wid) before += v;
        t += v;
    }
    __syncthreads(); // s_w is free for the next call
    tot = t;
    return before + incl - c;
}

// minimum of one value per thread over the block.  Every thread calls it.
__device__ inline int seq_block_min(int x, int *s_w) {
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    for (int d = 1; d < WAVE; d <<= 1) x = min(x, __shfl_xor(x, d, WAVE));
    if (lane == 0) s_w[wid] = x;
    __syncthreads();
    int m = INT_MAX;
    for (int w = 0; w < nw; ++w) m = min(m, s_w[w]);
    __syncthreads();
    return m;
}

// Run by every thread of the block, with the same control flow in every
// wave (all loop bounds and branches that reach a barrier come from LDS or
// from values every thread computes alike).
__device__ __attribute__((always_inline)) int seq_par_traverse(SeqArgs &a, SeqShared &sh) {
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int n = sh.n, rings = sh.rings, outn = sh.outn;
    const int n0 = a.nP + a.nQ; // nodes before any insertion; inserted nodes follow
    if (n - n0 > 4096) return PT_FALLBACK;
    __shared__ int s_w[SEQ_BLK / WAVE];
    __shared__ int64_t s_w64[SEQ_BLK / WAVE];
    __shared__ int s_next, s_qend, s_long, s_res;
    const int half = a.cap / 2;
    int *cand = a.cand, *info = a.cand + half;
    // the walks' node records (the walks change none of these fields)
    for (int v = tid; v < n; v += nthr) a.rec[v] = make_int4(a.nx[v], a.pv[v], a.nb[v], a.enex[v] | (a.vlen[v] << 2));
    // 1) the candidates in ring order from the root: original P nodes in index
    //    order, each followed by the nodes inserted after it
    auto is_cand = [&](int v) { return a.inter[v] != 0 && a.lab[v] == L_CROSSING; };
    int ncand = 0;
    // back to all-zero state words (listed candidates only carry marks)
    auto clear = [&]() {
        __syncthreads();
        for (int q = tid; q < ncand; q += nthr) {
            const int v = cand[q];
            a.seen[2 * v] = 0;
            a.seen[2 * v + 1] = 0;
        }
        __syncthreads();
    };
    for (int w = 0; w < a.nP; w += nthr) {
        const int v = w + tid;
        int c = 0;
        if (v < a.nP) {
            c = is_cand(v) ? 1 : 0;
            for (int x = a.nx[v], g = 0; x >= n0 && g <= n - n0; x = a.nx[x], ++g) c += is_cand(x) ? 1 : 0;
        }
        int tot;
        const int ex = seq_block_scan(c, tot, s_w);
        if (ncand + tot > half) {
            clear();
            return PT_FALLBACK;
        }
        int q = ncand + ex;
        if (v < a.nP) {
            if (is_cand(v)) {
                cand[q] = v;
                a.seen[2 * v] = ++q;
            }
            for (int x = a.nx[v], g = 0; x >= n0 && g <= n - n0; x = a.nx[x], ++g)
                if (is_cand(x)) {
                    cand[q] = x;
                    a.seen[2 * x] = ++q;
                }
        }
        ncand += tot;
    }
    __syncthreads();
    // 2)-4) in chunks of PT_CHUNK candidates in rank order, so that the
    // single walk's first failing ring is found without walking every
    // candidate after it.
    int nring = 0;    // rings of the chunks before
    int64_t nout = 0; // their vertices
    for (int c0 = 0; c0 < ncand; c0 += PT_CHUNK) {
        const int c1 = min(ncand, c0 + PT_CHUNK);
        // 2) every candidate walks without writing until it marks a candidate
        //    of smaller rank (then it does not own its ring); an owner's walk
        //    gives its ring's vertex count or its failure.  The single walk
        //    ends at the first failing ring: nothing after that owner's rank
        //    matters.  Threads take candidates from a counter in LDS as their
        //    walks end, so a long walk holds one lane, not the others.
        if (tid == 0) {
            s_next = c0 + nthr;
            s_qend = c1;
            s_long = 0;
        }
        __syncthreads();
        for (int q = c0 + tid; q < c1;) {
            if (q >= __hip_atomic_load(&s_qend, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
            int cnt = 0, nst = 0;
            const int res = pt_walk(
                a, cand[q],
                [&](int v) {
                    const int r = a.seen[2 * v];
                    return r > 0 && r < q + 1;
                },
                [&](int, bool, int len) {
                    cnt += len;
                    ++nst;
                });
            if (res == 2) atomicOr(&s_long, 1);
            if (res == 1) atomicMin(&s_qend, q + 1);
            info[q] = res == 3 ? -1 : (res == 1 ? -2 : cnt);
            a.pstep[q] = nst;
            q = atomicAdd(&s_next, 1);
        }
        __syncthreads();
        const int qend = s_qend;
        const bool lng = s_long != 0;
        if (lng) {
            clear();
            return PT_FALLBACK;
        }
        const bool fail = qend < c1 || info[c1 - 1] == -2;
        // 3) ring order: owner r (in the order of the starts) gets ring index
        //    r (kept in its seen[2v], still positive) and its output place
        //    from a running sum (kept in info); then the owners walk again in
        //    any order, and every candidate a walk marks keeps the first ring
        //    that marked it
        const int nring0 = nring;
        const int64_t out0 = int64_t(outn) + nout; // the chunk's first output vertex
        int nrec = 0;                              // the chunk's step records
        bool allw = !fail;                         // every owner of the chunk writes
        for (int q0 = c0; q0 < qend; q0 += nthr) {
            const int q = q0 + tid;
            const int inf = q < qend ? info[q] : -1;
            const bool owner = inf != -1;
            int otot;
            const int oex = seq_block_scan(owner ? 1 : 0, otot, s_w);
            const int64_t c = owner && inf > 0 ? inf : 0;
            int64_t ctot;
            const int64_t cex = seq_block_scan(c, ctot, s_w64);
            int stot;
            const int sex = seq_block_scan(owner ? a.pstep[q] : 0, stot, s_w);
            const int r = nring + oex;
            const int64_t at = int64_t(outn) + nout + cex;
            bool write = true;
            if (owner) {
                write = !fail && rings + r < a.ring_cap && at + c <= a.out_cap;
                if (write) a.ring_off[rings + r] = int(at);
                a.seen[2 * cand[q]] = r + 1;
                info[q] = write ? int(at) : -3; // -3: an owner that writes nothing
                a.pstep[q] = nrec + sex;
            }
            int wtot;
            (void)seq_block_scan(write ? 0 : 1, wtot, s_w);
            if (wtot) allw = false;
            nring += otot;
            nout += ctot;
            nrec += stot;
        }
        // rings that share (node, direction) states can exceed the records'
        // room (2 per node); the single walk then decides
        if (allw && nrec > 2 * a.cap) {
            clear();
            return PT_FALLBACK;
        }
        if (tid == 0) s_next = c0 + nthr;
        __syncthreads();
        for (int q = c0 + tid; q < qend; q = atomicAdd(&s_next, 1)) {
            const int inf = info[q];
            if (inf == -1) continue;
            const int X = cand[q];
            const int tag = SEEN_FIRST - (a.seen[2 * X] - 1);
            const bool write = inf >= 0 && allw;
            int at = inf, k = a.pstep[q];
            (void)pt_walk(
                a, X,
                [&](int v) {
                    if (a.seen[2 * v] > 0) atomicMax(&a.seen[2 * v + 1], tag);
                    return false;
                },
                [&](int v, bool fwd, int len) {
                    if (write) a.emit[k++] = make_int2(v << 1 | (fwd ? 1 : 0), at);
                    at += len;
                });
        }
        __syncthreads();
        // the rings' vertices, one thread per output vertex: the step record
        // of vertex o is the last one placed at or before it (records follow
        // the output order)
        if (allw && nrec > 0) {
            const int O = int(int64_t(outn) + nout - out0);
            for (int o = tid; o < O; o += nthr) {
                const int oo = int(out0) + o;
                int lo = 0, hi = nrec - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (a.emit[mid].y <= oo) lo = mid;
                    else hi = mid - 1;
                }
                const int2 e = a.emit[lo];
                const int v = e.x >> 1, t = oo - e.y, len = a.rec[v].w >> 2;
                const int u = (e.x & 1) ? t : len - 1 - t;
                const int vs = a.vslot[v];
                const bool onP = v < a.nP;
                double x, y;
                if (vs >= 0 && len > 1) {
                    x = onP ? a.px[vs + u] : a.qx[vs + u];
                    y = onP ? a.py[vs + u] : a.qy[vs + u];
                } else {
                    x = a.vx[v];
                    y = a.vy[v];
                }
                a.ox[oo] = x;
                a.oy[oo] = y;
                a.osrc[oo] = vs >= 0 ? (onP ? a.P0 + vs + u : ~(a.Q0 + vs + u)) : a.vsrc[v];
            }
        }
        __syncthreads();
        // 4) the check: an owner is first marked by its own ring, any other
        //    candidate by a ring whose start comes before it.  At the first
        //    candidate where that does not hold, a candidate no earlier ring
        //    marked is where the single walk starts its next ring: if a full
        //    walk from it fails, the single walk fails there too.
        int before = nring0;
        for (int q0 = c0; q0 < qend; q0 += nthr) {
            const int q = q0 + tid;
            const int inf = q < qend ? info[q] : -1;
            const bool owner = inf != -1;
            int otot;
            const int r = before + seq_block_scan(owner ? 1 : 0, otot, s_w);
            bool bad = false, unmarked = false;
            if (q < qend) {
                const int fv = a.seen[2 * cand[q] + 1];
                const int first = fv > 0 ? SEEN_FIRST - fv : INT_MAX;
                bad = owner ? first != r : first >= r;
                unmarked = !owner && first >= r;
            }
            const int fb = seq_block_min(bad ? tid : INT_MAX, s_w);
            if (fb != INT_MAX) {
                if (tid == fb) {
                    int res = 0;
                    if (unmarked)
                        res = pt_walk(
                            a, cand[q], [&](int) { return false; }, [&](int, bool, int) {});
                    s_res = unmarked && res == 1 ? PT_FAIL : PT_FALLBACK;
                }
                __syncthreads();
                const int out = s_res;
                clear();
                return out;
            }
            before += otot;
        }
        if (fail) {
            clear();
            return PT_FAIL;
        }
    }
    clear();
    if (rings + nring > a.ring_cap || int64_t(outn) + nout > a.out_cap) return PT_FAIL;
    if (tid == 0) {
        sh.rings = rings + nring;
        sh.outn = outn + int(nout);
    }
    return PT_DONE;
}

// the entry/exit walk's state (status, first of a delayed pair) as 2 bits
// x = st | first << 1 (EXIT = 0, ENTRY = 1); a node's effect on it is a
// 4-entry table of 2-bit states, and tables compose along the ring
__device__ inline int seq_tapply(uint32_t t, int x) { return int((t >> (2 * x)) & 3u); }
__device__ inline uint32_t seq_tcompose(uint32_t f, uint32_t g) { // f first, then g
    uint32_t h = 0;
#pragma unroll
    for (int x = 0; x < 4; ++x) h |= ((g >> (2 * ((f >> (2 * x)) & 3u))) & 3u) << (2 * x);
    return h;
}

// The entry/exit walks of the rings marked in sh.scan, on every thread of the
// block: ring s from the node after sh.sV[s] around to sh.sV[s], WAVE x
// waves nodes per round.  Each wave scans its nodes' tables, the waves'
// composed tables give every wave the state the walk enters its nodes in
// (LDS), and each lane applies its node's effect; list appends keep ring
// order (ballot ranks within a wave, counts across waves) and the lists'
// bounds.  Every branch that reaches a barrier depends on LDS values only.
__device__ __attribute__((always_inline)) void seq_enex_block(SeqArgs &a, SeqShared &sh) {
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wid = tid / WAVE, nw = blockDim.x / WAVE;
    const int nthr = blockDim.x;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (WAVE - lane));
    __shared__ uint32_t s_tab[SEQ_BLK / WAVE];
    __shared__ int s_cs[SEQ_BLK / WAVE], s_cc[SEQ_BLK / WAVE];
    const int half = a.cap / 2, lim = half / 2;
    const bool uni = a.uni != 0;
    for (int s = 0; s < 2; ++s) {
        if (!sh.scan[s]) continue;
        int *splitL = s ? a.cand + half / 2 : a.cand;
        int *crossL = s ? a.cand + half + half / 2 : a.cand + half;
        const int r0 = s ? a.nP : 0, m = s ? a.nQ : a.nP, v0 = sh.sV[s] - r0;
        int state = sh.sst[s] | 2; // first = true
        int nsplit = sh.nsplit[s], ncross = sh.ncross[s];
        for (int pb = 1; pb < m; pb += nthr) {
            const int p = pb + wid * WAVE + lane;
            const bool on = p < m;
            int I = -1, lab = L_NONE;
            bool it = false;
            if (on) {
                I = r0 + (v0 + p >= m ? v0 + p - m : v0 + p);
                lab = a.lab[I];
                it = a.inter[I] != 0;
            }
            uint32_t tab = 0xE4u; // identity
            if (it) {
                if (lab == L_CROSSING) tab = 0xB1u;     // status toggles
                else if (lab == L_DCROSS) tab = 0x4Bu;  // first: clear it; else set it, toggle
                else if (lab == L_DBOUNCE) tab = 0x1Bu; // first flips, status toggles
            }
            uint32_t P = tab;
            for (int d = 1; d < WAVE; d <<= 1) {
                const uint32_t o = uint32_t(__shfl_up(int(P), d, WAVE));
                if (lane >= d) P = seq_tcompose(o, P);
            }
            // lanes past the ring's end hold the identity: lane 63 has the wave's table
            if (lane == WAVE - 1) s_tab[wid] = P;
            __syncthreads();
            int sw = state, send = state;
            for (int w = 0; w < nw; ++w) {
                const uint32_t t = s_tab[w];
                if (w < wid) sw = seq_tapply(t, sw);
                send = seq_tapply(t, send);
            }
            const uint32_t Pex = uint32_t(__shfl_up(int(P), 1, WAVE));
            const int sin = lane == 0 ? sw : seq_tapply(Pex, sw);
            const int sst = sin & 1;
            const bool sfirst = (sin >> 1) != 0;
            bool addS = false, addC = false;
            if (it) {
                if (lab == L_CROSSING) {
                    a.enex[I] = sst;
                } else if (lab == L_BOUNCING) {
                    addS = (sst == EXIT) != uni;
                } else if (lab == L_DCROSS) {
                    a.enex[I] = sst;
                    if (sfirst ? ((sst == EXIT) != uni) : ((sst == ENTRY) != uni)) a.lab[I] = L_CROSSING;
                } else if (lab == L_DBOUNCE) {
                    a.enex[I] = sst;
                    addC = sfirst ? ((sst == EXIT) != uni) : ((sst == ENTRY) != uni);
                }
            }
            const uint64_t ms = __ballot(addS), mc = __ballot(addC);
            if (lane == 0) {
                s_cs[wid] = int(__popcll(ms));
                s_cc[wid] = int(__popcll(mc));
            }
            __syncthreads();
            int bs = 0, bc = 0, ts = 0, tc = 0;
            for (int w = 0; w < nw; ++w) {
                const int cs = s_cs[w], cc = s_cc[w];
                if (w < wid) {
                    bs += cs;
                    bc += cc;
                }
                ts += cs;
                tc += cc;
            }
            if (addS) {
                const int q = nsplit + bs + int(__popcll(ms & lt));
                if (q < lim) splitL[q] = I;
            }
            if (addC) {
                const int q = ncross + bc + int(__popcll(mc & lt));
                if (q < lim) crossL[q] = I;
            }
            nsplit = min(nsplit + ts, lim);
            ncross = min(ncross + tc, lim);
            state = send;
            __syncthreads(); // s_tab, s_cs and s_cc are rewritten next round
        }
        if (tid == 0) {
            sh.nsplit[s] = nsplit;
            sh.ncross[s] = ncross;
        }
    }
}

// The reference's sequential algorithm on the job's rings, run by one wave
// with uniform control flow: every lane holds the same walk state.  Part 0
// runs the steps before the entry/exit walks and leaves its state in sh
// (sh.go = 2), part 2 the steps after them up to createResult2's walks
// (sh.go = 1); either writes the job's result instead when it fails (sh.go =
// 0).  Part 1 takes the walks' outcome (sh.ptr), runs the single walk when
// they fell back, and writes the result.  Lane i
// caches the hot fields of node w0 + i (a window of 64 consecutive nodes,
// w0 a multiple of 64), so a walk in ring order reads registers and reloads
// one coalesced window every 64 nodes.  Every write goes to memory (from all
// lanes, the same value) and to the window.  Output runs are written one
// vertex per lane.
__device__ __attribute__((always_inline)) void seq_run(SeqArgs &a, SeqShared &sh, int part) {
    const int lane = threadIdx.x & (WAVE - 1);
    int status = 0;
    // ring nodes already laid out by seq_init: P = [0,nP), Q = [nP,nP+nQ)
    int n = a.nP + a.nQ;
    int w0 = -(1 << 30);
    int c_nx = 0, c_pv = 0, c_nb = 0, c_lab = 0, c_enex = 0, c_inter = 0;
    int c_seen0 = 0, c_seen1 = 0, c_vlen = 0, c_vslot = 0;
    auto fill = [&](int I) {
        w0 = I & ~(WAVE - 1);
        const int j = w0 + lane;
        if (j < n) {
            c_nx = a.nx[j];
            c_pv = a.pv[j];
            c_nb = a.nb[j];
            c_lab = a.lab[j];
            c_enex = a.enex[j];
            c_inter = a.inter[j];
            const int2 sn = reinterpret_cast<const int2 *>(a.seen)[j];
            c_seen0 = sn.x;
            c_seen1 = sn.y;
            c_vlen = a.vlen[j];
            c_vslot = a.vslot[j];
        }
        // wait for the window here, so that its later reads wait for nothing
        // (the walk's own stores share the same counter)
        __builtin_amdgcn_s_waitcnt(0x0f70); // vmcnt(0), other counters free
    };
    auto slot = [&](int I) {
        if (unsigned(I - w0) >= unsigned(WAVE)) fill(I);
        return I - w0;
    };
    auto NX_ = [&](int I) { const int k = slot(I); return __builtin_amdgcn_readlane(c_nx, k); };
    auto PV_ = [&](int I) { const int k = slot(I); return __builtin_amdgcn_readlane(c_pv, k); };
    auto NB_ = [&](int I) { const int k = slot(I); return __builtin_amdgcn_readlane(c_nb, k); };
    auto LAB_ = [&](int I) { const int k = slot(I); return __builtin_amdgcn_readlane(c_lab, k); };
    auto ENEX_ = [&](int I) { const int k = slot(I); return __builtin_amdgcn_readlane(c_enex, k); };
    auto INTER_ = [&](int I) { const int k = slot(I); return __builtin_amdgcn_readlane(c_inter, k); };
    auto VLEN_ = [&](int I) { const int k = slot(I); return __builtin_amdgcn_readlane(c_vlen, k); };
    auto VSLOT_ = [&](int I) { const int k = slot(I); return __builtin_amdgcn_readlane(c_vslot, k); };
    // the traversal's state words, 2 per node (direction d)
    auto SEEN_ = [&](int I, int d) {
        const int k = slot(I);
        return d ? __builtin_amdgcn_readlane(c_seen1, k) : __builtin_amdgcn_readlane(c_seen0, k);
    };
    auto set_seen = [&](int I, int d, int v) {
        a.seen[2 * I + d] = v;
        if (I - w0 == lane) {
            if (d) c_seen1 = v;
            else c_seen0 = v;
        }
    };
    auto set_nx = [&](int I, int v) { a.nx[I] = v; if (I - w0 == lane) c_nx = v; };
    auto set_pv = [&](int I, int v) { a.pv[I] = v; if (I - w0 == lane) c_pv = v; };
    auto set_nb = [&](int I, int v) { a.nb[I] = v; if (I - w0 == lane) c_nb = v; };
    auto set_lab = [&](int I, int v) { a.lab[I] = v; if (I - w0 == lane) c_lab = v; };
    auto set_enex = [&](int I, int v) { a.enex[I] = v; if (I - w0 == lane) c_enex = v; };
    auto set_inter = [&](int I, int v) { a.inter[I] = uint8_t(v); if (I - w0 == lane) c_inter = v; };
    // a new single-vertex node (its own ring until inserted), -1 when full
    auto snew = [&](double x, double y, int src) -> int {
        if (n >= a.cap) return -1;
        const int v = n++;
        a.vx[v] = x;
        a.vy[v] = y;
        a.vx1[v] = x;
        a.vy1[v] = y;
        a.vlen[v] = 1;
        a.vslot[v] = -1;
        a.vsrc[v] = src;
        if (v - w0 == lane) {
            c_vlen = 1;
            c_vslot = -1;
            c_seen0 = c_seen1 = 0; // zeroed with the rest at set-up
        }
        set_nx(v, v);
        set_pv(v, v);
        set_nb(v, -1);
        set_lab(v, L_NONE);
        set_enex(v, NEITHER);
        set_inter(v, 0);
        return v;
    };
    auto sinsert = [&](int V, int curr) {
        const int c = NX_(curr);
        set_nx(PV_(c), V);
        set_pv(V, PV_(c));
        set_nx(V, c);
        set_pv(c, V);
    };
    // the vertices of node V, walking forward or backward, one per lane
    auto semit = [&](int &outn, int V, bool forward) -> bool {
        const int len = VLEN_(V);
        if (outn + len > a.out_cap) return false;
        const int vs = VSLOT_(V);
        const bool onP = V < a.nP;
        for (int t0 = 0; t0 < len; t0 += WAVE) {
            const int t = t0 + lane;
            if (t < len) {
                const int u = forward ? t : len - 1 - t;
                double x, y;
                seq_vertex(a, V, u, x, y);
                a.ox[outn + t] = x;
                a.oy[outn + t] = y;
                a.osrc[outn + t] = vs >= 0 ? (onP ? a.P0 + vs + u : ~(a.Q0 + vs + u)) : a.vsrc[V];
            }
        }
        outn += len;
        return true;
    };
    const int rootP = 0, rootQ = a.nP;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (WAVE - lane));
    if (part == 1) {
        n = sh.n;
        status = sh.status;
        int rings = sh.rings, outn = sh.outn;
        if (sh.ptr == PT_FAIL) goto done;
        if (sh.ptr == PT_FALLBACK) {
            int root = rootP, I = root;
            bool have = INTER_(I) && LAB_(I) == L_CROSSING;
            if (!have) {
                do { I = NX_(I); } while (!(INTER_(I) && LAB_(I) == L_CROSSING) && I != root);
                have = I != root;
            }
            long bound = 4L * n + 16;
            int comp = 0;
            while (have) {
                if (rings >= a.ring_cap) goto done;
                a.ring_off[rings++] = outn;
                int V = I;
                set_inter(V, 0);
                long steps = 0;
                ++comp;
                do {
                    int st = seq_toggle(ENEX_(V));
                    do {
                        const bool fwd = (st == EXIT) != (a.uni != 0); // (status == EXIT) ^ UNION, polyclip.h:740
                        V = fwd ? NX_(V) : PV_(V);
                        set_inter(V, 0);
                        // within one component the walk is determined by (vertex, direction):
                        // a repeated state can never reach I again (the reference loops forever)
                        const int dir = fwd ? 1 : 0;
                        if (SEEN_(V, dir) == comp) goto done;
                        set_seen(V, dir, comp);
                        if (++steps > bound || !semit(outn, V, fwd)) goto done;
                    } while (!(ENEX_(V) == st) && V != I);
                    if (V != I) {
                        V = NB_(V);
                        if (V < 0) goto done;
                        set_inter(V, 0);
                    }
                } while (V != I);
                do { I = NX_(I); } while (!(INTER_(I) && LAB_(I) == L_CROSSING) && I != root);
                have = I != root;
            }
        }
        a.ring_off[rings] = outn;
        a.result[0] = rings;
        a.result[1] = outn;
        a.result[2] = status;
        return;
    }
    if (part == 2) {
        n = sh.n;
        status = sh.status;
        const int rings = sh.rings, outn = sh.outn;
        const int half = a.cap / 2;
        const int nsplit[2] = {sh.nsplit[0], sh.nsplit[1]}, ncross[2] = {sh.ncross[0], sh.ncross[1]};
        int *splitL[2] = {a.cand, a.cand + half / 2};
        int *crossL[2] = {a.cand + half, a.cand + half + half / 2};
        // membership of the Q lists, marked in the traversal's (still zero)
        // state words: -1 in slot 2V for the split list, in 2V+1 for the crossing list
        const int nmark = n;
        for (int j = 0; j < nsplit[1]; ++j) set_seen(splitL[1][j], 0, -1);
        for (int j = 0; j < ncross[1]; ++j) set_seen(crossL[1][j], 1, -1);
        // 5) split interior bouncing pairs
        for (int i = 0; i < nsplit[0]; ++i) {
            int IP = splitL[0][i], IQ = NB_(IP);
            if (!(IQ >= 0 && IQ < nmark && SEEN_(IQ, 0) == -1)) continue;
            int VP = snew(a.vx[IP], a.vy[IP], seq_src(a, IP));
            int VQ = snew(a.vx[IQ], a.vy[IQ], seq_src(a, IQ));
            if (VP < 0 || VQ < 0) { status |= UB_TRAVERSAL; goto done; }
            // the vertex before IP is the last vertex of its previous node, the one after the first of the next
            double sP = areaA(a.vx1[PV_(IP)], a.vy1[PV_(IP)], a.vx[IP], a.vy[IP], a.vx[NX_(IP)], a.vy[NX_(IP)]);
            double sQ = areaA(a.vx1[PV_(IQ)], a.vy1[PV_(IQ)], a.vx[IQ], a.vy[IQ], a.vx[NX_(IQ)], a.vy[NX_(IQ)]);
            if (sP * sQ > 0) {
                set_nb(IP, VQ); set_nb(VQ, IP); set_inter(IP, 1); set_inter(VQ, 1);
                set_nb(IQ, VP); set_nb(VP, IQ); set_inter(IQ, 1); set_inter(VP, 1);
            } else {
                set_nb(VP, VQ); set_nb(VQ, VP); set_inter(VP, 1); set_inter(VQ, 1);
            }
            sinsert(VP, IP);
            sinsert(VQ, IQ);
            if (!a.uni) {
                set_enex(IP, EXIT); set_enex(VP, ENTRY); set_enex(IQ, EXIT); set_enex(VQ, ENTRY);
            } else {
                set_enex(IP, ENTRY); set_enex(VP, EXIT); set_enex(IQ, ENTRY); set_enex(VQ, EXIT);
            }
            set_lab(IP, L_CROSSING); set_lab(VP, L_CROSSING); set_lab(IQ, L_CROSSING); set_lab(VQ, L_CROSSING);
        }
        // 6) crossing candidates
        for (int i = 0; i < ncross[0]; ++i) {
            int IP = crossL[0][i], IQ = NB_(IP);
            if (IQ >= 0 && IQ < nmark && SEEN_(IQ, 1) == -1) {
                set_lab(IP, L_CROSSING);
                set_lab(IQ, L_CROSSING);
            }
        }
        for (int j = 0; j < nsplit[1]; ++j) set_seen(splitL[1][j], 0, 0);
        for (int j = 0; j < ncross[1]; ++j) set_seen(crossL[1][j], 1, 0);
        // createResult2.  First the walks run one per lane (seq_par_traverse,
        // every wave of the block): every candidate start (a P crossing, in
        // ring order from the root) walks its ring without writing; the
        // candidate with the smallest rank on a ring owns it; the owners walk
        // again and write the rings.  The reference's single walk starts a
        // ring at the first candidate that no earlier ring visited, so the
        // result is the same when every owner is first visited by its own
        // ring and every other candidate by a ring that starts before it;
        // that is checked, and otherwise (or when a walk is too long or too
        // many nodes were inserted) the single walk of part 1 runs.
        __threadfence_block(); // this wave's stores, before the other waves read them
        if (lane == 0) {
            sh.n = n;
            sh.status = status;
            sh.rings = rings;
            sh.outn = outn;
            sh.go = 1;
        }
        return;
    }
    // 2) intersection chains.  Before any insertion ring P is nodes [0, nP) in
    // ring order from its root, so the chain starts (intersections labelled
    // LEFT_ON or RIGHT_ON) are found 64 nodes per ballot, in that order, and
    // walked one after another.  A start that an earlier chain relabelled is
    // checked again and skipped, and a later block reads the labels the chains
    // before it wrote: the reference's single walk, step for step.
    for (int w = 0; w < a.nP; w += WAVE) {
        const int v = w + lane;
        bool start = false;
        if (v < a.nP) {
            const int lb = a.lab[v];
            start = a.inter[v] != 0 && (lb == L_LEFT_ON || lb == L_RIGHT_ON);
        }
        uint64_t starts = __ballot(start);
        while (starts) {
            const int I = w + __builtin_ctzll(starts);
            starts &= starts - 1ull;
            if (INTER_(I) && (LAB_(I) == L_LEFT_ON || LAB_(I) == L_RIGHT_ON)) {
                int x = LAB_(I) == L_LEFT_ON ? 0 : 1;
                int X = I, J = I, guard = 0;
                do {
                    set_lab(J, L_NONE);
                    if (NB_(J) >= 0) set_lab(NB_(J), L_NONE);
                    J = NX_(J);
                    if (++guard > n) { status |= UB_TRAVERSAL; goto done; }
                } while (LAB_(J) == L_ON_ON);
                int y = LAB_(J) == L_ON_LEFT ? 0 : 1;
                int ct = x != y ? L_DCROSS : L_DBOUNCE;
                set_lab(X, ct);
                set_lab(J, ct);
                if (NB_(X) >= 0) set_lab(NB_(X), ct);
                if (NB_(J) >= 0) set_lab(NB_(J), ct);
            }
        }
    }
    {
        int rings = 0, outn = 0;
        int roots[2] = {rootP, rootQ};
        bool noInter[2] = {false, false}, ident[2] = {false, false};
        // rings in index order (no insertion yet): ring s is nodes [ring0[s], ring0[s] + ringn[s])
        const int ring0[2] = {rootP, rootQ}, ringn[2] = {a.nP, a.nQ};
        // the ring's first node in ring order that is not an intersection, or -1
        auto first_plain = [&](int s) {
            for (int w = 0; w < ringn[s]; w += WAVE) {
                const int v = ring0[s] + w + lane;
                const uint64_t m = __ballot(w + lane < ringn[s] && a.inter[v] == 0);
                if (m) return ring0[s] + w + int(__builtin_ctzll(m));
            }
            return -1;
        };
        // 3.5) rings without crossing vertices (64 nodes per ballot)
        for (int s = 0; s < 2; ++s) {
            int A = roots[s], B = roots[1 - s];
            bool nocross = true, allon = true;
            int V = A;
            for (int w = 0; w < ringn[s] && nocross; w += WAVE) {
                const int v = ring0[s] + w + lane;
                bool cr = false, notOn = false;
                if (w + lane < ringn[s]) {
                    const int lb = a.lab[v];
                    const bool it = a.inter[v] != 0;
                    cr = it && (lb == L_CROSSING || lb == L_DCROSS);
                    // polygon::noCrossingVertex(union_case), polygon.h:465-475
                    if (a.uni && it && (lb == L_BOUNCING || lb == L_DBOUNCE)) cr = true;
                    notOn = lb != L_ON_ON;
                }
                if (__ballot(cr)) nocross = false;
                if (__ballot(notOn)) allon = false;
            }
            if (!nocross) continue;
            noInter[s] = true;
            if (allon) {
                ident[s] = true;
                continue;
            }
            double qx = 0, qy = 0;
            bool found = false, plain = false;
            V = first_plain(s);
            if (V >= 0) {
                qx = a.vx[V];
                qy = a.vy[V];
                found = plain = true;
            }
            if (!found) {
                V = A;
                do {
                    int nxv = NX_(V), nbv = NB_(V);
                    if (nbv >= 0 && NB_(nxv) != PV_(nbv) && NB_(nxv) != NX_(nbv)) {
                        qx = 0.5 * (a.vx[V] + a.vx[nxv]);
                        qy = 0.5 * (a.vy[V] + a.vy[nxv]);
                        found = true;
                        break;
                    }
                    V = NX_(V);
                } while (V != A);
            }
            if (!found) { status |= UB_TRAVERSAL; goto done; }
            // the ring's first plain vertex: its winding against the other ring is known
            const bool in = plain ? ((s == 0 ? a.windP : a.windQ) % 2) != 0 : seq_point_in(a, B, qx, qy);
            if (in != (a.uni != 0)) { // isInside ^ UNION, polyclip.h:470
                if (rings >= a.ring_cap) { status |= UB_TRAVERSAL; goto done; }
                a.ring_off[rings++] = outn;
                int W = A;
                do {
                    if (!semit(outn, W, true)) { status |= UB_TRAVERSAL; goto done; }
                    W = NX_(W);
                } while (W != A);
            }
        }
        if (ident[0] && ident[1]) {
            int target = NB_(rootP), V = rootQ;
            bool found = false;
            do {
                if (V == target) found = true;
                V = NX_(V);
            } while (V != rootQ);
            if (found) {
                if (rings >= a.ring_cap) { status |= UB_TRAVERSAL; goto done; }
                a.ring_off[rings++] = outn;
                int W = rootP;
                do {
                    if (!semit(outn, W, true)) { status |= UB_TRAVERSAL; goto done; }
                    W = NX_(W);
                } while (W != rootP);
            }
        }
        // 4) entry/exit flags; split and crossing candidates in scratch
        int nsplit[2] = {0, 0}, ncross[2] = {0, 0};
        bool inserted[2] = {false, false};
        int half = a.cap / 2; // cand[0..half) split lists, [half..cap) crossing lists
        int *splitL[2] = {a.cand, a.cand + half / 2};
        int *crossL[2] = {a.cand + half, a.cand + half + half / 2};
        // the entry/exit walk's state (status, first of a delayed pair) as 2 bits
        // x = st | first << 1 (EXIT = 0, ENTRY = 1); a node's effect on it is a
        // 4-entry table of 2-bit states, and tables compose along the ring
        if (lane == 0) sh.scan[0] = sh.scan[1] = 0;
        for (int s = 0; s < 2; ++s) {
            if (noInter[s]) continue;
            int A = roots[s], B = roots[1 - s];
            int V = first_plain(s), W = A; // ring s has no inserted node yet
            if (V < 0) {
                W = A;
                do {
                    int nxv = NX_(W), nbv = NB_(W);
                    if (nbv >= 0 && NB_(nxv) != PV_(nbv) && NB_(nxv) != NX_(nbv)) {
                        int T = snew(0.5 * (a.vx[W] + a.vx[nxv]), 0.5 * (a.vy[W] + a.vy[nxv]), SRC_NONE);
                        if (T < 0) { status |= UB_TRAVERSAL; goto done; }
                        sinsert(T, W);
                        inserted[s] = true;
                        V = T;
                        break;
                    }
                    W = NX_(W);
                } while (W != A);
            }
            if (V < 0) { status |= UB_TRAVERSAL; goto done; }
            int st = ENTRY;
            // V is the ring's first plain vertex unless a midpoint had to be inserted;
            // its winding against ring B is known while B has no inserted vertex
            const bool known = a.vslot[V] >= 0 && !inserted[1 - s];
            const bool in = known ? ((s == 0 ? a.windP : a.windQ) % 2) != 0 : seq_point_in(a, B, a.vx[V], a.vy[V]);
            if (in) st = seq_toggle(st);
            bool first = true;
            if (!inserted[s]) {
                // the walk from the node after V around to V runs on every wave
                // (seq_enex_block); no other step before it reads this ring's
                // labels or lists
                if (lane == 0) {
                    sh.scan[s] = 1;
                    sh.sV[s] = V;
                    sh.sst[s] = st;
                }
                continue;
            }
            int I = NX_(V);
            while (I != V) {
                if (INTER_(I)) {
                    int lab = LAB_(I);
                    if (lab == L_CROSSING) {
                        set_enex(I, st);
                        st = seq_toggle(st);
                    }
                    if (lab == L_BOUNCING && ((st == EXIT) != (a.uni != 0)) && nsplit[s] < half / 2)
                        splitL[s][nsplit[s]++] = I;
                    if (lab == L_DCROSS) {
                        set_enex(I, st);
                        if (first) {
                            if ((st == EXIT) != (a.uni != 0)) set_lab(I, L_CROSSING);
                            first = false;
                        } else {
                            if ((st == ENTRY) != (a.uni != 0)) set_lab(I, L_CROSSING);
                            first = true;
                            st = seq_toggle(st);
                        }
                    }
                    if (lab == L_DBOUNCE) {
                        set_enex(I, st);
                        if (first) {
                            if (((st == EXIT) != (a.uni != 0)) && ncross[s] < half / 2) crossL[s][ncross[s]++] = I;
                            first = false;
                        } else {
                            if (((st == ENTRY) != (a.uni != 0)) && ncross[s] < half / 2) crossL[s][ncross[s]++] = I;
                            first = true;
                        }
                        st = seq_toggle(st);
                    }
                }
                I = NX_(I);
            }
        }
        __threadfence_block(); // this wave's stores, before the other waves read them
        if (lane == 0) {
            sh.n = n;
            sh.status = status;
            sh.rings = rings;
            sh.outn = outn;
            sh.nsplit[0] = nsplit[0];
            sh.nsplit[1] = nsplit[1];
            sh.ncross[0] = ncross[0];
            sh.ncross[1] = ncross[1];
            sh.go = 2;
        }
        return;
    }
done:
    if (lane == 0) sh.go = 0;
    a.result[0] = 0;
    a.result[1] = 0;
    a.result[2] = status | UB_TRAVERSAL;
    a.ring_off[0] = 0;
}


// sequential route for several jobs: one block per job, its first wave runs seq_run
struct SeqJob {
    int j;                     // job
    int nP, nQ, P0, Q0;        // ring sizes (slots) and first slots in the dense ring arrays
    int node0, nPn, nQn;       // first node, node counts of P and Q
    int cap, out_cap, ring_cap;
    int64_t vbase, obase, rbase; // scratch bases: vertices, output vertices, ring offsets
};

struct SeqBatch {
    const SeqJob *jobs;
    const double *px, *py, *qx, *qy;
    const double *nx0, *ny0, *nx1, *ny1;
    const int *nlen, *nslot, *ninter, *nlab, *nnb;
    const int *windP, *windQ;
    double *vx, *vy, *vx1, *vy1;
    int *vlen, *vslot, *nx, *pv, *nb, *lab, *enex, *vsrc, *cand, *seen;
    uint8_t *inter;
    int4 *rec;
    int *pstep;
    int2 *emit;
    double *ox, *oy;
    int *osrc, *ring_off, *result;
    int uni;
    int par;
};

// one block per sequential job: its threads lay out the rings, its first wave
// runs the algorithm (no barrier after that point)
__global__ __launch_bounds__(SEQ_BLK) void k_sequential(SeqBatch b) {
    const SeqJob jb = b.jobs[blockIdx.x];
    SeqArgs a;
    a.nP = jb.nPn;
    a.nQ = jb.nQn;
    a.cap = jb.cap;
    a.P0 = jb.P0;
    a.Q0 = jb.Q0;
    a.uni = b.uni;
    a.par = b.par;
    a.nx0 = b.nx0 + jb.node0; a.ny0 = b.ny0 + jb.node0; a.nx1 = b.nx1 + jb.node0; a.ny1 = b.ny1 + jb.node0;
    a.nlen = b.nlen + jb.node0; a.nslot = b.nslot + jb.node0; a.ninter = b.ninter + jb.node0;
    a.nlab = b.nlab + jb.node0; a.nnb = b.nnb + jb.node0;
    a.px = b.px + jb.P0; a.py = b.py + jb.P0; a.qx = b.qx + jb.Q0; a.qy = b.qy + jb.Q0;
    a.windP = b.windP[jb.j];
    a.windQ = b.windQ[jb.j];
    a.vx = b.vx + jb.vbase; a.vy = b.vy + jb.vbase; a.vx1 = b.vx1 + jb.vbase; a.vy1 = b.vy1 + jb.vbase;
    a.vlen = b.vlen + jb.vbase; a.vslot = b.vslot + jb.vbase;
    a.nx = b.nx + jb.vbase; a.pv = b.pv + jb.vbase; a.nb = b.nb + jb.vbase; a.lab = b.lab + jb.vbase;
    a.enex = b.enex + jb.vbase; a.vsrc = b.vsrc + jb.vbase; a.cand = b.cand + jb.vbase;
    a.seen = b.seen + 2 * jb.vbase;
    a.rec = b.rec + jb.vbase;
    a.pstep = b.pstep + jb.vbase;
    a.emit = b.emit + 2 * jb.vbase;
    a.inter = b.inter + jb.vbase;
    a.ox = b.ox + jb.obase; a.oy = b.oy + jb.obase; a.osrc = b.osrc + jb.obase; a.out_cap = jb.out_cap;
    a.ring_off = b.ring_off + jb.rbase; a.ring_cap = jb.ring_cap;
    a.result = b.result + 3 * int64_t(blockIdx.x);
    if (jb.nPn + jb.nQn > jb.cap) { // cannot happen: cap is sized from the node counts
        if (threadIdx.x == 0) {
            a.result[0] = 0;
            a.result[1] = 0;
            a.result[2] = UB_TRAVERSAL;
            a.ring_off[0] = 0;
        }
        return;
    }
    __shared__ SeqShared sh;
    if (threadIdx.x == 0) sh.go = 0;
    __syncthreads();
    if (threadIdx.x < WAVE) seq_run(a, sh, 0); // wave 0: the steps before the entry/exit walks
    __syncthreads();
    if (!sh.go) return; // the job's result is written
    seq_enex_block(a, sh); // every wave
    __syncthreads();
    if (threadIdx.x < WAVE) seq_run(a, sh, 2); // wave 0: split and crossing candidates
    __syncthreads();
    if (!sh.go) return;
    const int ptr = a.par ? seq_par_traverse(a, sh) : PT_FALLBACK; // every wave
    __syncthreads();
    if (threadIdx.x == 0) sh.ptr = ptr;
    __syncthreads();
    if (threadIdx.x < WAVE) seq_run(a, sh, 1); // wave 0: the single walk if needed, the result
}

// Every sequential job's working rings before any insertion: its P and Q
// rings in node order (node v of the job at slot vbase + v), nothing
// inserted, no walk state.  Job q owns nodes [nodeoff[2q], nodeoff[2q+2]).
__global__ void k_seq_init(SeqBatch b, const int64_t *__restrict__ nodeoff, int nseq, int64_t N) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= N) return;
    int lo = 0, hi = nseq - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (nodeoff[2 * mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const SeqJob &jb = b.jobs[lo];
    const int v = int(t - jb.node0);
    const int64_t w = jb.vbase + v;
    b.vx[w] = b.nx0[t];
    b.vy[w] = b.ny0[t];
    b.vx1[w] = b.nx1[t];
    b.vy1[w] = b.ny1[t];
    b.vlen[w] = b.nlen[t];
    b.vslot[w] = b.nslot[t];
    b.vsrc[w] = SRC_NONE;
    b.lab[w] = b.nlab[t];
    const bool it = b.ninter[t] != 0;
    b.inter[w] = it ? 1 : 0;
    b.enex[w] = NEITHER;
    b.nb[w] = it ? b.nnb[t] : -1;
    const bool onP = v < jb.nPn;
    const int r0 = onP ? 0 : jb.nPn, m = onP ? jb.nPn : jb.nQn, i = v - r0;
    b.nx[w] = r0 + (i + 1 == m ? 0 : i + 1);
    b.pv[w] = r0 + (i == 0 ? m - 1 : i - 1);
}

// Ring nodes of the sequential jobs: every intersection slot is a node, every
// maximal run of plain slots is one node.  Segment g = 2q (ring P of sequential
// job q) or 2q+1 (its ring Q) covers flat slots [segoff[g], segoff[g+1]).
struct NodeArgs {
    int nseg;
    const int64_t *segoff;
    const SeqJob *jobs;
    const double *px, *py, *qx, *qy;
    const int *avP, *labP, *nbrP, *labQ, *nbrQ;
};

__device__ inline bool node_inter(const NodeArgs &a, int g, int s) {
    const SeqJob &jb = a.jobs[g >> 1];
    return (g & 1) ? a.nbrQ[jb.Q0 + s] != -100 : a.avP[jb.P0 + s] != -100;
}

// node offset of every segment: the node-start scan at the segment's first slot
__global__ void k_seg_node_offsets(int nseg, const int64_t *segoff, const uint32_t *nidx, int64_t *nodeoff) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g <= nseg) nodeoff[g] = int64_t(nidx[segoff[g]]);
}

__global__ void k_node_flags(NodeArgs a, int64_t S, uint32_t *flag) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= S) return;
    int lo = 0, hi = a.nseg - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (a.segoff[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const int g = lo, s = int(t - a.segoff[g]);
    flag[t] = (s == 0 || node_inter(a, g, s) || node_inter(a, g, s - 1)) ? 1u : 0u;
}

// node records (start slot, interior flag, first vertex, label, neighbour slot)
__global__ void k_node_fill(NodeArgs a, int64_t S, const uint32_t *flag, const uint32_t *nidx, int *nslot,
                            int *ninter, double *nx0, double *ny0, int *nlab, int *nnb) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= S || !flag[t]) return;
    int lo = 0, hi = a.nseg - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (a.segoff[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    const int g = lo, s = int(t - a.segoff[g]);
    const SeqJob &jb = a.jobs[g >> 1];
    const uint32_t k = nidx[t];
    const bool onQ = g & 1;
    const bool in = node_inter(a, g, s);
    nslot[k] = s;
    ninter[k] = in ? 1 : 0;
    nx0[k] = onQ ? a.qx[jb.Q0 + s] : a.px[jb.P0 + s];
    ny0[k] = onQ ? a.qy[jb.Q0 + s] : a.py[jb.P0 + s];
    nlab[k] = onQ ? a.labQ[jb.Q0 + s] : a.labP[jb.P0 + s];
    nnb[k] = in ? (onQ ? a.nbrQ[jb.Q0 + s] : a.nbrP[jb.P0 + s]) - 1 : -1;
}

// run lengths, last vertices and neighbour nodes (node indices local to the job)
__global__ void k_node_finish(NodeArgs a, int64_t N, const int64_t *nodeoff, const uint32_t *flag,
                              const uint32_t *nidx, const int *nslot, const int *ninter, int *nlen, double *nx1,
                              double *ny1, int *nnb) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= N) return;
    int lo = 0, hi = a.nseg - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (nodeoff[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    const int g = lo;
    const SeqJob &jb = a.jobs[g >> 1];
    const bool onQ = g & 1;
    const int nsl = int(a.segoff[g + 1] - a.segoff[g]);
    const int end = k + 1 < nodeoff[g + 1] ? nslot[k + 1] : nsl;
    const int len = end - nslot[k];
    nlen[k] = len;
    const int sl = nslot[k] + len - 1;
    nx1[k] = onQ ? a.qx[jb.Q0 + sl] : a.px[jb.P0 + sl];
    ny1[k] = onQ ? a.qy[jb.Q0 + sl] : a.py[jb.P0 + sl];
    if (ninter[k]) {
        const int og = g ^ 1; // the other ring of the same job
        const int w = nnb[k];
        const int on = int(a.segoff[og + 1] - a.segoff[og]);
        if (w < 0 || w >= on) {
            nnb[k] = -1;
        } else {
            const int64_t t = a.segoff[og] + w;
            const int64_t node = int64_t(nidx[t]) + int64_t(flag[t]) - 1; // node holding that slot
            nnb[k] = int(node - nodeoff[g & ~1]);
        }
    }
}

// rings of the sequential jobs into the batch output (one block per job)
__global__ void k_jseq_copy(const SeqJob *jobs, const int *result, const uint32_t *JV, const uint32_t *JR,
                            const double *sx, const double *sy, const int *ssrc, const int *sroff, double *ox,
                            double *oy, int *osrc, int *off) {
    const SeqJob jb = jobs[blockIdx.x];
    const int rings = result[3 * blockIdx.x], verts = result[3 * blockIdx.x + 1];
    const int64_t o = JV[jb.j];
    for (int i = threadIdx.x; i < verts; i += blockDim.x) {
        ox[o + i] = sx[jb.obase + i];
        oy[o + i] = sy[jb.obase + i];
        osrc[o + i] = ssrc[jb.obase + i];
    }
    for (int r = threadIdx.x; r < rings; r += blockDim.x) off[JR[jb.j] + r] = int(o + sroff[jb.rbase + r]);
}

// scratch owned by this stage
struct FosterScratch {
    DBuf<uint32_t> crossP, crossQ, ccP, ccQ;
    DBuf<uint32_t> jring; // per job first ring after clean-up, then the output totals
    DBuf<int> enexP, enexQ, cposP, cposQ;
    DBuf<uint8_t> splitP, splitQ;
    DBuf<int> ppos, plen, pdir, pj;
    DBuf<uint4> jrec, jrec2; // doubling records of the crossing cycles
    DBuf<uint32_t> succ, val, val2, inode, dist, tot, isstart, ringidx, head, piece_at;
    DBuf<int> jst;
    DBuf<uint32_t> jv, jr;
    DBuf<double> ox, oy, qpts, mx, my, nx0, ny0, nx1, ny1, vx1, vy1;
    DBuf<int> osrc, qoks, mav, mnbr, mlab, mjob, nslot, ninter, nlab, nnb, nlen, vlen, vslot, seen;
    DBuf<int64_t> moff, segoff, nodeoff;
    DBuf<uint32_t> nflag, nidx;
    DBuf<int> off, rstart, ringid;
    DBuf<uint32_t> keep, kscan, rcnt, nonempty, kctl, ktstate;
    DBuf<uint8_t> keep0;
    DBuf<unsigned long long> kstatus; // clean-up rounds: per tile status words
    // sequential route
    DBuf<SeqJob> sjobs;
    DBuf<int> sres, sroff, sosrc;
    DBuf<double> vx, vy, sox, soy;
    DBuf<int> nx, pv, nb, lab, enex, vsrc, cand;
    DBuf<uint8_t> inter;
    DBuf<int4> srec;
    DBuf<uint32_t> jmoved; // per doubling round: some cycle minimum moved
    DBuf<int> spstep;
    DBuf<int2> semit;
};

FosterScratch &fscratch() {
    thread_local FosterScratch *p = new FosterScratch();
    return *p;
}

template <class T> void d2h_sync(Ctx &c, T *h, const T *d, int64_t n) {
    ReadBack rb(c);
    rb.add(h, d, size_t(n));
    rb.run();
}

// clean-up of nr rings laid out at off[0..nr] in (ox, oy, osrc); result into
// Ctx; ring_idx[r] = index of ring r among the non-empty rings (nr+1 entries)
// clean-up of the rings; job_ring gets, per job, its first ring after it (J+1
// entries, from the jobs' first rings before it, jr on the device)
void cleanup_rings(Ctx &c, FosterScratch &f, int nr, int64_t n, int J, const uint32_t *jr,
                   std::vector<uint32_t> &job_ring) {
    hipStream_t s = c.stream;
    if (nr == 0 || n == 0) {
        job_ring.assign(size_t(J) + 1, 0u);
        c.nrings = 0;
        c.n_out = 0;
        c.d_ring_off.ensure(1);
        SC_HIP(hipMemsetAsync(c.d_ring_off.p, 0, sizeof(int), s));
        return;
    }
    int *rstart = f.rstart.ensure(nr);
    int *ringid = f.ringid.ensure(size_t(n));
    k_root_loop<<<grid_for(nr), BLK, 0, s>>>(nr, f.off, f.ox, f.oy, rstart);
    {
        FillBatch fb(s);
        fb.add(ringid, n, 0u);
        fb.flush();
    }
    k_ring_heads<<<grid_for(nr), BLK, 0, s>>>(nr, f.off, reinterpret_cast<uint32_t *>(ringid));
    scan_inclusive_max_u32(reinterpret_cast<uint32_t *>(ringid), reinterpret_cast<uint32_t *>(ringid), n, s);
    uint32_t *keep = f.keep.ensure(size_t(n) + 1);
    uint8_t *keep0 = f.keep0.ensure(size_t(n) + KEEP_ITEMS);
    k_keep0<<<grid_for(n), BLK, 0, s>>>(n, ringid, f.off, rstart, f.ox, f.oy, keep, keep0);
    SC_HIP(hipGetLastError());
    // clean-up rounds, KEEP_BATCH per look at the per-round change flags; a
    // round after one without change changes nothing either
    const int64_t ntiles = (n + KEEP_TILE - 1) / KEEP_TILE;
    uint32_t *ctl = f.kctl.ensure(KEEP_ROUNDS + 1); // changed flags, error word
    unsigned long long *kst = f.kstatus.ensure(size_t(ntiles));
    uint32_t *nonempty = f.nonempty.ensure(nr + 1);
    {
        FillBatch fb(s);
        fb.add(ctl, KEEP_ROUNDS + 1, 0u);
        fb.add(kst, 2 * ntiles, 0u);
        fb.add(keep + n, 1, 0u);         // closing entries of the two scans below
        fb.add(nonempty + nr, 1, 0u);
        fb.flush();
    }
    KeepRound kr;
    kr.n = n;
    kr.ringid = ringid;
    kr.off = f.off;
    kr.rstart = rstart;
    kr.x = f.ox;
    kr.y = f.oy;
    kr.keep = keep;
    kr.keep0 = keep0;
    kr.status = kst;
    kr.err = ctl + KEEP_ROUNDS;
    kr.tstate = f.ktstate.ensure(3 * size_t(ntiles)); // written by every tile in the first round
    bool converged = false;
    for (int it = 0; it < KEEP_ROUNDS && !converged;) {
        const int end = std::min(it + KEEP_BATCH, KEEP_ROUNDS);
        for (; it < end; ++it) {
            kr.epoch = (unsigned long long)(it + 1) << KS_EPOCH_SHIFT;
            kr.changed = ctl + it;
            kr.first = it == 0 ? 1 : 0;
            k_keep_round<<<unsigned(ntiles), BLK, 0, s>>>(kr);
            SC_HIP(hipGetLastError());
        }
        // one copy from the batch's last change flag up to the error word
        uint32_t h[KEEP_ROUNDS + 1];
        d2h_sync(c, h, ctl + it - 1, int64_t(KEEP_ROUNDS - it + 2));
        SC_REQUIRE(h[KEEP_ROUNDS - it + 1] == 0, SC_EHIP, "clean-up scan gave up waiting for an earlier tile");
        converged = h[0] == 0;
    }
    if (!converged) {
        k_cleanup_seq<<<grid_for(nr), BLK, 0, s>>>(nr, f.off, rstart, f.ox, f.oy, keep);
        SC_HIP(hipGetLastError());
    }
    // compaction
    uint32_t *kscan = f.kscan.ensure(size_t(n) + 1);
    uint32_t *tot = c.d_u32_total.ensure(2);
    scan_exclusive_u32(keep, kscan, n + 1, tot, s);
    uint32_t *rcnt = f.rcnt.ensure(nr + 1);
    k_ring_counts<<<grid_for(nr), BLK, 0, s>>>(nr, f.off, kscan, rcnt);
    SC_HIP(hipGetLastError());
    // surviving rings: exclusive scan of the non-empty flags; both totals and
    // the jobs' first rings come back in one copy
    k_nonempty<<<grid_for(nr), BLK, 0, s>>>(nr, rcnt, nonempty);
    SC_HIP(hipGetLastError());
    scan_exclusive_u32(nonempty, nonempty, int64_t(nr) + 1, tot + 1, s);
    uint32_t *jring = f.jring.ensure(size_t(J) + 3);
    k_job_rings<<<grid_for(int64_t(J) + 1), BLK, 0, s>>>(J, jr, nonempty, jring);
    SC_HIP(hipGetLastError());
    SC_HIP(hipMemcpyAsync(jring + J + 1, tot, sizeof(uint32_t) * 2, hipMemcpyDeviceToDevice, s));
    job_ring.resize(size_t(J) + 3);
    d2h_sync(c, job_ring.data(), jring, int64_t(J) + 3);
    const uint32_t nout = job_ring[size_t(J) + 1];
    c.nrings = int(job_ring[size_t(J) + 2]);
    job_ring.resize(size_t(J) + 1);
    c.n_out = nout;
    c.d_ring_x.ensure(nout);
    c.d_ring_y.ensure(nout);
    c.d_ring_src.ensure(nout);
    int *off2 = c.d_ring_off.ensure(size_t(c.nrings) + 1);
    k_compact_out<<<grid_for(n), BLK, 0, s>>>(n, keep, kscan, f.ox, f.oy, f.osrc, c.d_ring_x, c.d_ring_y,
                                              c.d_ring_src);
    k_ring_off_final<<<grid_for(nr), BLK, 0, s>>>(nr, rcnt, nullptr, nonempty, f.off, kscan, off2);
    SC_HIP(hipGetLastError());
    {
        FillBatch fb(s);
        fb.add(off2 + c.nrings, 1, uint32_t(int(nout)));
        fb.flush();
    }
    c.wait();
}

} // namespace

void release_foster_scratch() {
    FosterScratch &f = fscratch();
    for (DBuf<uint32_t> *b : {&f.jring, &f.crossP, &f.crossQ, &f.ccP, &f.ccQ, &f.succ, &f.val, &f.val2,
                              &f.inode, &f.dist, &f.tot, &f.isstart, &f.ringidx, &f.head, &f.piece_at,
                              &f.jv, &f.jr, &f.keep, &f.kscan, &f.rcnt, &f.nonempty, &f.kctl, &f.ktstate,
                              &f.jmoved})
        b->release();
    f.keep0.release();
    f.srec.release();
    f.spstep.release();
    f.semit.release();
    f.jrec.release();
    f.kstatus.release();
    f.jrec2.release();
    for (DBuf<int> *b : {&f.enexP, &f.enexQ, &f.cposP, &f.cposQ, &f.ppos, &f.plen, &f.pdir, &f.pj, &f.jst,
                         &f.osrc, &f.off, &f.rstart, &f.ringid, &f.sres, &f.sroff, &f.sosrc, &f.nx, &f.pv, &f.nb,
                         &f.lab, &f.enex, &f.vsrc, &f.cand})
        b->release();
    for (DBuf<double> *b : {&f.ox, &f.oy, &f.vx, &f.vy, &f.sox, &f.soy, &f.qpts, &f.mx, &f.my, &f.nx0, &f.ny0,
                             &f.nx1, &f.ny1, &f.vx1, &f.vy1})
        b->release();
    for (DBuf<int> *b : {&f.qoks, &f.mav, &f.mnbr, &f.mlab, &f.mjob, &f.nslot, &f.ninter, &f.nlab, &f.nnb, &f.nlen,
                         &f.vlen, &f.vslot, &f.seen})
        b->release();
    f.moff.release();
    f.segoff.release();
    f.nodeoff.release();
    f.nflag.release();
    f.nidx.release();
    f.splitP.release();
    f.splitQ.release();
    f.inter.release();
    f.sjobs.release();
}

void label_and_traverse(Ctx &c) {
    JobSet &js = c.jobs;
    SC_REQUIRE(js.J >= 1 && int(js.AB.size()) == js.J + 1, SC_ESTATE, "segclip_intersections must run first");
    hipStream_t s = c.stream;
    FosterScratch &f = fscratch();
    const int J = js.J;
    const bool sparse = c.sparseP;
    const int *AB = js.d_AB.p, *AC = js.d_AC.p;
    // the base side is either its dense rings (slots) or, in the sparse layout,
    // its intersections (positions); the overlay side is always dense
    const SparseP sp = sparse ? sparse_view(c) : SparseP{};
    const int *PO = sparse ? js.d_p0 : AB; // offsets of the base-side arrays
    const int64_t nP = sparse ? js.p0[J] : js.AB[J], nQ = js.AC[J];
    const Side P = sparse ? Side{PO, AC, c.sp_x, c.sp_y, c.sp_lab, c.sp_nbr, c.sp_av}
                          : Side{AB, AC, c.aug_bx, c.aug_by, c.lab_b, c.nbr_b, c.av_b};
    const Side Q{AC, AB, c.aug_cx, c.aug_cy, c.lab_c, c.nbr_c, nullptr};
    // per-job words
    int *jw = f.jst.ensure(11 * int64_t(J) + 4);
    {
        FillBatch fb(s);
        fb.add(jw, 11 * int64_t(J) + 4, 0u);
        fb.flush();
    }
    JobState st;
    st.J = J;
    st.uni = c.uni;
    st.AB = PO;
    st.AC = AC;
    st.slow = jw;
    st.firstP = jw + J;
    st.firstQ = jw + 2 * J;
    st.windP = jw + 3 * J;
    st.windQ = jw + 4 * J;
    st.kind = jw + 5 * J;
    st.mP = jw + 6 * J;
    st.mQ = jw + 7 * J;
    st.cP0 = jw + 8 * J;          // J+1
    st.cQ0 = jw + 9 * J + 1;      // J+1
    int *sidx = jw + 10 * J + 2;  // J
    uint32_t *crossP = f.crossP.ensure(nP + 1), *crossQ = f.crossQ.ensure(nQ + 1);
    if (nP > 0) k_jclassify<<<grid_for(nP), BLK, 0, s>>>(P, J, nP, crossP, st.slow);
    k_jclassify<<<grid_for(nQ), BLK, 0, s>>>(Q, J, nQ, crossQ, st.slow);
    if (sparse) k_sp_first_plain<<<grid_for(int64_t(J) * WAVE), BLK, 0, s>>>(sp, J, st.firstP);
    else k_first_plain<<<grid_for(int64_t(J) * WAVE), BLK, 0, s>>>(P, J, st.firstP);
    k_first_plain<<<grid_for(int64_t(J) * WAVE), BLK, 0, s>>>(Q, J, st.firstQ);
    SC_HIP(hipGetLastError());
    {
        FillBatch fb(s);
        fb.add(crossP + nP, 1, 0u);
        fb.add(crossQ + nQ, 1, 0u);
        fb.flush();
    }
    uint32_t *ccP = f.ccP.ensure(nP + 1), *ccQ = f.ccQ.ensure(nQ + 1);
    uint32_t *tot = c.d_u32_total.ensure(2);
    scan_exclusive_u32_2(crossP, ccP, nP + 1, tot, crossQ, ccQ, nQ + 1, tot + 1, s);
    // point-in-polygon of each ring's first plain vertex w.r.t. the other ring
    double *qpts = f.qpts.ensure(4 * size_t(J));
    int *qoks = f.qoks.ensure(2 * size_t(J));
    double *pq = qpts, *qq = qpts + 2 * size_t(J);
    int *pok = qoks, *qok = qoks + J;
    k_query_pts<<<grid_for(J), BLK, 0, s>>>(J, P, Q, sparse ? 1 : 0, sp, st.firstP, st.firstQ, pq, pok, qq, qok);
    SC_HIP(hipGetLastError());
    k_jwinding<<<grid_for(nQ), BLK, 0, s>>>(Q, J, nQ, pq, pq + 1, pok, st.windP);
    if (sparse) {
        const int chunks = (c.nb + WCH - 1) / WCH;
        k_sp_wind_base<<<unsigned(int64_t(J) * chunks), BLK, 0, s>>>(J, c.nb, c.bpoly, qq, qok, st.windQ);
        if (nP > 0) k_sp_wind_fix<<<grid_for(nP), BLK, 0, s>>>(int(nP), J, sp, qq, qok, st.windQ);
    } else {
        k_jwinding<<<grid_for(nP), BLK, 0, s>>>(P, J, nP, qq, qq + 1, qok, st.windQ);
    }
    k_jdecide<<<grid_for(J + 1), BLK, 0, s>>>(st, ccP, ccQ);
    SC_HIP(hipGetLastError());
    uint32_t htot[2];
    {
        ReadBack rb(c); // with the post stage's job flags, when still on the device
        rb.add(htot, static_cast<const uint32_t *>(tot), 2);
        const bool fl = c.flags_pending;
        if (fl) rb.add(js.flags.data(), static_cast<const int *>(c.flags.p), size_t(J));
        rb.run();
        if (fl) {
            c.flags_pending = false;
            int any = 0;
            for (int v : js.flags) any |= v;
            c.undefined = any;
        }
    }
    const int mP = int(htot[0]), mQ = int(htot[1]);
    const int M = mP + mQ;
    // entry/exit status, split pairs and pieces of the crossing jobs
    int *enexP = f.enexP.ensure(nP), *enexQ = f.enexQ.ensure(nQ);
    uint8_t *splitP = f.splitP.ensure(nP), *splitQ = f.splitQ.ensure(nQ);
    PieceArgs pa;
    if (M > 0) {
        if (sparse)
            k_sp_enex<<<grid_for(nP), BLK, 0, s>>>(int(nP), J, sp, c.sp_lab, crossP, ccP, st.firstP, st.windP,
                                                   st.kind, enexP, splitP);
        else
            k_jenex<<<grid_for(nP), BLK, 0, s>>>(P, J, nP, crossP, ccP, st.firstP, st.windP, st.kind, enexP,
                                                 splitP);
        k_jenex<<<grid_for(nQ), BLK, 0, s>>>(Q, J, nQ, crossQ, ccQ, st.firstQ, st.windQ, st.kind, enexQ, splitQ);
        k_jsplit_pairs<<<grid_for(nP), BLK, 0, s>>>(P, J, nP, splitP, splitQ, st.slow);
        int *cposP = f.cposP.ensure(mP + 1), *cposQ = f.cposQ.ensure(mQ + 1);
        k_compact_cross<<<grid_for(nP), BLK, 0, s>>>(nP, crossP, ccP, cposP);
        k_compact_cross<<<grid_for(nQ), BLK, 0, s>>>(nQ, crossQ, ccQ, cposQ);
        pa.J = J;
        pa.P = P;
        pa.Q = Q;
        pa.PAB = AB;
        pa.sparse = sparse ? 1 : 0;
        pa.sp = sp;
        pa.cposP = cposP;
        pa.cposQ = cposQ;
        pa.crossP = crossP;
        pa.crossQ = crossQ;
        pa.ccountP = ccP;
        pa.ccountQ = ccQ;
        pa.enexP = enexP;
        pa.enexQ = enexQ;
        pa.cP0 = st.cP0;
        pa.cQ0 = st.cQ0;
        pa.kind = st.kind;
        pa.mP = mP;
        pa.mQ = mQ;
        pa.ppos = f.ppos.ensure(M);
        pa.plen = f.plen.ensure(M);
        pa.pdir = f.pdir.ensure(M);
        pa.pj = f.pj.ensure(M);
        pa.succ = f.succ.ensure(M);
        pa.val = f.val.ensure(M);
        pa.slow = st.slow;
        k_jpieces<<<grid_for(M), BLK, 0, s>>>(pa);
        k_jredecide<<<grid_for(J), BLK, 0, s>>>(J, st.slow, st.kind);
        k_jmask<<<grid_for(M), BLK, 0, s>>>(M, pa.pj, st.kind, pa.succ, pa.val);
        SC_HIP(hipGetLastError());
    }
    // routes and crossing counts on the host
    std::vector<int> hk(3 * size_t(J));
    d2h_sync(c, hk.data(), st.kind, 3 * int64_t(J));
    const int *hkind = hk.data(), *hmP = hk.data() + J, *hmQ = hk.data() + 2 * J;
    js.route.assign(hkind, hkind + J);
    js.reason.resize(size_t(J));
    SC_HIP(hipMemcpyAsync(js.reason.data(), st.slow, sizeof(int) * size_t(J), hipMemcpyDeviceToHost, s));
    int maxm = 0, maxmP = 0; // crossings of the largest traversed job, and its P crossings
    std::vector<int> seq_jobs;
    bool any_nocross = false;
    for (int j = 0; j < J; ++j) {
        if (hkind[j] == K_CROSS) {
            maxm = std::max(maxm, hmP[j] + hmQ[j]);
            maxmP = std::max(maxmP, hmP[j]);
        }
        else if (hkind[j] == K_SEQ) seq_jobs.push_back(j);
        else any_nocross = true;
    }
    // sequential route
    const int nseq = int(seq_jobs.size());
    std::vector<SeqJob> sj(static_cast<size_t>(nseq));
    if (nseq > 0) {
        std::vector<int> hsidx(static_cast<size_t>(J), -1);
        int64_t mb = 0;
        std::vector<int64_t> moff(static_cast<size_t>(nseq)), segoff(2 * size_t(nseq) + 1, 0);
        for (int q = 0; q < nseq; ++q) {
            const int j = seq_jobs[q];
            SeqJob &b = sj[q];
            b.j = j;
            b.nP = js.AB[j + 1] - js.AB[j];
            b.nQ = js.AC[j + 1] - js.AC[j];
            // sparse layout: the base ring is materialised densely for this job
            moff[q] = mb;
            b.P0 = sparse ? int(mb) : js.AB[j];
            mb += b.nP;
            b.Q0 = js.AC[j];
            segoff[2 * q + 1] = segoff[2 * q] + b.nP;
            segoff[2 * q + 2] = segoff[2 * q + 1] + b.nQ;
            hsidx[j] = q;
        }
        SeqJob *djobs = f.sjobs.ensure(nseq);
        SC_HIP(hipMemcpyAsync(djobs, sj.data(), sizeof(SeqJob) * nseq, hipMemcpyHostToDevice, s));
        SC_HIP(hipMemcpyAsync(sidx, hsidx.data(), sizeof(int) * J, hipMemcpyHostToDevice, s));
        SeqBatch b;
        b.jobs = djobs;
        b.uni = c.uni;
        b.par = c.seq_par;
        b.px = c.aug_bx; b.py = c.aug_by; b.qx = c.aug_cx; b.qy = c.aug_cy;
        const int *avP = c.av_b, *labP = c.lab_b, *nbrP = c.nbr_b;
        if (sparse) {
            double *mx = f.mx.ensure(mb), *my = f.my.ensure(mb);
            int *mav = f.mav.ensure(mb), *mnbr = f.mnbr.ensure(mb), *mlab = f.mlab.ensure(mb);
            int *djob = f.mjob.ensure(nseq);
            int64_t *dmoff = f.moff.ensure(nseq);
            SC_HIP(hipMemcpyAsync(djob, seq_jobs.data(), sizeof(int) * nseq, hipMemcpyHostToDevice, s));
            SC_HIP(hipMemcpyAsync(dmoff, moff.data(), sizeof(int64_t) * nseq, hipMemcpyHostToDevice, s));
            for (int q0 = 0; q0 < nseq; q0 += 65535) {
                dim3 g(64, unsigned(std::min(65535, nseq - q0)));
                k_sp_materialize<<<g, BLK, 0, s>>>(q0, nseq, djob, dmoff, sp, c.sp_av, c.sp_nbr, c.sp_lab, mx, my,
                                                   mav, mnbr, mlab);
                SC_HIP(hipGetLastError());
            }
            b.px = mx; b.py = my;
            avP = mav; labP = mlab; nbrP = mnbr;
        }
        // ring nodes
        const int nseg = 2 * nseq;
        const int64_t S = segoff[nseg];
        int64_t *dseg = f.segoff.ensure(nseg + 1);
        SC_HIP(hipMemcpyAsync(dseg, segoff.data(), sizeof(int64_t) * (nseg + 1), hipMemcpyHostToDevice, s));
        NodeArgs na;
        na.nseg = nseg;
        na.segoff = dseg;
        na.jobs = djobs;
        na.px = b.px; na.py = b.py; na.qx = b.qx; na.qy = b.qy;
        na.avP = avP; na.labP = labP; na.nbrP = nbrP; na.labQ = c.lab_c; na.nbrQ = c.nbr_c;
        uint32_t *flag = f.nflag.ensure(S + 1), *nidx = f.nidx.ensure(S + 1);
        k_node_flags<<<grid_for(S), BLK, 0, s>>>(na, S, flag);
        SC_HIP(hipGetLastError());
        SC_HIP(hipMemsetAsync(flag + S, 0, sizeof(uint32_t), s));
        scan_exclusive_u32(flag, nidx, S + 1, nullptr, s);
        // node offsets of the segments, gathered on the device and read back at once
        std::vector<int64_t> nodeoff(size_t(nseg) + 1);
        int64_t *dnodeoff = f.nodeoff.ensure(nseg + 1);
        k_seg_node_offsets<<<grid_for(int64_t(nseg) + 1), BLK, 0, s>>>(nseg, dseg, nidx, dnodeoff);
        SC_HIP(hipGetLastError());
        {
            ReadBack rb(c);
            rb.add(nodeoff.data(), static_cast<const int64_t *>(dnodeoff), size_t(nseg) + 1);
            rb.run();
        }
        const int64_t N = nodeoff[nseg];
        int *nslot = f.nslot.ensure(N), *ninter = f.ninter.ensure(N), *nlab = f.nlab.ensure(N);
        int *nnb = f.nnb.ensure(N), *nlen = f.nlen.ensure(N);
        double *nx0 = f.nx0.ensure(N), *ny0 = f.ny0.ensure(N), *nx1 = f.nx1.ensure(N), *ny1 = f.ny1.ensure(N);
        k_node_fill<<<grid_for(S), BLK, 0, s>>>(na, S, flag, nidx, nslot, ninter, nx0, ny0, nlab, nnb);
        k_node_finish<<<grid_for(N), BLK, 0, s>>>(na, N, dnodeoff, flag, nidx, nslot, ninter, nlen, nx1, ny1, nnb);
        SC_HIP(hipGetLastError());
        // scratch per job, sized by its nodes
        int64_t vb = 0, ob = 0, rb = 0;
        for (int q = 0; q < nseq; ++q) {
            SeqJob &j = sj[q];
            j.node0 = int(nodeoff[2 * q]);
            j.nPn = int(nodeoff[2 * q + 1] - nodeoff[2 * q]);
            j.nQn = int(nodeoff[2 * q + 2] - nodeoff[2 * q + 1]);
            j.cap = 2 * (j.nPn + j.nQn) + 64;
            j.out_cap = 2 * (j.nP + j.nQ) + 64;
            j.ring_cap = j.nP + j.nQ + 8;
            j.vbase = vb;
            j.obase = ob;
            j.rbase = rb;
            vb += j.cap;
            ob += j.out_cap;
            rb += j.ring_cap + 1;
        }
        SC_HIP(hipMemcpyAsync(djobs, sj.data(), sizeof(SeqJob) * nseq, hipMemcpyHostToDevice, s));
        b.nx0 = nx0; b.ny0 = ny0; b.nx1 = nx1; b.ny1 = ny1;
        b.nlen = nlen; b.nslot = nslot; b.ninter = ninter; b.nlab = nlab; b.nnb = nnb;
        b.windP = st.windP;
        b.windQ = st.windQ;
        b.vx = f.vx.ensure(vb); b.vy = f.vy.ensure(vb); b.vx1 = f.vx1.ensure(vb); b.vy1 = f.vy1.ensure(vb);
        b.vlen = f.vlen.ensure(vb); b.vslot = f.vslot.ensure(vb);
        b.nx = f.nx.ensure(vb); b.pv = f.pv.ensure(vb); b.nb = f.nb.ensure(vb); b.lab = f.lab.ensure(vb);
        b.enex = f.enex.ensure(vb); b.vsrc = f.vsrc.ensure(vb); b.cand = f.cand.ensure(vb);
        b.seen = f.seen.ensure(2 * vb);
        b.inter = f.inter.ensure(vb);
        b.rec = f.srec.ensure(vb);
        b.pstep = f.spstep.ensure(vb);
        b.emit = f.semit.ensure(2 * vb);
        b.ox = f.sox.ensure(ob); b.oy = f.soy.ensure(ob); b.osrc = f.sosrc.ensure(ob);
        b.ring_off = f.sroff.ensure(rb);
        b.result = f.sres.ensure(3 * int64_t(nseq));
        // the jobs' working rings and zeroed walk state, spread over the device
        {
            FillBatch fb(s);
            fb.add(b.seen, 2 * vb, 0u);
            fb.flush();
        }
        k_seq_init<<<grid_for(N), BLK, 0, s>>>(b, dnodeoff, nseq, N);
        SC_HIP(hipGetLastError());
        // the sequential jobs run beside the crossing jobs' labelling
        c.ensure_side_stream();
        SC_HIP(hipEventRecord(c.ev_fork, s));
        SC_HIP(hipStreamWaitEvent(c.stream2, c.ev_fork, 0));
        k_sequential<<<nseq, SEQ_BLK, 0, c.stream2>>>(b);
        SC_HIP(hipGetLastError());
        SC_HIP(hipEventRecord(c.ev_join, c.stream2));
    }
    // cycles of the crossing jobs: minimum and distance to it by doubling
    uint32_t *cmin = nullptr, *dist = nullptr, *inode = nullptr;
    uint32_t *base = f.tot.ensure(int64_t(mP) + 1), *ridx = f.ringidx.ensure(int64_t(mP) + 1);
    if (maxm > 0) {
        uint4 *rec = f.jrec.ensure(mP > 0 ? mP : 1), *rec2 = f.jrec2.ensure(mP > 0 ? mP : 1);
        if (mP > 0) {
            k_jump_init<<<grid_for(mP), BLK, 0, s>>>(mP, M, pa.val, pa.succ, pa.plen, rec);
            // a cycle of a traversed job holds at most maxmP P crossings
            const int rounds = ilog2_exact(maxmP);
            uint32_t *moved = f.jmoved.ensure(size_t(rounds > 0 ? rounds : 1));
            {
                FillBatch fb(s);
                fb.add(moved, rounds, 0u);
                fb.flush();
            }
            for (int r = 0; r < rounds; ++r) {
                k_jump_rank<<<grid_for(mP), BLK, 0, s>>>(mP, rec, rec2, moved, r);
                std::swap(rec, rec2);
            }
        }
        cmin = f.val2.ensure(M);
        inode = f.inode.ensure(M);
        dist = f.dist.ensure(M);
        k_rank_final<<<grid_for(M), BLK, 0, s>>>(M, mP, rec, pa.succ, pa.pj, AB, sparse ? 1 : 0, sp, ccP, pa.plen, cmin,
                                                 inode, dist);
        SC_HIP(hipGetLastError());
        uint32_t *isstart = f.isstart.ensure(int64_t(mP) + 1);
        k_start_totals<<<grid_for(mP), BLK, 0, s>>>(mP, pa.cposP, pa.pj, AB, sparse ? c.sp_slot.p : nullptr, cmin,
                                                    dist, base, isstart);
        {
            FillBatch fb(s);
            fb.add(base + mP, 1, 0u);
            fb.add(isstart + mP, 1, 0u);
            fb.flush();
        }
        SC_HIP(hipGetLastError());
        scan_exclusive_u32_2(base, base, int64_t(mP) + 1, nullptr, isstart, ridx, int64_t(mP) + 1, nullptr, s);
    }
    // per-job output sizes before clean-up, and their offsets
    uint32_t *jv = f.jv.ensure(J + 1), *jr = f.jr.ensure(J + 1);
    CountArgs ca;
    ca.J = J;
    ca.AB = AB;
    ca.AC = AC;
    ca.kind = st.kind;
    ca.windP = st.windP;
    ca.windQ = st.windQ;
    ca.cP0 = st.cP0;
    ca.sidx = sidx;
    ca.seq_res = nseq > 0 ? f.sres.p : nullptr;
    ca.base = base;
    ca.ridx = ridx;
    ca.jv = jv;
    ca.jr = jr;
    if (nseq > 0) SC_HIP(hipStreamWaitEvent(s, c.ev_join, 0));
    k_jcounts<<<grid_for(J), BLK, 0, s>>>(ca);
    SC_HIP(hipGetLastError());
    {
        FillBatch fb(s);
        fb.add(jv + J, 1, 0u);
        fb.add(jr + J, 1, 0u);
        fb.flush();
    }
    scan_exclusive_u32_2(jv, jv, int64_t(J) + 1, tot, jr, jr, int64_t(J) + 1, tot + 1, s);
    d2h_sync(c, htot, tot, 2);
    const int64_t NV = htot[0];
    const int NR = int(htot[1]);
    f.ox.ensure(NV + 1);
    f.oy.ensure(NV + 1);
    f.osrc.ensure(NV + 1);
    int *off = f.off.ensure(int64_t(NR) + 1);
    if (maxm > 0 && NV > 0) {
        k_jring_offsets<<<grid_for(mP), BLK, 0, s>>>(mP, f.isstart, ridx, base, pa.pj, st.cP0, jv, jr, off);
        uint32_t *head = f.head.ensure(NV + 1), *piece_at = f.piece_at.ensure(NV + 1);
        {
            FillBatch fb(s);
            fb.add(head, NV + 1, 0u);
            fb.flush();
        }
        k_jheads<<<grid_for(M), BLK, 0, s>>>(M, inode, dist, base, pa.pj, st.cP0, jv, head, piece_at);
        SC_HIP(hipGetLastError());
        scan_inclusive_max_u32(head, head, NV, s);
        EmitArgs ea;
        ea.J = J;
        ea.mP = mP;
        ea.PAB = AB;
        ea.sparse = sparse ? 1 : 0;
        ea.sp = sp;
        ea.JV = jv;
        ea.kind = st.kind;
        ea.ppos = pa.ppos;
        ea.pdir = pa.pdir;
        ea.pj = pa.pj;
        ea.head = head;
        ea.piece_at = piece_at;
        ea.P = P;
        ea.Q = Q;
        ea.ox = f.ox;
        ea.oy = f.oy;
        ea.osrc = f.osrc;
        k_jemit<<<grid_for(NV), BLK, 0, s>>>(ea, NV);
        SC_HIP(hipGetLastError());
    }
    if (any_nocross) {
        if (sparse)
            k_sp_nocross_p<<<grid_for(NV), BLK, 0, s>>>(NV, J, jv, st.kind, st.windP, sp, f.ox, f.oy, f.osrc);
        else
            k_jcopy_nocross<<<grid_for(nP), BLK, 0, s>>>(P, true, J, nP, st.kind, st.windP, st.windQ, AB, jv, f.ox,
                                                         f.oy, f.osrc);
        k_jcopy_nocross<<<grid_for(nQ), BLK, 0, s>>>(Q, false, J, nQ, st.kind, st.windP, st.windQ, AB, jv, f.ox,
                                                     f.oy, f.osrc);
        k_jnocross_off<<<grid_for(J), BLK, 0, s>>>(J, st.kind, st.windP, st.windQ, AB, jv, jr, off);
        SC_HIP(hipGetLastError());
    }
    std::vector<int> hres(3 * size_t(nseq));
    if (nseq > 0) {
        k_jseq_copy<<<nseq, BLK, 0, s>>>(f.sjobs, f.sres, jv, jr, f.sox, f.soy, f.sosrc, f.sroff, f.ox, f.oy,
                                          f.osrc, off);
        SC_HIP(hipGetLastError());
    }
    {
        FillBatch fb(s);
        fb.add(off + NR, 1, uint32_t(int(NV)));
        fb.flush();
    }
    if (nseq > 0) {
        ReadBack rb(c);
        rb.add(hres.data(), static_cast<const int *>(f.sres.p), 3 * size_t(nseq));
        rb.run();
    }
    for (int q = 0; q < nseq; ++q) {
        js.flags[sj[q].j] |= hres[3 * q + 2];
        c.undefined |= hres[3 * q + 2];
    }
    std::vector<uint32_t> job_ring;
    cleanup_rings(c, f, NR, NV, J, jr, job_ring);
    js.ring_off.resize(size_t(J) + 1);
    for (int j = 0; j <= J; ++j) js.ring_off[j] = int(job_ring[j]);
}

void clip_jobs(Ctx &c, int k0, int k1) {
    SC_REQUIRE(k0 >= 1 && k1 <= c.K && k0 <= k1, SC_EINVAL, "overlay index out of range");
    StageTimes t;
    // independent small pairs: one persistent launch (sc_small.hip); a batch
    // it does not cover runs on the staged path below
    if (small_batch_eligible(c, k0, k1)) {
        for (int i = 0; i <= 6; ++i) c.mark(i);
        if (clip_small_batch(c, k0, k1)) {
            c.mark(7);
            if (c.timing) {
                SC_HIP(hipEventSynchronize(c.ev[7]));
                t.label = c.elapsed(6, 7);
                c.times = t;
            }
            return;
        }
    }
    const bool tree = c.T == 0, cand = tree || int(c.pair_off.size()) != c.K + 1;
    c.mark(0);
    if (tree) {
        build_edges(c);
        build_elementary(c);
    }
    c.mark(1);
    if (tree) build_node_intervals(c);
    c.mark(2);
    if (tree) build_leaf_end_entries(c);
    c.mark(3);
    if (tree) build_cover_lists(c);
    c.mark(4);
    if (cand) find_candidates(c);
    c.mark(5);
    post_jobs(c, k0, k1);
    c.mark(6);
    label_and_traverse(c);
    c.mark(7);
    if (c.timing) {
        SC_HIP(hipEventSynchronize(c.ev[7]));
        t.elementary = c.elapsed(0, 1);
        t.tree = c.elapsed(1, 2);
        t.endlist = c.elapsed(2, 3);
        t.cover = c.elapsed(3, 4);
        t.candidates = c.elapsed(4, 5);
        t.stab_kernel = cand ? c.elapsed(8, 9) : 0.0f;
        t.post = c.elapsed(5, 6);
        t.label = c.elapsed(6, 7);
        c.times = t;
    }
}

void clip_all(Ctx &c, int k) { clip_jobs(c, k, k); }

} // namespace sc
