// sc_foster.hip -- Foster/Greiner-Hormann labelling, traversal and clean-up
// (rows A16-A17) on the GPU.
//
// Reference: copyToLinkedList (src/clip.cpp:606-695), labelIntersections2,
// createResult2, cleanUpResult2 (foster/polyclip.h:337-787), intersection
// case (UNION = false).  The augmented rings built by segclip_intersections
// are the linked lists of copyToLinkedList laid out in ring order: slot i of
// the base ring P is followed by slot i+1 (mod n), likewise for the overlay
// ring Q; a slot is an intersection vertex when its alpha (P) or neighbour
// (Q) is set.
//
// Fast path (every label CROSSING or BOUNCING, a plain vertex on both rings,
// crossings alternate, no interior bouncing pair): the entry/exit status of
// a crossing vertex is the parity of the crossings met since the first plain
// vertex (one prefix count); a traversal piece runs from a crossing vertex to
// the next crossing vertex in the walking direction, then jumps to the
// neighbour of that vertex on the other ring.  Pieces form cycles; a cycle's
// output component is started at its base-ring vertex that comes first in
// base-ring order (pointer jumping for the minimum), and the pieces of a
// component are laid out by list ranking.  Everything else -- intersection
// chains (LEFT_ON ... ON_RIGHT), split bouncing pairs, rings made only of
// intersections, inconsistent labels -- runs the reference's sequential
// algorithm in one device thread.
//
// Clean-up: the reference removes, in one pass from the root, every vertex
// nearly collinear (|A| < 1e-9) with the last kept vertex and its successor,
// after first dropping collinear roots.  The decisions are computed as the
// unique fixed point of keep[i] = |A(last kept before i, i, i+1)| >= eps,
// iterated in parallel from the guess "predecessor kept"; the fixed point
// equals the sequential result because it is determined left to right.
#include "sc_ctx.hpp"
#include "sc_fp64.hpp"

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

// reasons for taking the sequential path
enum : int { SLOW_LABEL = 1, SLOW_NOPLAIN = 2, SLOW_ODD = 4, SLOW_SPLIT = 8, SLOW_LINK = 16 };

// status words
enum : int { ST_SLOW = 0, ST_PLAIN_P = 1, ST_PLAIN_Q = 2, ST_WIND_P = 3, ST_WIND_Q = 4, ST_CHANGED = 5, ST_N = 8 };

struct Ring {
    int n;
    const double *x, *y;
    const int *lab, *nbr, *av; // av null: intersection iff nbr set
};

__device__ inline bool inter_of(const Ring &r, int i) { return r.av ? r.av[i] != -100 : r.nbr[i] != -100; }

__global__ __launch_bounds__(BLK) void k_classify(Ring r, int side, int n_other, uint32_t *cross, int *st) {
    __shared__ int s_min;
    if (threadIdx.x == 0) s_min = INT_MAX;
    __syncthreads();
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < r.n) {
        bool in = inter_of(r, i);
        int lab = r.lab[i];
        if (in) {
            if (lab != L_CROSSING && lab != L_BOUNCING) atomicOr(&st[ST_SLOW], SLOW_LABEL);
            int nb = r.nbr[i];
            if (nb < 1 || nb > n_other) atomicOr(&st[ST_SLOW], SLOW_LINK);
        } else {
            atomicMin(&s_min, i);
        }
        cross[i] = (in && lab == L_CROSSING) ? 1u : 0u;
    }
    __syncthreads();
    // first plain vertex: one global atomic per block
    if (threadIdx.x == 0 && s_min != INT_MAX) atomicMin(&st[side == 0 ? ST_PLAIN_P : ST_PLAIN_Q], s_min);
}

// winding test of polygon::pointInPoly (polygon.h:435-457): the query point is
// the vertex at index st[qslot] of ring q; the sum is accumulated in st[out]
__global__ void k_winding(Ring r, Ring q, const int *st_idx, int *out) {
    __shared__ int s_sum[BLK / WAVE];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int qi = *st_idx;
    int w = 0;
    if (i < r.n && qi >= 0 && qi < q.n) {
        double rx = q.x[qi], ry = q.y[qi];
        int j = i + 1 == r.n ? 0 : i + 1;
        double p0x = r.x[i], p0y = r.y[i], p1x = r.x[j], p1y = r.y[j];
        if ((p0y < ry) != (p1y < ry)) {
            int s = 2 * (p1y > p0y) - 1;
            if (p0x >= rx) {
                if (p1x > rx) w = s;
                else if ((areaA(p0x, p0y, p1x, p1y, rx, ry) > 0) == (p1y > p0y)) w = s;
            } else if (p1x > rx) {
                if ((areaA(p0x, p0y, p1x, p1y, rx, ry) > 0) == (p1y > p0y)) w = s;
            }
        }
    }
    // wave then block reduction
    for (int d = WAVE / 2; d > 0; d >>= 1) w += __shfl_down(w, d, WAVE);
    if ((threadIdx.x & (WAVE - 1)) == 0) s_sum[threadIdx.x / WAVE] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < BLK / WAVE; ++k) t += s_sum[k];
        if (t) atomicAdd(out, t);
    }
}

// entry/exit status of every crossing vertex from the parity of crossings
// since the first plain vertex V; bouncing vertices met with status EXIT are
// split candidates (labelIntersections2 step 4, polyclip.h:551-635)
__global__ void k_enex(Ring r, const uint32_t *cross, const uint32_t *ccount, uint32_t total, int V, int st0,
                       int *enex, uint8_t *split) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.n) return;
    uint32_t c = ccount[i] - ccount[V] + (i < V ? total : 0u);
    int st = st0 ^ int(c & 1u);
    enex[i] = cross[i] ? st : NEITHER;
    split[i] = (inter_of(r, i) && r.lab[i] == L_BOUNCING && st == EXIT) ? 1 : 0;
}

__global__ void k_split_pairs(int n, const uint8_t *splitP, const int *nbrP, const uint8_t *splitQ, int *st) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !splitP[i]) return;
    int q = nbrP[i] - 1;
    if (splitQ[q]) atomicOr(&st[ST_SLOW], SLOW_SPLIT);
}

__global__ void k_compact_cross(int n, const uint32_t *cross, const uint32_t *ccount, int *cpos) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && cross[i]) cpos[ccount[i]] = i;
}

struct PieceArgs {
    Ring P, Q;
    const int *cposP, *cposQ;
    const uint32_t *crossP, *crossQ, *ccountP, *ccountQ;
    const int *enexP, *enexQ;
    int mP, mQ;
    int *ppos, *plen, *pdir;
    uint32_t *succ, *val;
    int *st;
};

__global__ void k_pieces(PieceArgs a) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int M = a.mP + a.mQ;
    if (x >= M) return;
    bool onP = x < a.mP;
    const Ring &R = onP ? a.P : a.Q;
    const int *cpos = onP ? a.cposP : a.cposQ;
    const int *enex = onP ? a.enexP : a.enexQ;
    int m = onP ? a.mP : a.mQ;
    int j = onP ? x : x - a.mP;
    int pos = cpos[j];
    int e = enex[pos];
    int d = e == ENTRY ? 1 : -1;
    int jw = j + d;
    if (jw < 0) jw += m;
    if (jw >= m) jw -= m;
    int posW = cpos[jw];
    if (enex[posW] != (e ^ 1)) atomicOr(&a.st[ST_SLOW], SLOW_ODD);
    int len = d > 0 ? posW - pos : pos - posW;
    if (len <= 0) len += R.n;
    int w = R.nbr[posW] - 1; // slot on the other ring
    const uint32_t *ocross = onP ? a.crossQ : a.crossP;
    const uint32_t *occ = onP ? a.ccountQ : a.ccountP;
    uint32_t s = 0;
    if (!ocross[w]) atomicOr(&a.st[ST_SLOW], SLOW_LINK);
    else s = (onP ? uint32_t(a.mP) : 0u) + occ[w];
    a.ppos[x] = pos;
    a.plen[x] = len;
    a.pdir[x] = d;
    a.succ[x] = s;
    a.val[x] = onP ? min(2u * uint32_t(pos), 2u * uint32_t(posW) + 1u) : 0xffffffffu;
}

__global__ void k_jump_min(int M, const uint32_t *val, const uint32_t *jmp, uint32_t *val2, uint32_t *jmp2) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= M) return;
    uint32_t j = jmp[x];
    uint32_t v = val[x], w = val[j];
    val2[x] = v < w ? v : w;
    jmp2[x] = jmp[j];
}

constexpr uint32_t NIL = 0xffffffffu;

__global__ void k_rank_init(int M, const uint32_t *cmin, const uint32_t *succ, const uint32_t *ccountP,
                            const int *plen, uint32_t *inode, uint32_t *nxt, uint32_t *dist) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= M) return;
    uint32_t m = cmin[x];
    bool trav = m != 0xffffffffu && (m & 1u) == 0u;
    uint32_t I = trav ? ccountP[m >> 1] : NIL;
    inode[x] = I;
    nxt[x] = trav ? (succ[x] == I ? NIL : succ[x]) : NIL;
    dist[x] = trav ? uint32_t(plen[x]) : 0u;
}

__global__ void k_rank_step(int M, const uint32_t *dist, const uint32_t *nxt, uint32_t *dist2, uint32_t *nxt2) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= M) return;
    uint32_t n = nxt[x];
    if (n == NIL) {
        dist2[x] = dist[x];
        nxt2[x] = NIL;
    } else {
        dist2[x] = dist[x] + dist[n];
        nxt2[x] = nxt[n];
    }
}

// component totals at their starts (base-ring order)
__global__ void k_start_totals(int mP, const int *cposP, const uint32_t *cmin, const uint32_t *dist,
                               uint32_t *tot, uint32_t *isstart) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= mP) return;
    uint32_t m = cmin[j];
    bool s = m != 0xffffffffu && (m & 1u) == 0u && (m >> 1) == uint32_t(cposP[j]);
    tot[j] = s ? dist[j] : 0u;
    isstart[j] = s ? 1u : 0u;
}

__global__ void k_ring_offsets(int mP, const uint32_t *isstart, const uint32_t *ringidx, const uint32_t *base,
                               int *ring_off, uint32_t out_base) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= mP) return;
    if (isstart[j]) ring_off[ringidx[j]] = int(out_base + base[j]);
}

__global__ void k_heads(int M, const uint32_t *inode, const uint32_t *dist, const uint32_t *base, uint32_t out_base,
                        uint32_t *head, uint32_t *piece_at, uint32_t *ostart) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= M) return;
    uint32_t I = inode[x];
    if (I == NIL) return;
    uint32_t total = dist[I];
    uint32_t o = out_base + base[I] + (total - dist[x]);
    head[o] = o;
    piece_at[o] = uint32_t(x);
    ostart[x] = o;
}

__global__ void k_emit(uint32_t out_base, uint32_t n_out, const uint32_t *hs, const uint32_t *piece_at, int mP,
                       const int *ppos, const int *pdir, Ring P, Ring Q, double *ox, double *oy) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x + out_base;
    if (i >= int64_t(out_base) + n_out) return;
    uint32_t s = hs[i];
    uint32_t x = piece_at[s];
    int k = int(i - s) + 1;
    const Ring &R = int(x) < mP ? P : Q;
    int pos = ppos[x] + pdir[x] * k;
    pos %= R.n;
    if (pos < 0) pos += R.n;
    ox[i] = R.x[pos];
    oy[i] = R.y[pos];
}

__global__ void k_copy_ring(Ring r, double *ox, double *oy, uint32_t at) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.n) return;
    ox[at + i] = r.x[i];
    oy[at + i] = r.y[i];
}

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

__global__ void k_keep0(int64_t n, const int *ringid, const int *off, const int *rstart, const double *x,
                        const double *y, uint32_t *keep) {
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
}

__global__ void k_keep_marks(int64_t n, const uint32_t *keep, uint32_t *mark) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) mark[i] = keep[i] ? uint32_t(i) + 1u : 0u;
}

__global__ void k_keep_step(int64_t n, const int *ringid, const int *off, const int *rstart, const double *x,
                            const double *y, const uint32_t *lastk, uint32_t *keep, int *changed) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int r = ringid[i];
    int rs = rstart[r], last = off[r + 1] - 1;
    if (i <= rs) return;
    int pv = int(lastk[i - 1]) - 1; // last kept index before i (>= rs)
    int nx = i == last ? rs : int(i) + 1;
    uint32_t k = collinear(x, y, pv, int(i), nx) ? 0u : 1u;
    if (k != keep[i]) {
        keep[i] = k;
        atomicOr(changed, 1);
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
                              const double *y, double *x2, double *y2) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    x2[kscan[i]] = x[i];
    y2[kscan[i]] = y[i];
}

__global__ void k_ring_off_final(int nr, const uint32_t *cnt, const uint32_t *rscan, const uint32_t *nonempty,
                                 const int *off, const uint32_t *kscan, int *off2) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    if (cnt[r]) off2[nonempty[r]] = int(kscan[off[r]]);
}

// ---------------------------------------------------------------------------
// sequential path: the reference's algorithm on index-linked vertices, one thread
// ---------------------------------------------------------------------------
struct SeqArgs {
    int nP, nQ, cap;
    const double *px, *py, *qx, *qy;
    const int *avP, *labP, *nbrP, *labQ, *nbrQ;
    double *vx, *vy;
    int *nx, *pv, *nb, *lab, *enex;
    uint8_t *inter;
    int *cand; // scratch list, cap entries
    double *ox, *oy;
    int out_cap;
    int *ring_off;
    int ring_cap;
    int *result; // [0] rings, [1] vertices, [2] status bits
};

__device__ int seq_new(SeqArgs &a, int &n, double x, double y) {
    if (n >= a.cap) return -1;
    int v = n++;
    a.vx[v] = x;
    a.vy[v] = y;
    a.nx[v] = a.pv[v] = v;
    a.nb[v] = -1;
    a.lab[v] = L_NONE;
    a.enex[v] = NEITHER;
    a.inter[v] = 0;
    return v;
}

__device__ void seq_insert_after(SeqArgs &a, int V, int curr) {
    int c = a.nx[curr];
    a.nx[a.pv[c]] = V;
    a.pv[V] = a.pv[c];
    a.nx[V] = c;
    a.pv[c] = V;
}

__device__ bool seq_point_in(SeqArgs &a, int root, double rx, double ry) {
    int w = 0, one = root;
    do {
        int two = a.nx[one];
        double p0x = a.vx[one], p0y = a.vy[one], p1x = a.vx[two], p1y = a.vy[two];
        if ((p0y < ry) != (p1y < ry)) {
            if (p0x >= rx) {
                if (p1x > rx) w = w + 2 * (p1y > p0y) - 1;
                else if ((areaA(p0x, p0y, p1x, p1y, rx, ry) > 0) == (p1y > p0y)) w = w + 2 * (p1y > p0y) - 1;
            } else if (p1x > rx) {
                if ((areaA(p0x, p0y, p1x, p1y, rx, ry) > 0) == (p1y > p0y)) w = w + 2 * (p1y > p0y) - 1;
            }
        }
        one = two;
    } while (one != root);
    return (w % 2) != 0;
}

__device__ int seq_toggle(int s) { return s == ENTRY ? EXIT : (s == EXIT ? ENTRY : s); }

__global__ void k_sequential(SeqArgs a) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int n = 0, status = 0;
    // rings in slot order; P = [0,nP), Q = [nP,nP+nQ)
    for (int i = 0; i < a.nP; ++i) {
        int v = seq_new(a, n, a.px[i], a.py[i]);
        a.lab[v] = a.labP[i];
        a.inter[v] = a.avP[i] != -100;
    }
    for (int i = 0; i < a.nQ; ++i) {
        int v = seq_new(a, n, a.qx[i], a.qy[i]);
        a.lab[v] = a.labQ[i];
        a.inter[v] = a.nbrQ[i] != -100;
    }
    for (int i = 0; i < a.nP; ++i) {
        a.nx[i] = i + 1 == a.nP ? 0 : i + 1;
        a.pv[i] = i == 0 ? a.nP - 1 : i - 1;
        if (a.inter[i]) a.nb[i] = a.nP + a.nbrP[i] - 1;
    }
    for (int i = 0; i < a.nQ; ++i) {
        int v = a.nP + i;
        a.nx[v] = a.nP + (i + 1 == a.nQ ? 0 : i + 1);
        a.pv[v] = a.nP + (i == 0 ? a.nQ - 1 : i - 1);
        if (a.inter[v]) a.nb[v] = a.nbrQ[i] - 1;
    }
    const int rootP = 0, rootQ = a.nP;
    // 2) intersection chains
    {
        int I = rootP;
        do {
            if (a.inter[I] && (a.lab[I] == L_LEFT_ON || a.lab[I] == L_RIGHT_ON)) {
                int x = a.lab[I] == L_LEFT_ON ? 0 : 1;
                int X = I, J = I, guard = 0;
                do {
                    a.lab[J] = L_NONE;
                    if (a.nb[J] >= 0) a.lab[a.nb[J]] = L_NONE;
                    J = a.nx[J];
                    if (++guard > n) { status |= UB_TRAVERSAL; goto done; }
                } while (a.lab[J] == L_ON_ON);
                int y = a.lab[J] == L_ON_LEFT ? 0 : 1;
                int ct = x != y ? L_DCROSS : L_DBOUNCE;
                a.lab[X] = ct;
                a.lab[J] = ct;
                if (a.nb[X] >= 0) a.lab[a.nb[X]] = ct;
                if (a.nb[J] >= 0) a.lab[a.nb[J]] = ct;
            }
            I = a.nx[I];
        } while (I != rootP);
    }
    {
        int rings = 0, outn = 0;
        int roots[2] = {rootP, rootQ};
        bool noInter[2] = {false, false}, ident[2] = {false, false};
        // 3.5) rings without crossing vertices
        for (int s = 0; s < 2; ++s) {
            int A = roots[s], B = roots[1 - s];
            bool nocross = true, allon = true;
            int V = A;
            do {
                if (a.inter[V] && (a.lab[V] == L_CROSSING || a.lab[V] == L_DCROSS)) nocross = false;
                if (a.lab[V] != L_ON_ON) allon = false;
                V = a.nx[V];
            } while (V != A);
            if (!nocross) continue;
            noInter[s] = true;
            if (allon) {
                ident[s] = true;
                continue;
            }
            double qx = 0, qy = 0;
            bool found = false;
            V = A;
            do {
                if (!a.inter[V]) {
                    qx = a.vx[V];
                    qy = a.vy[V];
                    found = true;
                    break;
                }
                V = a.nx[V];
            } while (V != A);
            if (!found) {
                V = A;
                do {
                    int nxv = a.nx[V], nbv = a.nb[V];
                    if (nbv >= 0 && a.nb[nxv] != a.pv[nbv] && a.nb[nxv] != a.nx[nbv]) {
                        qx = 0.5 * (a.vx[V] + a.vx[nxv]);
                        qy = 0.5 * (a.vy[V] + a.vy[nxv]);
                        found = true;
                        break;
                    }
                    V = a.nx[V];
                } while (V != A);
            }
            if (!found) { status |= UB_TRAVERSAL; goto done; }
            if (seq_point_in(a, B, qx, qy)) {
                if (rings >= a.ring_cap) { status |= UB_TRAVERSAL; goto done; }
                a.ring_off[rings++] = outn;
                int W = A;
                do {
                    if (outn >= a.out_cap) { status |= UB_TRAVERSAL; goto done; }
                    a.ox[outn] = a.vx[W];
                    a.oy[outn] = a.vy[W];
                    outn++;
                    W = a.nx[W];
                } while (W != A);
            }
        }
        if (ident[0] && ident[1]) {
            int target = a.nb[rootP], V = rootQ;
            bool found = false;
            do {
                if (V == target) found = true;
                V = a.nx[V];
            } while (V != rootQ);
            if (found) {
                if (rings >= a.ring_cap) { status |= UB_TRAVERSAL; goto done; }
                a.ring_off[rings++] = outn;
                int W = rootP;
                do {
                    if (outn >= a.out_cap) { status |= UB_TRAVERSAL; goto done; }
                    a.ox[outn] = a.vx[W];
                    a.oy[outn] = a.vy[W];
                    outn++;
                    W = a.nx[W];
                } while (W != rootP);
            }
        }
        // 4) entry/exit flags; split and crossing candidates in scratch
        int nsplit[2] = {0, 0}, ncross[2] = {0, 0};
        int half = a.cap / 2; // cand[0..half) split lists, [half..cap) crossing lists
        int *splitL[2] = {a.cand, a.cand + half / 2};
        int *crossL[2] = {a.cand + half, a.cand + half + half / 2};
        for (int s = 0; s < 2; ++s) {
            if (noInter[s]) continue;
            int A = roots[s], B = roots[1 - s];
            int V = -1, W = A;
            do {
                if (!a.inter[W]) { V = W; break; }
                W = a.nx[W];
            } while (W != A);
            if (V < 0) {
                W = A;
                do {
                    int nxv = a.nx[W], nbv = a.nb[W];
                    if (nbv >= 0 && a.nb[nxv] != a.pv[nbv] && a.nb[nxv] != a.nx[nbv]) {
                        int T = seq_new(a, n, 0.5 * (a.vx[W] + a.vx[nxv]), 0.5 * (a.vy[W] + a.vy[nxv]));
                        if (T < 0) { status |= UB_TRAVERSAL; goto done; }
                        seq_insert_after(a, T, W);
                        V = T;
                        break;
                    }
                    W = a.nx[W];
                } while (W != A);
            }
            if (V < 0) { status |= UB_TRAVERSAL; goto done; }
            int st = ENTRY;
            if (seq_point_in(a, B, a.vx[V], a.vy[V])) st = seq_toggle(st);
            bool first = true;
            int I = a.nx[V];
            while (I != V) {
                if (a.inter[I]) {
                    int lab = a.lab[I];
                    if (lab == L_CROSSING) {
                        a.enex[I] = st;
                        st = seq_toggle(st);
                    }
                    if (lab == L_BOUNCING && st == EXIT && nsplit[s] < half / 2) splitL[s][nsplit[s]++] = I;
                    if (lab == L_DCROSS) {
                        a.enex[I] = st;
                        if (first) {
                            if (st == EXIT) a.lab[I] = L_CROSSING;
                            first = false;
                        } else {
                            if (st == ENTRY) a.lab[I] = L_CROSSING;
                            first = true;
                            st = seq_toggle(st);
                        }
                    }
                    if (lab == L_DBOUNCE) {
                        a.enex[I] = st;
                        if (first) {
                            if (st == EXIT && ncross[s] < half / 2) crossL[s][ncross[s]++] = I;
                            first = false;
                        } else {
                            if (st == ENTRY && ncross[s] < half / 2) crossL[s][ncross[s]++] = I;
                            first = true;
                        }
                        st = seq_toggle(st);
                    }
                }
                I = a.nx[I];
            }
        }
        // 5) split interior bouncing pairs
        for (int i = 0; i < nsplit[0]; ++i) {
            int IP = splitL[0][i], IQ = a.nb[IP];
            bool found = false;
            for (int j = 0; j < nsplit[1]; ++j)
                if (splitL[1][j] == IQ) found = true;
            if (!found) continue;
            int VP = seq_new(a, n, a.vx[IP], a.vy[IP]);
            int VQ = seq_new(a, n, a.vx[IQ], a.vy[IQ]);
            if (VP < 0 || VQ < 0) { status |= UB_TRAVERSAL; goto done; }
            double sP = areaA(a.vx[a.pv[IP]], a.vy[a.pv[IP]], a.vx[IP], a.vy[IP], a.vx[a.nx[IP]], a.vy[a.nx[IP]]);
            double sQ = areaA(a.vx[a.pv[IQ]], a.vy[a.pv[IQ]], a.vx[IQ], a.vy[IQ], a.vx[a.nx[IQ]], a.vy[a.nx[IQ]]);
            if (sP * sQ > 0) {
                a.nb[IP] = VQ; a.nb[VQ] = IP; a.inter[IP] = a.inter[VQ] = 1;
                a.nb[IQ] = VP; a.nb[VP] = IQ; a.inter[IQ] = a.inter[VP] = 1;
            } else {
                a.nb[VP] = VQ; a.nb[VQ] = VP; a.inter[VP] = a.inter[VQ] = 1;
            }
            seq_insert_after(a, VP, IP);
            seq_insert_after(a, VQ, IQ);
            a.enex[IP] = EXIT; a.enex[VP] = ENTRY; a.enex[IQ] = EXIT; a.enex[VQ] = ENTRY;
            a.lab[IP] = a.lab[VP] = a.lab[IQ] = a.lab[VQ] = L_CROSSING;
        }
        // 6) crossing candidates
        for (int i = 0; i < ncross[0]; ++i) {
            int IP = crossL[0][i], IQ = a.nb[IP];
            for (int j = 0; j < ncross[1]; ++j)
                if (crossL[1][j] == IQ) {
                    a.lab[IP] = L_CROSSING;
                    a.lab[IQ] = L_CROSSING;
                }
        }
        // createResult2
        {
            int root = rootP, I = root;
            bool have = a.inter[I] && a.lab[I] == L_CROSSING;
            if (!have) {
                do { I = a.nx[I]; } while (!(a.inter[I] && a.lab[I] == L_CROSSING) && I != root);
                have = I != root;
            }
            long bound = 4L * n + 16;
            while (have) {
                if (rings >= a.ring_cap) { status |= UB_TRAVERSAL; goto done; }
                a.ring_off[rings++] = outn;
                int V = I;
                a.inter[V] = 0;
                long steps = 0;
                do {
                    int st = seq_toggle(a.enex[V]);
                    do {
                        V = st == EXIT ? a.nx[V] : a.pv[V];
                        a.inter[V] = 0;
                        if (outn >= a.out_cap || ++steps > bound) { status |= UB_TRAVERSAL; goto done; }
                        a.ox[outn] = a.vx[V];
                        a.oy[outn] = a.vy[V];
                        outn++;
                    } while (!(a.enex[V] == st) && V != I);
                    if (V != I) {
                        V = a.nb[V];
                        if (V < 0) { status |= UB_TRAVERSAL; goto done; }
                        a.inter[V] = 0;
                    }
                } while (V != I);
                do { I = a.nx[I]; } while (!(a.inter[I] && a.lab[I] == L_CROSSING) && I != root);
                have = I != root;
            }
        }
        a.ring_off[rings] = outn;
        a.result[0] = rings;
        a.result[1] = outn;
        a.result[2] = status;
        return;
    }
done:
    a.result[0] = 0;
    a.result[1] = 0;
    a.result[2] = status | UB_TRAVERSAL;
    a.ring_off[0] = 0;
}

// scratch owned by this stage
struct FosterScratch {
    DBuf<uint32_t> crossP, crossQ, ccP, ccQ;
    DBuf<int> enexP, enexQ, cposP, cposQ;
    DBuf<uint8_t> splitP, splitQ;
    DBuf<int> ppos, plen, pdir;
    DBuf<uint32_t> succ, val, val2, jmp2, inode, nxt, nxt2, dist, dist2, tot, isstart, ringidx, head, piece_at,
        ostart;
    DBuf<int> st;
    DBuf<double> ox, oy, ox2, oy2;
    DBuf<int> off, off2, rstart, ringid;
    DBuf<uint32_t> keep, mark, kscan, rcnt, rscan, nonempty;
    // sequential path
    DBuf<double> vx, vy;
    DBuf<int> nx, pv, nb, lab, enex, cand, result;
    DBuf<uint8_t> inter;
};

FosterScratch &fscratch() {
    thread_local FosterScratch *p = new FosterScratch();
    return *p;
}

void read_ints(Ctx &c, int *h, const int *d, int n) {
    SC_HIP(hipMemcpyAsync(h, d, sizeof(int) * size_t(n), hipMemcpyDeviceToHost, c.stream));
    SC_HIP(hipStreamSynchronize(c.stream));
}

// clean-up of nr rings laid out at off[0..nr] in (ox, oy); result into Ctx
void cleanup_rings(Ctx &c, FosterScratch &f, int nr, int64_t n) {
    hipStream_t s = c.stream;
    if (nr == 0 || n == 0) {
        c.nrings = 0;
        c.n_out = 0;
        c.d_ring_off.ensure(1);
        SC_HIP(hipMemsetAsync(c.d_ring_off.p, 0, sizeof(int), s));
        return;
    }
    int *rstart = f.rstart.ensure(nr);
    int *ringid = f.ringid.ensure(size_t(n));
    k_root_loop<<<grid_for(nr), BLK, 0, s>>>(nr, f.off, f.ox, f.oy, rstart);
    SC_HIP(hipMemsetAsync(ringid, 0, sizeof(int) * size_t(n), s));
    k_ring_heads<<<grid_for(nr), BLK, 0, s>>>(nr, f.off, reinterpret_cast<uint32_t *>(ringid));
    scan_inclusive_max_u32(reinterpret_cast<uint32_t *>(ringid), reinterpret_cast<uint32_t *>(ringid), n, s);
    uint32_t *keep = f.keep.ensure(size_t(n) + 1);
    uint32_t *mark = f.mark.ensure(size_t(n));
    k_keep0<<<grid_for(n), BLK, 0, s>>>(n, ringid, f.off, rstart, f.ox, f.oy, keep);
    SC_HIP(hipGetLastError());
    int *st = f.st.ensure(ST_N);
    bool converged = false;
    for (int it = 0; it < 48; ++it) {
        k_keep_marks<<<grid_for(n), BLK, 0, s>>>(n, keep, mark);
        scan_inclusive_max_u32(mark, mark, n, s);
        SC_HIP(hipMemsetAsync(st + ST_CHANGED, 0, sizeof(int), s));
        k_keep_step<<<grid_for(n), BLK, 0, s>>>(n, ringid, f.off, rstart, f.ox, f.oy, mark, keep, st + ST_CHANGED);
        SC_HIP(hipGetLastError());
        int ch = 0;
        read_ints(c, &ch, st + ST_CHANGED, 1);
        if (!ch) {
            converged = true;
            break;
        }
    }
    if (!converged) {
        k_cleanup_seq<<<grid_for(nr), BLK, 0, s>>>(nr, f.off, rstart, f.ox, f.oy, keep);
        SC_HIP(hipGetLastError());
    }
    // compaction
    SC_HIP(hipMemsetAsync(keep + n, 0, sizeof(uint32_t), s));
    uint32_t *kscan = f.kscan.ensure(size_t(n) + 1);
    uint32_t *tot = c.d_u32_total.ensure(2);
    scan_exclusive_u32(keep, kscan, n + 1, tot, s);
    uint32_t *rcnt = f.rcnt.ensure(nr + 1), *nonempty = f.nonempty.ensure(nr + 1);
    k_ring_counts<<<grid_for(nr), BLK, 0, s>>>(nr, f.off, kscan, rcnt);
    SC_HIP(hipGetLastError());
    // nonempty flags -> ring index
    uint32_t *rflag = f.rscan.ensure(nr + 1);
    {
        // rflag[r] = rcnt[r] > 0
        std::vector<uint32_t> h(static_cast<size_t>(nr));
        SC_HIP(hipMemcpyAsync(h.data(), rcnt, sizeof(uint32_t) * nr, hipMemcpyDeviceToHost, s));
        SC_HIP(hipStreamSynchronize(s));
        std::vector<uint32_t> idx(size_t(nr) + 1, 0);
        for (int r = 0; r < nr; ++r) idx[r + 1] = idx[r] + (h[r] ? 1u : 0u);
        SC_HIP(hipMemcpyAsync(nonempty, idx.data(), sizeof(uint32_t) * (nr + 1), hipMemcpyHostToDevice, s));
        c.nrings = int(idx[nr]);
        (void)rflag;
    }
    uint32_t nout = read_u32(c, tot);
    c.n_out = nout;
    c.d_ring_x.ensure(nout);
    c.d_ring_y.ensure(nout);
    int *off2 = c.d_ring_off.ensure(size_t(c.nrings) + 1);
    k_compact_out<<<grid_for(n), BLK, 0, s>>>(n, keep, kscan, f.ox, f.oy, c.d_ring_x, c.d_ring_y);
    k_ring_off_final<<<grid_for(nr), BLK, 0, s>>>(nr, rcnt, nullptr, nonempty, f.off, kscan, off2);
    SC_HIP(hipGetLastError());
    int last = int(nout);
    SC_HIP(hipMemcpyAsync(off2 + c.nrings, &last, sizeof(int), hipMemcpyHostToDevice, s));
    SC_HIP(hipStreamSynchronize(s));
}

Ring ring_P(Ctx &c) { return Ring{c.nbaug, c.aug_bx, c.aug_by, c.lab_b, c.nbr_b, c.av_b}; }
Ring ring_Q(Ctx &c) { return Ring{c.ncaug, c.aug_cx, c.aug_cy, c.lab_c, c.nbr_c, nullptr}; }

void sequential_path(Ctx &c, FosterScratch &f) {
    hipStream_t s = c.stream;
    int nP = c.nbaug, nQ = c.ncaug;
    int cap = 2 * (nP + nQ) + 64;
    int out_cap = 2 * (nP + nQ) + 64;
    int ring_cap = nP + nQ + 8;
    SeqArgs a;
    a.nP = nP; a.nQ = nQ; a.cap = cap;
    a.px = c.aug_bx; a.py = c.aug_by; a.qx = c.aug_cx; a.qy = c.aug_cy;
    a.avP = c.av_b; a.labP = c.lab_b; a.nbrP = c.nbr_b; a.labQ = c.lab_c; a.nbrQ = c.nbr_c;
    a.vx = f.vx.ensure(cap); a.vy = f.vy.ensure(cap);
    a.nx = f.nx.ensure(cap); a.pv = f.pv.ensure(cap); a.nb = f.nb.ensure(cap); a.lab = f.lab.ensure(cap);
    a.enex = f.enex.ensure(cap); a.inter = f.inter.ensure(cap); a.cand = f.cand.ensure(cap);
    a.ox = f.ox.ensure(out_cap); a.oy = f.oy.ensure(out_cap); a.out_cap = out_cap;
    a.ring_off = f.off.ensure(ring_cap + 1); a.ring_cap = ring_cap;
    a.result = f.result.ensure(4);
    k_sequential<<<1, 1, 0, s>>>(a);
    SC_HIP(hipGetLastError());
    int res[3];
    read_ints(c, res, a.result, 3);
    c.undefined |= res[2];
    cleanup_rings(c, f, res[0], res[1]);
}

} // namespace

void label_and_traverse(Ctx &c) {
    SC_REQUIRE(c.cur_k >= 1, SC_ESTATE, "segclip_intersections must run first");
    hipStream_t s = c.stream;
    FosterScratch &f = fscratch();
    Ring P = ring_P(c), Q = ring_Q(c);
    const int nP = P.n, nQ = Q.n;
    int *st = f.st.ensure(ST_N);
    int init[ST_N] = {0, INT_MAX, INT_MAX, 0, 0, 0, 0, 0};
    SC_HIP(hipMemcpyAsync(st, init, sizeof(init), hipMemcpyHostToDevice, s));
    uint32_t *crossP = f.crossP.ensure(nP + 1), *crossQ = f.crossQ.ensure(nQ + 1);
    k_classify<<<grid_for(nP), BLK, 0, s>>>(P, 0, nQ, crossP, st);
    k_classify<<<grid_for(nQ), BLK, 0, s>>>(Q, 1, nP, crossQ, st);
    SC_HIP(hipGetLastError());
    SC_HIP(hipMemsetAsync(crossP + nP, 0, sizeof(uint32_t), s));
    SC_HIP(hipMemsetAsync(crossQ + nQ, 0, sizeof(uint32_t), s));
    uint32_t *ccP = f.ccP.ensure(nP + 1), *ccQ = f.ccQ.ensure(nQ + 1);
    uint32_t *tot = c.d_u32_total.ensure(2);
    scan_exclusive_u32(crossP, ccP, nP + 1, tot, s);
    scan_exclusive_u32(crossQ, ccQ, nQ + 1, tot + 1, s);
    // point-in-polygon of each ring's first plain vertex w.r.t. the other ring
    k_winding<<<grid_for(nQ), BLK, 0, s>>>(Q, P, st + ST_PLAIN_P, st + ST_WIND_P);
    k_winding<<<grid_for(nP), BLK, 0, s>>>(P, Q, st + ST_PLAIN_Q, st + ST_WIND_Q);
    SC_HIP(hipGetLastError());
    int hst[ST_N];
    uint32_t htot[2];
    SC_HIP(hipMemcpyAsync(htot, tot, sizeof(htot), hipMemcpyDeviceToHost, s));
    read_ints(c, hst, st, ST_N);
    const uint32_t mP = htot[0], mQ = htot[1];
    int slow = hst[ST_SLOW];
    const bool plainP = hst[ST_PLAIN_P] != INT_MAX, plainQ = hst[ST_PLAIN_Q] != INT_MAX;
    if (!plainP || !plainQ) slow |= SLOW_NOPLAIN;
    if (mP != mQ || (mP & 1u)) slow |= SLOW_ODD;
    if (slow) {
        sequential_path(c, f);
        return;
    }
    const bool insideP = (hst[ST_WIND_P] % 2) != 0; // P's plain vertex inside Q
    const bool insideQ = (hst[ST_WIND_Q] % 2) != 0;
    if (mP == 0) {
        // no crossing vertex on either ring: each ring is output if it lies inside the other
        int nr = 0;
        int64_t n = 0;
        int hoff[3] = {0, 0, 0};
        f.ox.ensure(size_t(nP) + nQ);
        f.oy.ensure(size_t(nP) + nQ);
        if (insideP) {
            k_copy_ring<<<grid_for(nP), BLK, 0, s>>>(P, f.ox, f.oy, uint32_t(n));
            n += nP;
            hoff[++nr] = int(n);
        }
        if (insideQ) {
            k_copy_ring<<<grid_for(nQ), BLK, 0, s>>>(Q, f.ox, f.oy, uint32_t(n));
            n += nQ;
            hoff[++nr] = int(n);
        }
        SC_HIP(hipGetLastError());
        int *off = f.off.ensure(3);
        SC_HIP(hipMemcpyAsync(off, hoff, sizeof(int) * 3, hipMemcpyHostToDevice, s));
        cleanup_rings(c, f, nr, n);
        return;
    }
    // entry/exit status
    int *enexP = f.enexP.ensure(nP), *enexQ = f.enexQ.ensure(nQ);
    uint8_t *splitP = f.splitP.ensure(nP), *splitQ = f.splitQ.ensure(nQ);
    k_enex<<<grid_for(nP), BLK, 0, s>>>(P, crossP, ccP, mP, hst[ST_PLAIN_P], insideP ? EXIT : ENTRY, enexP, splitP);
    k_enex<<<grid_for(nQ), BLK, 0, s>>>(Q, crossQ, ccQ, mQ, hst[ST_PLAIN_Q], insideQ ? EXIT : ENTRY, enexQ, splitQ);
    k_split_pairs<<<grid_for(nP), BLK, 0, s>>>(nP, splitP, c.nbr_b, splitQ, st);
    // compact crossing vertices and build the pieces
    int *cposP = f.cposP.ensure(mP), *cposQ = f.cposQ.ensure(mQ);
    k_compact_cross<<<grid_for(nP), BLK, 0, s>>>(nP, crossP, ccP, cposP);
    k_compact_cross<<<grid_for(nQ), BLK, 0, s>>>(nQ, crossQ, ccQ, cposQ);
    const int M = int(mP + mQ);
    PieceArgs pa;
    pa.P = P; pa.Q = Q; pa.cposP = cposP; pa.cposQ = cposQ; pa.crossP = crossP; pa.crossQ = crossQ;
    pa.ccountP = ccP; pa.ccountQ = ccQ; pa.enexP = enexP; pa.enexQ = enexQ; pa.mP = int(mP); pa.mQ = int(mQ);
    pa.ppos = f.ppos.ensure(M); pa.plen = f.plen.ensure(M); pa.pdir = f.pdir.ensure(M);
    pa.succ = f.succ.ensure(M); pa.val = f.val.ensure(M); pa.st = st;
    k_pieces<<<grid_for(M), BLK, 0, s>>>(pa);
    SC_HIP(hipGetLastError());
    read_ints(c, hst, st, 1);
    if (hst[ST_SLOW]) {
        sequential_path(c, f);
        return;
    }
    // cycle minima by pointer jumping
    uint32_t *val = pa.val, *val2 = f.val2.ensure(M), *jmp = f.nxt.ensure(M), *jmp2 = f.jmp2.ensure(M);
    SC_HIP(hipMemcpyAsync(jmp, pa.succ, sizeof(uint32_t) * M, hipMemcpyDeviceToDevice, s));
    int rounds = ilog2_exact(M) + 1;
    for (int r = 0; r < rounds; ++r) {
        k_jump_min<<<grid_for(M), BLK, 0, s>>>(M, val, jmp, val2, jmp2);
        std::swap(val, val2);
        std::swap(jmp, jmp2);
    }
    SC_HIP(hipGetLastError());
    // list ranking of the traversed cycles, broken at their start vertex
    uint32_t *inode = f.inode.ensure(M), *nxt = jmp2, *nxt2 = jmp, *dist = f.dist.ensure(M),
             *dist2 = f.dist2.ensure(M);
    k_rank_init<<<grid_for(M), BLK, 0, s>>>(M, val, pa.succ, ccP, pa.plen, inode, nxt, dist);
    for (int r = 0; r < rounds; ++r) {
        k_rank_step<<<grid_for(M), BLK, 0, s>>>(M, dist, nxt, dist2, nxt2);
        std::swap(dist, dist2);
        std::swap(nxt, nxt2);
    }
    SC_HIP(hipGetLastError());
    // components in base-ring order
    uint32_t *ctot = f.tot.ensure(mP + 1), *isstart = f.isstart.ensure(mP + 1), *ridx = f.ringidx.ensure(mP + 1);
    k_start_totals<<<grid_for(mP), BLK, 0, s>>>(int(mP), cposP, val, dist, ctot, isstart);
    SC_HIP(hipMemsetAsync(ctot + mP, 0, sizeof(uint32_t), s));
    SC_HIP(hipMemsetAsync(isstart + mP, 0, sizeof(uint32_t), s));
    SC_HIP(hipGetLastError());
    uint32_t *base = ctot; // in place
    scan_exclusive_u32(ctot, base, int64_t(mP) + 1, tot, s);
    scan_exclusive_u32(isstart, ridx, int64_t(mP) + 1, tot + 1, s);
    SC_HIP(hipMemcpyAsync(htot, tot, sizeof(htot), hipMemcpyDeviceToHost, s));
    SC_HIP(hipStreamSynchronize(s));
    const uint32_t n_out = htot[0];
    const int nr = int(htot[1]);
    int *off = f.off.ensure(size_t(nr) + 1);
    k_ring_offsets<<<grid_for(mP), BLK, 0, s>>>(int(mP), isstart, ridx, base, off, 0u);
    int last = int(n_out);
    SC_HIP(hipMemcpyAsync(off + nr, &last, sizeof(int), hipMemcpyHostToDevice, s));
    uint32_t *head = f.head.ensure(size_t(n_out) + 1), *piece_at = f.piece_at.ensure(size_t(n_out) + 1),
             *ostart = f.ostart.ensure(M);
    SC_HIP(hipMemsetAsync(head, 0, sizeof(uint32_t) * (size_t(n_out) + 1), s));
    // base[] is indexed by base-ring compact index; inode is a node id < mP for starts
    k_heads<<<grid_for(M), BLK, 0, s>>>(M, inode, dist, base, 0u, head, piece_at, ostart);
    SC_HIP(hipGetLastError());
    scan_inclusive_max_u32(head, head, n_out, s);
    f.ox.ensure(size_t(n_out) + 1);
    f.oy.ensure(size_t(n_out) + 1);
    k_emit<<<grid_for(n_out), BLK, 0, s>>>(0u, n_out, head, piece_at, int(mP), pa.ppos, pa.pdir, P, Q, f.ox, f.oy);
    SC_HIP(hipGetLastError());
    cleanup_rings(c, f, nr, n_out);
}

void clip_all(Ctx &c, int k) {
    SC_REQUIRE(k >= 1 && k <= c.K, SC_EINVAL, "overlay index out of range");
    StageTimes t;
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
    post_overlay(c, k);
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

} // namespace sc
