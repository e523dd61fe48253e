/*
 * seg_oracle.c -- CPU oracle, tree and candidate stages (A1-A15).
 *
 * TEST INFRASTRUCTURE ONLY (see seg_oracle.h).  A single-threaded plain-C
 * restatement of the reference multicore path.  Citations are to the
 * reference repository: "oacc" = p_seg_tree/oacc_segment_tree.cpp,
 * "clip" = src/clip.cpp.
 *
 * Intended semantics used where the reference reads memory it never wrote:
 *  - constructElementaryIntervalsArrayMulticore allocates verticalLines and
 *    elementaryArrayIndex with malloc but its own comments say "use calloc"
 *    (oacc:4563,4566); the oracle zero-initialises them.
 *  - preparePointsToNodeMapMulticore reads elementaryArray[epc] and never
 *    fills boaders[epc-1]/intBoaders[epc-1] (oacc:1985-2006); findBoarder
 *    reads boaders[sortedIndex[-1]] and [epc] (oacc:422,432,454).  The
 *    intended lookup is: the first boundary equal to the query among
 *    elementaryArray[1..epc-1], none if absent; the point-interval skip
 *    only looks at existing points.
 *  - realToInt/radixsort of the boundary keys (oacc:203-206,616-689) sort an
 *    already sorted array; the intended permutation is the identity and the
 *    oracle does not form the int32-overflowing keys.
 *  - insertToLeafLevelEdgeListMulticore may address leaf node T when the
 *    padding holds a single point (oacc:2047); the oracle skips it.
 *  - countEdgeIntersections reads edgeIds[0] of an empty list (clip:534);
 *    with no intersections every count is zero.
 *  - copyIntersectionsParallel copies only min(bSize,cSize) source vertices
 *    of the overlay (oacc:7531-7545, "assume base is the largest polygon");
 *    the oracle copies all of them.
 */
#include "so_internal.h"
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Threads: the loops the reference runs with OpenMP (insertAllParallel
 * oacc:2571-2607, insertToLeafLevelEdgeListMulticore oacc:2010-2113,
 * buildInternalEndListsMulticore oacc:1717-1968, saveIntersectionsIdsMulticore
 * oacc:6010-6244, saveIntersectionsParallel oacc:7257-7341,
 * sortIntersectionsAtEdgesParallel oacc:7436-7473, calculateInitLabelParallel
 * oacc:7925-7988) run as OpenMP loops here too.  Every parallel loop writes
 * its results in the single-thread order (per-node lists re-sorted by edge
 * index, stab output concatenated in node order), so the results do not
 * depend on the thread count. */
void so_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int so_get_threads(void) { return omp_get_max_threads(); }

static int cmp_int(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

void *so_xmalloc(size_t n) {
    void *p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "seg_oracle: out of memory (%zu bytes)\n", n);
        abort();
    }
    return p;
}
void *so_xcalloc(size_t n, size_t s) {
    void *p = calloc(n ? n : 1, s ? s : 1);
    if (!p) {
        fprintf(stderr, "seg_oracle: out of memory\n");
        abort();
    }
    return p;
}

/* ------------------------------------------------------------------ */
/* FP64 predicates                                                      */
/* ------------------------------------------------------------------ */

/* A(P,Q,R), oacc:5334-5337 and foster/point2D.h:72-74 */
double so_A(double px, double py, double qx, double qy, double rx, double ry) {
    double t1 = (qx - px) * (ry - py);
    double t2 = (qy - py) * (rx - px);
    return t1 - t2;
}

/* LSMF bounding-box filter, oacc:5282-5316 */
int so_lsmf8(double l1p1x, double l1p1y, double l1p2x, double l1p2y, double l2p1x, double l2p1y,
             double l2p2x, double l2p2y) {
    double minPX = l1p1x, minPY = l1p1y, maxPX = l1p2x, maxPY = l1p2y;
    double minQX = l2p1x, minQY = l2p1y, maxQX = l2p2x, maxQY = l2p2y;
    if (minPX > l1p2x) { minPX = l1p2x; maxPX = l1p1x; }
    if (minPY > l1p2y) { minPY = l1p2y; maxPY = l1p1y; }
    if (minQX > l2p2x) { minQX = l2p2x; maxQX = l2p1x; }
    if (minQY > l2p2y) { minQY = l2p2y; maxQY = l2p1y; }
    if (minPX > maxQX || maxPX < minQX) return 0;
    if (minPY > maxQY || maxPY < minQY) return 0;
    return 1;
}

/* lineSegmentIntersection / lineSegmentIntersection2, oacc:5426-5611;
 * same classification as foster/polyclip.h:60-179 including the
 * beta_is_0-tests-alpha quirk of the collinear branch (oacc:5501,5595). */
int so_lsi(double b1x, double b1y, double b2x, double b2y, double c1x, double c1y, double c2x,
           double c2y, double *alpha_out, double *beta_out) {
    const double eps = SO_EPS;
    double AP1 = so_A(b1x, b1y, c1x, c1y, c2x, c2y);
    double AP2 = so_A(b2x, b2y, c1x, c1y, c2x, c2y);
    double alpha = 0.0, beta = 0.0;
    int ty = 0;
    if (fabs(AP1 - AP2) > eps) {
        double AQ1 = so_A(c1x, c1y, b1x, b1y, b2x, b2y);
        double AQ2 = so_A(c2x, c2y, b1x, b1y, b2x, b2y);
        alpha = AP1 / (AP1 - AP2);
        beta = AQ1 / (AQ1 - AQ2);
        int a0 = 0, a01 = 0, b0 = 0, b01 = 0;
        if (alpha > eps && alpha < 1.0 - eps) a01 = 1;
        else if (fabs(alpha) <= eps) a0 = 1;
        if (beta > eps && beta < 1.0 - eps) b01 = 1;
        else if (fabs(beta) <= eps) b0 = 1;
        if (a01 && b01) ty = 1;
        else if (a0 && b01) ty = 2;
        else if (b0 && a01) ty = 3;
        else if (a0 && b0) ty = 4;
    } else if (fabs(AP1) < eps) {
        double dPx = b2x - b1x, dPy = b2y - b1y;
        double dQx = c2x - c1x, dQy = c2y - c1y;
        double PQx = c1x - b1x, PQy = c1y - b1y;
        double m1 = PQx * dPx, m2 = PQy * dPy;
        double n1 = dPx * dPx, n2 = dPy * dPy;
        alpha = (m1 + m2) / (n1 + n2);
        double k1 = PQx * dQx, k2 = PQy * dQy;
        double l1 = dQx * dQx, l2 = dQy * dQy;
        beta = -(k1 + k2) / (l1 + l2);
        int a0 = 0, a01 = 0, an = 0, b0 = 0, b01 = 0, bn = 0;
        if (alpha > eps && alpha < 1.0 - eps) a01 = 1;
        else if (fabs(alpha) <= eps) a0 = 1;
        else an = 1;
        if (beta > eps && beta < 1.0 - eps) b01 = 1;
        else if (fabs(alpha) <= eps) b0 = 1;
        else bn = 1;
        if (a01 && b01) ty = 5;
        else if (an && b01) ty = 6;
        else if (bn && a01) ty = 7;
        else if (a0 && b0) ty = 8;
    }
    if (alpha_out) *alpha_out = alpha;
    if (beta_out) *beta_out = beta;
    return ty;
}

int so_lsmf(const double *p) { return so_lsmf8(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]); }
int so_classify(const double *p, double *alpha, double *beta) {
    return so_lsi(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], alpha, beta);
}

/* ------------------------------------------------------------------ */
/* small helpers                                                        */
/* ------------------------------------------------------------------ */

/* Util::getNPowOf2, p_seg_tree/util.cpp:18-28 */
static int next_pow2(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

/* stable merge sort of an index array by double key (thrust::stable_sort_by_key
 * with operator<, oacc:4556) */
static void msort_idx(int *idx, int *tmp, const double *key, int lo, int hi) {
    if (hi - lo < 2) return;
    int mid = lo + (hi - lo) / 2;
    msort_idx(idx, tmp, key, lo, mid);
    msort_idx(idx, tmp, key, mid, hi);
    int i = lo, j = mid, k = lo;
    while (i < mid && j < hi) {
        if (key[idx[j]] < key[idx[i]]) tmp[k++] = idx[j++];
        else tmp[k++] = idx[i++];
    }
    while (i < mid) tmp[k++] = idx[i++];
    while (j < hi) tmp[k++] = idx[j++];
    memcpy(idx + lo, tmp + lo, sizeof(int) * (size_t)(hi - lo));
}

static int decimal_digits(long long m) { /* passes of "for(exp=1; m/exp>0; exp*=10)" */
    int d = 0;
    long long e = 1;
    while (m / e > 0) {
        d++;
        if (e > (long long)4e17) break;
        e *= 10;
    }
    return d;
}

/* ------------------------------------------------------------------ */
/* A1/A2: input layout, clip:172-311                                    */
/* ------------------------------------------------------------------ */

static void edge_get_data(so_edge *e, int id, int type, double a, double b) {
    /* Edge::getData, edge.h:190-201 */
    e->id = id;
    e->type = type;
    e->y = a;
    if (a < b) {
        e->start = a;
        e->end = b;
    } else {
        e->end = a;
        e->start = b;
    }
}

static void build_input(so_tree *t, const double *bxy, const double *cxy) {
    int nb = t->nb, K = t->K;
    t->bpoly = so_xmalloc(sizeof(double) * 2 * (size_t)(nb + 1));
    memcpy(t->bpoly, bxy, sizeof(double) * 2 * (size_t)nb);
    t->bpoly[2 * nb] = bxy[0];
    t->bpoly[2 * nb + 1] = bxy[1];
    t->coff = so_xmalloc(sizeof(int) * (size_t)(K + 1));
    t->coff[0] = 0;
    t->nct = 0;
    for (int i = 0; i < K; i++) {
        t->coff[i + 1] = t->coff[i] + t->csz[i] + 1;
        t->nct += t->csz[i];
    }
    t->cpoly = so_xmalloc(sizeof(double) * 2 * (size_t)(t->nct + K));
    {
        int src = 0;
        for (int i = 0; i < K; i++) {
            int o = t->coff[i];
            memcpy(t->cpoly + 2 * (size_t)o, cxy + 2 * (size_t)src, sizeof(double) * 2 * (size_t)t->csz[i]);
            t->cpoly[2 * (size_t)(o + t->csz[i])] = cxy[2 * (size_t)src];
            t->cpoly[2 * (size_t)(o + t->csz[i]) + 1] = cxy[2 * (size_t)src + 1];
            src += t->csz[i];
        }
    }
    t->ne = nb + t->nct;
    t->edges = so_xmalloc(sizeof(so_edge) * (size_t)t->ne);
    int id = 0;
    for (int v = 0; v < nb; v++) {
        edge_get_data(&t->edges[id], 2 * v, 0, t->bpoly[2 * v], t->bpoly[2 * v + 2]);
        if (v < nb - 1) { /* min/max from the non-closing base edges, clip:206-218 */
            if (v == 0) {
                t->minv = t->edges[id].start;
                t->maxv = t->edges[id].end;
            }
            if (t->minv > t->edges[id].start) t->minv = t->edges[id].start;
            if (t->maxv < t->edges[id].end) t->maxv = t->edges[id].end;
        }
        id++;
    }
    for (int i = 0; i < K; i++) {
        int o = t->coff[i];
        for (int j = 0; j < t->csz[i]; j++) {
            int g = o + j;
            edge_get_data(&t->edges[id], 2 * g, i + 1, t->cpoly[2 * (size_t)g], t->cpoly[2 * (size_t)g + 2]);
            if (j < t->csz[i] - 1) {
                if (t->minv > t->edges[id].start) t->minv = t->edges[id].start;
                if (t->maxv < t->edges[id].end) t->maxv = t->edges[id].end;
            }
            id++;
        }
    }
}

/* ------------------------------------------------------------------ */
/* A3: elementary intervals, oacc:4513-4692                             */
/* ------------------------------------------------------------------ */

static int build_elementary(so_tree *t) {
    int ic = t->ne;
    size_t m = 2 * (size_t)ic;
    double *ele = so_xmalloc(sizeof(double) * m);
    int *sidx = so_xmalloc(sizeof(int) * m);
    int *tmp = so_xmalloc(sizeof(int) * m);
    for (int j = 0; j < ic; j++) {
        ele[2 * j] = t->edges[j].start;
        ele[2 * j + 1] = t->edges[j].end;
        sidx[2 * j] = 2 * j;
        sidx[2 * j + 1] = 2 * j + 1;
    }
    msort_idx(sidx, tmp, ele, 0, (int)m);
    free(tmp);

    /* vertical-line flags with the reference's indexing (oacc:4569-4577),
     * zero-initialised as its comment intends */
    int *vl = so_xcalloc(m + 1, sizeof(int));
    for (int j = 0; j < ic; j++) {
        int i = sidx[j] / 2;
        vl[i] = 0;
        if (i == ic - 1) vl[i + 1] = 0;
        if (t->edges[i].start == t->edges[i].end) vl[i] = 1;
    }
    { /* exclusive scan over ic*2+1 entries, oacc:4579 */
        int run = 0;
        for (size_t i = 0; i <= m; i++) {
            int v = vl[i];
            vl[i] = run;
            run += v;
        }
    }
    int nvl = vl[m];
    int *vloc = so_xcalloc((size_t)(nvl > 0 ? nvl : 1), sizeof(int));
    for (int i = 0; i < ic; i++)
        if (vl[i] != vl[i + 1]) vloc[vl[i]] = i;

    /* unique with the last-of-group rule and the vertical duplicate rule,
     * oacc:4601-4635 */
    int *eai = so_xcalloc(m + 1, sizeof(int));
    double *ea3 = so_xcalloc(m, sizeof(double));
    for (size_t i = 0; i + 1 < m; i++) {
        eai[i] = 0;
        if (i == m - 2) {
            eai[i + 1] = 0;
            eai[i + 2] = 0;
        }
        int intvi = sidx[i] / 2;
        int comid = vl[i / 2];
        int lookup = comid > 0 ? vloc[comid - 1] : 0;
        if (ele[sidx[i]] != ele[sidx[i + 1]]) {
            ea3[i] = ele[sidx[i]];
            eai[i] = 1;
        } else if (t->edges[intvi].start == t->edges[intvi].end &&
                   (comid == 0 || (t->edges[lookup].start != t->edges[intvi].start &&
                                   t->edges[lookup].end != t->edges[intvi].start))) {
            ea3[i] = t->edges[intvi].start;
            eai[i] = 1;
        }
        if (i == m - 2) {
            if (ele[sidx[m - 2]] != ele[sidx[m - 1]]) {
                ea3[m - 1] = ele[sidx[m - 2]];
                eai[m - 1] = 1;
            }
        }
    }
    {
        int run = 0;
        for (size_t i = 0; i <= m; i++) {
            int v = eai[i];
            eai[i] = run;
            run += v;
        }
    }
    int cnt = eai[m];
    if (cnt < 1) {
        free(ele); free(sidx); free(vl); free(vloc); free(eai); free(ea3);
        return -1;
    }
    int round = next_pow2(cnt) + 1;
    t->ea = so_xmalloc(sizeof(double) * (size_t)round);
    for (size_t i = 0; i < m; i++)
        if (eai[i] != eai[i + 1]) t->ea[eai[i]] = ea3[i];
    for (int i = 0; i < round - cnt; i++) t->ea[cnt + i] = t->ea[cnt - 1];
    t->epc = round;
    t->nleaves = round - 1;
    t->T = 2 * t->nleaves; /* init(), seg.h:129 */
    free(ele); free(sidx); free(vl); free(vloc); free(eai); free(ea3);
    return 0;
}

/* ------------------------------------------------------------------ */
/* A4: node interval array, oacc:4948-4994                              */
/* ------------------------------------------------------------------ */

static void build_node_intervals(so_tree *t) {
    int T = t->T;
    t->niv = so_xcalloc(2 * (size_t)T, sizeof(double));
    for (int l = 0; l < t->nleaves; l++) { /* leaves from elementaryEdgeArray */
        t->niv[T + 2 * l] = t->ea[l];
        t->niv[T + 2 * l + 1] = t->ea[l + 1];
    }
    for (int lstart = T / 4; lstart >= 1; lstart /= 2)
        for (int v = lstart; v < 2 * lstart; v++) {
            t->niv[2 * v] = t->niv[2 * (2 * v)];
            t->niv[2 * v + 1] = t->niv[2 * (2 * v + 1) + 1];
        }
}

/* ------------------------------------------------------------------ */
/* A5: cover lists by recursive insertion, oacc:2496-2607               */
/* ------------------------------------------------------------------ */

static int edge_contains(const so_edge *x, double s, double e) { /* edge.h:170 */
    return x->start <= s && x->end >= e;
}
static int edge_intersects(const so_edge *x, double es, double ee) { /* edge.h:181 */
    if ((x->start <= es && x->end >= ee) || (es <= x->start && ee >= x->end)) return 1;
    if ((x->start <= es && x->end >= es) || (ee <= x->end && ee >= x->start)) return 1;
    return 0;
}

static void insert_rec(const so_tree *t, const so_edge *x, int eidx, int v, int64_t *cnt, int *out,
                       int64_t *fill) {
    if (edge_contains(x, t->niv[2 * v], t->niv[2 * v + 1])) {
        if (out) {
            int64_t at;
#pragma omp atomic capture
            at = fill[v]++;
            out[at] = eidx;
        } else {
#pragma omp atomic
            cnt[v]++;
        }
        return;
    }
    int l = 2 * v, r = 2 * v + 1;
    if (l < t->T && edge_intersects(x, t->niv[2 * l], t->niv[2 * l + 1]))
        insert_rec(t, x, eidx, l, cnt, out, fill);
    if (r < t->T && edge_intersects(x, t->niv[2 * r], t->niv[2 * r + 1]))
        insert_rec(t, x, eidx, r, cnt, out, fill);
}

/* insertAllParallel: every edge inserted recursively, in parallel over
 * edges (the reference locks the node, oacc:2527; here the node's counter
 * is advanced atomically).  Each node's list is then sorted by edge index,
 * the order of a single-thread insertion in edge order. */
static void build_cover(so_tree *t) {
    int T = t->T;
    int64_t *cnt = so_xcalloc((size_t)T + 1, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 1024)
    for (int e = 0; e < t->ne; e++) insert_rec(t, &t->edges[e], e, 1, cnt, NULL, NULL);
    t->cov_off = so_xmalloc(sizeof(int64_t) * ((size_t)T + 1));
    int64_t run = 0;
    for (int v = 0; v <= T; v++) {
        t->cov_off[v] = run;
        if (v < T) run += cnt[v];
    }
    t->cov_e = so_xmalloc(sizeof(int) * (size_t)(run > 0 ? run : 1));
    int64_t *fill = cnt;
    for (int v = 0; v < T; v++) fill[v] = t->cov_off[v];
#pragma omp parallel for schedule(dynamic, 1024)
    for (int e = 0; e < t->ne; e++) insert_rec(t, &t->edges[e], e, 1, NULL, t->cov_e, fill);
#pragma omp parallel for schedule(dynamic, 256)
    for (int v = 1; v < T; v++) {
        const int64_t n = t->cov_off[v + 1] - t->cov_off[v];
        if (n > 1) qsort(t->cov_e + t->cov_off[v], (size_t)n, sizeof(int), cmp_int);
    }
    free(cnt);
}

/* ------------------------------------------------------------------ */
/* A6/A7: end lists, oacc:1717-2135, 863-922                            */
/* ------------------------------------------------------------------ */

/* first boundary index k in [0, epc-2] with elementaryArray[k+1] == q
 * (findBoarder, oacc:420-460, intended lookup over the filled boundaries) */
static int find_boundary(const so_tree *t, double q) {
    int lo = 0, hi = t->epc - 1; /* boundaries b[k] = ea[k+1] */
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if (t->ea[mid + 1] < q) lo = mid + 1;
        else hi = mid;
    }
    if (lo < t->epc - 1 && t->ea[lo + 1] == q) return lo;
    return -1;
}

static int edge_key_less(const so_tree *t, int a, int b) {
    const so_edge *x = &t->edges[a], *y = &t->edges[b];
    if (x->id != y->id) return x->id < y->id;
    return x->type < y->type;
}

/* decimal LSD sort of one leaf list: one pass on the type's units digit
 * (MAX_POLY_TYPE=2), then passes over the id digits of 'maxv'
 * (edgeRadixsort, oacc:906-922) */
static void leaf_radix(const so_tree *t, int *a, int n, int *tmp, int maxv) {
    for (int pass = -1; pass < 12; pass++) {
        long long exp = 1;
        if (pass >= 0) {
            for (int q = 0; q < pass; q++) exp *= 10;
            if (maxv / exp <= 0) break;
        }
        int count[10] = {0};
        for (int i = 0; i < n; i++) {
            const so_edge *e = &t->edges[a[i]];
            int d = pass < 0 ? (e->type % 10) : (int)((e->id / exp) % 10);
            count[d]++;
        }
        for (int i = 1; i < 10; i++) count[i] += count[i - 1];
        for (int i = n - 1; i >= 0; i--) {
            const so_edge *e = &t->edges[a[i]];
            int d = pass < 0 ? (e->type % 10) : (int)((e->id / exp) % 10);
            tmp[--count[d]] = a[i];
        }
        memcpy(a, tmp, sizeof(int) * (size_t)n);
    }
}

static void build_end_lists(so_tree *t) {
    int T = t->T, half = T / 2, epc = t->epc;
    /* preparePointsToNodeMapMulticore, oacc:1970-2008 */
    int *left = so_xmalloc(sizeof(int) * (size_t)epc);
    int *right = so_xmalloc(sizeof(int) * (size_t)epc);
    for (int i = 1; i < epc; i++) {
        int ll = half + i - 1, rl = half + i;
        if (i != 1 && i + 1 <= epc - 1 && t->ea[i] == t->ea[i + 1]) rl = half + i + 1;
        left[i - 1] = ll;
        right[i - 1] = rl;
    }
    /* insertToLeafLevelEdgeListMulticore, oacc:2010-2113 */
    t->after_leaf = so_xmalloc(sizeof(int) * (size_t)t->ne);
    t->before_leaf = so_xmalloc(sizeof(int) * (size_t)t->ne);
    int64_t *lcnt = so_xcalloc((size_t)T + 1, sizeof(int64_t));
#pragma omp parallel for schedule(static)
    for (int e = 0; e < t->ne; e++) {
        int en = find_boundary(t, t->edges[e].end);
        int st = find_boundary(t, t->edges[e].start);
        int after = -1, before = -1;
        if (en >= 0) {
            int a = right[en];
            if (a < T && t->niv[2 * a] != t->niv[2 * a + 1]) after = a;
        }
        if (st >= 0) {
            int b = left[st];
            if (t->niv[2 * b] != t->niv[2 * b + 1]) before = b;
        }
        t->after_leaf[e] = after;
        t->before_leaf[e] = before;
        if (after >= 0) {
#pragma omp atomic
            lcnt[after]++;
        }
        if (before >= 0) {
#pragma omp atomic
            lcnt[before]++;
        }
    }
    /* per node lists, built level by level (buildInternalEndListsMulticore) */
    int64_t *off = so_xcalloc((size_t)T + 1, sizeof(int64_t));
    int **lists = so_xcalloc((size_t)T, sizeof(int *));
    int64_t *lens = so_xcalloc((size_t)T, sizeof(int64_t));
    for (int v = half; v < T; v++) {
        lists[v] = so_xmalloc(sizeof(int) * (size_t)(lcnt[v] > 0 ? lcnt[v] : 1));
        lens[v] = 0;
    }
    for (int e = 0; e < t->ne; e++) { /* scatter in edge order: after, then before */
        if (t->after_leaf[e] >= 0) { int v = t->after_leaf[e]; lists[v][lens[v]++] = e; }
        if (t->before_leaf[e] >= 0) { int v = t->before_leaf[e]; lists[v][lens[v]++] = e; }
    }
    int maxId = t->nb >= t->nct + t->K ? t->nb : t->nct + t->K; /* clip:1581-1585 */
    {
        /* (type, id) keys are distinct within a leaf, so the order after the
         * sort does not depend on the scatter order */
        int64_t mx = 0;
        for (int v = half; v < T; v++) if (lens[v] > mx) mx = lens[v];
#pragma omp parallel
        {
            int *tmp = so_xmalloc(sizeof(int) * (size_t)(mx > 0 ? mx : 1));
#pragma omp for schedule(dynamic, 1024)
            for (int v = half; v < T; v++) leaf_radix(t, lists[v], (int)lens[v], tmp, 2 * maxId);
            free(tmp);
        }
    }
    /* internal levels: id-merge with per-type fill buffer (oacc:1758-1950) */
    int ntypes = t->K + 1; /* POLY_TYPE_COUNT = argc-1 */
    for (int lstart = half / 2; lstart >= 1; lstart /= 2) {
#pragma omp parallel
        {
        int *fill = so_xmalloc(sizeof(int) * (size_t)ntypes);
#pragma omp for schedule(dynamic, 256)
        for (int v = lstart; v < 2 * lstart; v++) {
            int c1 = 2 * v, c2 = 2 * v + 1;
            int64_t n1 = lens[c1], n2 = lens[c2];
            int *out = so_xmalloc(sizeof(int) * (size_t)(n1 + n2 > 0 ? n1 + n2 : 1));
            int64_t ne_ = 0;
            for (int f = 0; f < ntypes; f++) fill[f] = -1;
#define FLUSH()                                                   \
    do {                                                          \
        for (int f = 0; f < ntypes; f++)                          \
            if (fill[f] >= 0) out[ne_++] = fill[f];               \
        for (int f = 0; f < ntypes; f++) fill[f] = -1;            \
    } while (0)
            if (n1 >= 1 && n2 >= 1) {
                int *A = lists[c1], *B = lists[c2];
                int64_t i = 0, j = 0;
                int previd = t->edges[A[0]].id <= t->edges[B[0]].id ? t->edges[A[0]].id : t->edges[B[0]].id;
                while (i < n1 && j < n2) {
                    int ia = t->edges[A[i]].id, ib = t->edges[B[j]].id;
                    int curr = ia <= ib ? ia : ib;
                    if (curr != previd) { FLUSH(); previd = curr; }
                    if (ia < ib) { fill[t->edges[A[i]].type] = A[i]; i++; }
                    else if (ia > ib) { fill[t->edges[B[j]].type] = B[j]; j++; }
                    else {
                        fill[t->edges[A[i]].type] = A[i];
                        fill[t->edges[B[j]].type] = B[j];
                        i++; j++;
                    }
                }
                for (; i < n1; i++) {
                    int curr = t->edges[A[i]].id;
                    if (curr != previd) { FLUSH(); previd = curr; }
                    fill[t->edges[A[i]].type] = A[i];
                }
                for (; j < n2; j++) {
                    int curr = t->edges[B[j]].id;
                    if (curr != previd) { FLUSH(); previd = curr; }
                    fill[t->edges[B[j]].type] = B[j];
                }
                FLUSH();
            } else if (n1 >= 1) {
                memcpy(out, lists[c1], sizeof(int) * (size_t)n1);
                ne_ = n1;
            } else if (n2 >= 1) {
                memcpy(out, lists[c2], sizeof(int) * (size_t)n2);
                ne_ = n2;
            }
#undef FLUSH
            lists[v] = out;
            lens[v] = ne_;
        }
        free(fill);
        }
    }
    int64_t run = 0;
    for (int v = 0; v <= T; v++) {
        off[v] = run;
        if (v < T && v >= 1) run += lens[v];
    }
    t->end_off = off;
    t->end_e = so_xmalloc(sizeof(int) * (size_t)(run > 0 ? run : 1));
    for (int v = 1; v < T; v++) {
        if (lens[v]) memcpy(t->end_e + off[v], lists[v], sizeof(int) * (size_t)lens[v]);
        free(lists[v]);
    }
    free(lists);
    free(lens);
    free(lcnt);
    free(left);
    free(right);
    (void)edge_key_less;
}

/* ------------------------------------------------------------------ */
/* build / getters                                                      */
/* ------------------------------------------------------------------ */

so_tree *so_build(const double *bxy, int nb, const double *cxy, const int *csz, int K) {
    if (!bxy || !cxy || !csz || nb < 3 || K < 1) return NULL;
    for (int i = 0; i < K; i++)
        if (csz[i] < 3) return NULL;
    so_tree *t = so_xcalloc(1, sizeof(so_tree));
    t->nb = nb;
    t->K = K;
    t->csz = so_xmalloc(sizeof(int) * (size_t)K);
    memcpy(t->csz, csz, sizeof(int) * (size_t)K);
    build_input(t, bxy, cxy);
    if (build_elementary(t) != 0) {
        so_free(t);
        return NULL;
    }
    build_node_intervals(t);
    build_cover(t);
    build_end_lists(t);
    return t;
}

void so_free(so_tree *t) {
    if (!t) return;
    free(t->csz); free(t->coff); free(t->bpoly); free(t->cpoly); free(t->edges);
    free(t->ea); free(t->niv); free(t->cov_off); free(t->cov_e); free(t->end_off);
    free(t->end_e); free(t->after_leaf); free(t->before_leaf);
    free(t);
}

int so_num_edges(const so_tree *t) { return t->ne; }
const so_edge *so_edges(const so_tree *t) { return t->edges; }
const double *so_bpoly(const so_tree *t) { return t->bpoly; }
const double *so_cpoly(const so_tree *t) { return t->cpoly; }
const int *so_coff(const so_tree *t) { return t->coff; }
int so_epc(const so_tree *t) { return t->epc; }
const double *so_elementary(const so_tree *t) { return t->ea; }
int so_tree_size(const so_tree *t) { return t->T; }
const double *so_node_intervals(const so_tree *t) { return t->niv; }
const int64_t *so_cover_off(const so_tree *t) { return t->cov_off; }
const int *so_cover_edges(const so_tree *t) { return t->cov_e; }
const int64_t *so_end_off(const so_tree *t) { return t->end_off; }
const int *so_end_edges(const so_tree *t) { return t->end_e; }
const int *so_after_leaf(const so_tree *t) { return t->after_leaf; }
const int *so_before_leaf(const so_tree *t) { return t->before_leaf; }

/* ------------------------------------------------------------------ */
/* A8: stab and candidate emission, oacc:6010-6244                      */
/* ------------------------------------------------------------------ */

typedef struct {
    int *b, *c, *ty;
    int64_t n, cap;
} trip_vec;

static void trip_push(trip_vec *v, int b, int c, int ty) {
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 1024;
        v->b = realloc(v->b, sizeof(int) * (size_t)v->cap);
        v->c = realloc(v->c, sizeof(int) * (size_t)v->cap);
        v->ty = realloc(v->ty, sizeof(int) * (size_t)v->cap);
        if (!v->b || !v->c || !v->ty) abort();
    }
    v->b[v->n] = b;
    v->c[v->n] = c;
    v->ty[v->n] = ty;
    v->n++;
}

static void test_pair(const so_tree *t, int bId, int cId, trip_vec *out) {
    const double *B = t->bpoly, *C = t->cpoly;
    if (so_lsmf8(B[bId], B[bId + 1], B[bId + 2], B[bId + 3], C[cId], C[cId + 1], C[cId + 2], C[cId + 3])) {
        int ty = so_lsi(B[bId], B[bId + 1], B[bId + 2], B[bId + 3], C[cId], C[cId + 1], C[cId + 2],
                        C[cId + 3], NULL, NULL);
        if (ty) trip_push(out, bId, cId, ty);
    }
}

/* One node's pairs in the reference's loop order (oacc:6149-6241). */
static void stab_node(const so_tree *t, int v, int cType, trip_vec *out) {
    const int bType = 0;
    {
        int64_t cs = t->cov_off[v], ce = t->cov_off[v + 1];
        if (cs == ce) return; /* no Node allocated, oacc:6149-6150 */
        for (int64_t i = cs; i < ce; i++) {
            const so_edge *it = &t->edges[t->cov_e[i]];
            for (int64_t j = cs; j < ce; j++) {
                const so_edge *in = &t->edges[t->cov_e[j]];
                if (it->type == bType && in->type == cType) test_pair(t, it->id, in->id, out);
                else if (it->type == cType && in->type == bType) test_pair(t, in->id, it->id, out);
            }
            for (int64_t j = t->end_off[v]; j < t->end_off[v + 1]; j++) {
                const so_edge *en = &t->edges[t->end_e[j]];
                if (it->type == bType && en->type == cType) test_pair(t, it->id, en->id, out);
                else if (it->type == cType && en->type == bType) test_pair(t, en->id, it->id, out);
            }
        }
    }
}

/* saveIntersectionsIdsMulticore: every node, in parallel over chunks of
 * consecutive nodes (the reference appends under omp critical,
 * oacc:6170,6226); the chunks' pairs are concatenated in node order, which is
 * the single-thread emission order. */
static void stab(const so_tree *t, int cType, trip_vec *out) {
    const int CH = 64;
    const int nch = (t->T + CH - 1) / CH;
    trip_vec *part = so_xcalloc((size_t)nch, sizeof(trip_vec));
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < nch; q++) {
        const int v0 = q * CH > 1 ? q * CH : 1, v1 = (q + 1) * CH < t->T ? (q + 1) * CH : t->T;
        for (int v = v0; v < v1; v++) stab_node(t, v, cType, &part[q]);
    }
    int64_t total = out->n;
    for (int q = 0; q < nch; q++) total += part[q].n;
    if (total > out->cap) {
        out->cap = total > 0 ? total : 1;
        out->b = realloc(out->b, sizeof(int) * (size_t)out->cap);
        out->c = realloc(out->c, sizeof(int) * (size_t)out->cap);
        out->ty = realloc(out->ty, sizeof(int) * (size_t)out->cap);
        if (!out->b || !out->c || !out->ty) abort();
    }
    for (int q = 0; q < nch; q++) {
        if (part[q].n) {
            memcpy(out->b + out->n, part[q].b, sizeof(int) * (size_t)part[q].n);
            memcpy(out->c + out->n, part[q].c, sizeof(int) * (size_t)part[q].n);
            memcpy(out->ty + out->n, part[q].ty, sizeof(int) * (size_t)part[q].n);
            out->n += part[q].n;
        }
        free(part[q].b);
        free(part[q].c);
        free(part[q].ty);
    }
    free(part);
}

/* ------------------------------------------------------------------ */
/* A10: decimal radix sort + duplicate removal, clip:368-493            */
/* ------------------------------------------------------------------ */

/* radixsort(int*,int*,n,m), clip:422-435: stable, one pass per decimal digit
 * of m; sets *trunc when some key has more digits than m */
static void radix_int(const int *data, int *oidx, int n, long long m, int *trunc) {
    int *in = so_xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *o = so_xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *oi = so_xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *ls = so_xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        in[i] = data[i];
        oidx[i] = i;
    }
    int passes = decimal_digits(m);
    for (int p = 0; p < passes; p++) {
        int count[10] = {0};
        for (int i = 0; i < n; i++) {
            ls[i] = in[i] % 10;
            count[ls[i]]++;
        }
        for (int i = 1; i < 10; i++) count[i] += count[i - 1];
        for (int i = n - 1; i >= 0; i--) {
            o[count[ls[i]] - 1] = in[i];
            oi[count[ls[i]] - 1] = oidx[i];
            count[ls[i]]--;
        }
        for (int i = 0; i < n; i++) {
            in[i] = o[i] / 10;
            oidx[i] = oi[i];
        }
    }
    for (int i = 0; i < n; i++)
        if (in[i] != 0 && trunc) *trunc = 1;
    free(in); free(o); free(oi); free(ls);
}

/* ------------------------------------------------------------------ */
/* A11-A15                                                              */
/* ------------------------------------------------------------------ */

static int non_degenerate(char poly, int ty) { /* isNonDegenerateCase, clip:508-517 */
    if (poly == 'b' && (ty == 1 || ty == 3 || ty == 5 || ty == 7)) return 1;
    if (poly == 'c' && (ty == 1 || ty == 2 || ty == 6)) return 1;
    return 0;
}

/* countEdgeIntersections, clip:522-601; returns unique edge count */
static int count_edge_intersections(const int *ids, const int *sidx, int cStart, const int *types,
                                    int n, char poly, int *all, int *nd, int *uids, int polysize) {
    for (int i = 0; i < polysize; i++) all[i] = nd[i] = 0;
    if (n <= 0) return 0;
    int *ec = so_xmalloc(sizeof(int) * (size_t)n), *ndc = so_xmalloc(sizeof(int) * (size_t)n);
    int j = 0, edgeCount = 1, ndCount = 0;
    int cur = ids[sidx[0]];
    ec[0] = edgeCount;
    uids[0] = cur;
    if (non_degenerate(poly, types[sidx[0]])) ndCount = 1;
    ndc[0] = ndCount;
    for (int i = 1; i < n; i++) {
        if (ids[sidx[i]] == cur) {
            edgeCount++;
            if (non_degenerate(poly, types[sidx[i]])) ndCount++;
        } else {
            ec[j] = edgeCount;
            ndc[j] = ndCount;
            uids[j + 1] = ids[sidx[i]];
            cur = ids[sidx[i]];
            j++;
            edgeCount = 1;
            ndCount = non_degenerate(poly, types[sidx[i]]) ? 1 : 0;
        }
        if (i == n - 1) {
            ec[j] = edgeCount;
            ndc[j] = ndCount;
        }
    }
    int nu = j + 1;
    j = 0;
    for (int i = 0; i < polysize; i++) {
        if (j < nu && i + cStart == uids[j] / 2) {
            all[i] = ec[j];
            nd[i] = ndc[j];
            j++;
        }
    }
    free(ec);
    free(ndc);
    return j;
}

static void exclusive_scan_int(int *a, int n) {
    int run = 0;
    for (int i = 0; i < n; i++) {
        int v = a[i];
        a[i] = run;
        run += v;
    }
}

/* alphaToInt, oacc:7247-7252.  The reference converts 1e9*a to int without
 * a range check; beyond INT_MAX that conversion is undefined (x86 gives
 * INT_MIN, and the 9-digit radix sort then indexes its counts with a negative
 * digit).  Defined here as saturation to INT_MAX, the value the GPU's
 * conversion gives, and reported through *oor (status bit 16). */
static int alpha_to_int(double a, int *oor) {
    if (a < SO_EPS) return 0;
    const double q = (double)1000000000 * a;
    if (!(q < 2147483647.0)) {
        *oor = 1;
        return 2147483647;
    }
    return (int)q;
}

/* radixsort(LONG*,...) over [s,e) with 9 passes, oacc:616-689 as called by
 * sortIntersectionsAtEdgesParallel (oacc:7466-7468) */
static void radix_long_range(long long *vals, int *idx, int s, int e) {
    int n = e - s;
    if (n <= 1) return;
    long long *in = so_xmalloc(sizeof(long long) * (size_t)n);
    long long *o = so_xmalloc(sizeof(long long) * (size_t)n);
    int *oi = so_xmalloc(sizeof(int) * (size_t)n), *ls = so_xmalloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; i++) in[i] = vals[s + i];
    for (int p = 0; p < SO_EPS_POSITIONS; p++) {
        int count[10] = {0};
        for (int i = 0; i < n; i++) {
            ls[i] = (int)(in[i] % 10);
            count[ls[i]]++;
        }
        for (int i = 1; i < 10; i++) count[i] += count[i - 1];
        for (int i = n - 1; i >= 0; i--) {
            o[count[ls[i]] - 1] = in[i];
            oi[count[ls[i]] - 1] = idx[s + i];
            count[ls[i]]--;
        }
        for (int i = 0; i < n; i++) {
            in[i] = o[i] / 10;
            idx[s + i] = oi[i];
        }
    }
    free(in); free(o); free(oi); free(ls);
}

/* sortIntersectionsAtEdgesParallel, oacc:7436-7473 */
static void sort_at_edges(const int *qa, const int *sidx, int cStart, const int *all, const int *nd,
                          const int *uids, int nu, int n, int *out, int always) {
    long long *cp = so_xmalloc(sizeof(long long) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        out[i] = i;
        cp[i] = qa[sidx[i]];
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int u = 0; u < nu; u++) {
        int eid = uids[u] / 2 - cStart;
        int s = all[eid], e = all[eid + 1];
        int ns = nd[eid], ne_ = nd[eid + 1];
        if (always ? (e - s) > 1 : (ne_ - ns) > 1) radix_long_range(cp, out, s, e);
    }
    free(cp);
}

so_result *so_clip(so_tree *t, int k, int canonical) { return so_clip_op(t, k, canonical, 0); }

so_result *so_clip_op(so_tree *t, int k, int canonical, int uni) {
    if (!t || k < 1 || k > t->K) return NULL;
    so_result *r = so_xcalloc(1, sizeof(so_result));
    int bSize = t->nb, cSize = t->csz[k - 1], cStart = t->coff[k - 1];
    const double *bPoly = t->bpoly, *cPoly = t->cpoly;

    /* A8 */
    trip_vec cand = {0};
    stab(t, k, &cand);
    r->ncand = cand.n;
    int nd_ = (int)cand.n;
    int trunc = 0;

    int *bIds = so_xmalloc(sizeof(int) * (size_t)(nd_ > 0 ? nd_ : 1));
    int *cIds = so_xmalloc(sizeof(int) * (size_t)(nd_ > 0 ? nd_ : 1));
    int *types = so_xmalloc(sizeof(int) * (size_t)(nd_ > 0 ? nd_ : 1));
    int I = 0;
    if (!canonical) {
        /* A10 literal: stable radix by bId, keep first of each (bId,cId) */
        int *sb = so_xmalloc(sizeof(int) * (size_t)(nd_ > 0 ? nd_ : 1));
        radix_int(cand.b, sb, nd_, 2LL * bSize, &trunc);
        char *dup = so_xcalloc((size_t)(nd_ > 0 ? nd_ : 1), 1);
        for (int id = 0; id < nd_;) {
            int sid = sb[id];
            if (!dup[sid]) {
                for (int id2 = id + 1; id2 < nd_; id2++) {
                    int sid2 = sb[id2];
                    if (cand.b[sid] == cand.b[sid2] && cand.c[sid] == cand.c[sid2]) dup[sid2] = 1;
                    if (cand.b[sid] != cand.b[sid2]) break;
                }
                bIds[I] = cand.b[sid];
                cIds[I] = cand.c[sid];
                types[I] = cand.ty[sid];
                I++;
                id++;
            } else {
                ++id;
            }
        }
        free(sb);
        free(dup);
    } else {
        /* canonical: unique (bId,cId) in lexicographic order */
        int64_t *key = so_xmalloc(sizeof(int64_t) * (size_t)(nd_ > 0 ? nd_ : 1));
        int *ord = so_xmalloc(sizeof(int) * (size_t)(nd_ > 0 ? nd_ : 1));
        for (int i = 0; i < nd_; i++) {
            key[i] = ((int64_t)cand.b[i] << 32) | (uint32_t)cand.c[i];
            ord[i] = i;
        }
        /* insertion-free: sort indices by key with a simple bottom-up merge */
        int *tmp = so_xmalloc(sizeof(int) * (size_t)(nd_ > 0 ? nd_ : 1));
        for (int w = 1; w < nd_; w *= 2) {
            for (int lo = 0; lo < nd_; lo += 2 * w) {
                int mid = lo + w < nd_ ? lo + w : nd_, hi = lo + 2 * w < nd_ ? lo + 2 * w : nd_;
                int i = lo, j = mid, q = lo;
                while (i < mid && j < hi) tmp[q++] = key[ord[j]] < key[ord[i]] ? ord[j++] : ord[i++];
                while (i < mid) tmp[q++] = ord[i++];
                while (j < hi) tmp[q++] = ord[j++];
            }
            memcpy(ord, tmp, sizeof(int) * (size_t)nd_);
        }
        for (int i = 0; i < nd_; i++) {
            if (i > 0 && key[ord[i]] == key[ord[i - 1]]) continue;
            bIds[I] = cand.b[ord[i]];
            cIds[I] = cand.c[ord[i]];
            types[I] = cand.ty[ord[i]];
            I++;
        }
        free(key); free(ord); free(tmp);
    }
    free(cand.b); free(cand.c); free(cand.ty);
    r->I = I;
    r->pb = bIds;
    r->pc = cIds;
    r->pt = types;

    int *bSorted = so_xmalloc(sizeof(int) * (size_t)(I > 0 ? I : 1));
    for (int i = 0; i < I; i++) bSorted[i] = i; /* bPolLineIdsSortedIndex[j]=j, clip:486 */
    int *cSorted = so_xmalloc(sizeof(int) * (size_t)(I > 0 ? I : 1));
    radix_int(cIds, cSorted, I, 2LL * t->nct, &trunc); /* clip:1805 */
    if (trunc) r->status |= 8;

    /* A11 */
    int *bAll = so_xcalloc((size_t)bSize + 1, sizeof(int)), *bNd = so_xcalloc((size_t)bSize + 1, sizeof(int));
    int *cAll = so_xcalloc((size_t)cSize + 1, sizeof(int)), *cNd = so_xcalloc((size_t)cSize + 1, sizeof(int));
    int *bU = so_xmalloc(sizeof(int) * (size_t)(I > 0 ? I : 1)), *cU = so_xmalloc(sizeof(int) * (size_t)(I > 0 ? I : 1));
    int nuB = count_edge_intersections(bIds, bSorted, 0, types, I, 'b', bAll, bNd, bU, bSize);
    int nuC = count_edge_intersections(cIds, cSorted, cStart, types, I, 'c', cAll, cNd, cU, cSize);
    exclusive_scan_int(bAll, bSize + 1);
    exclusive_scan_int(cAll, cSize + 1);
    exclusive_scan_int(bNd, bSize + 1);
    exclusive_scan_int(cNd, cSize + 1);
    int nbAug = bSize + bNd[bSize], ncAug = cSize + cNd[cSize];

    /* A12: saveIntersectionsParallel, oacc:7257-7341 */
    int oor = 0; /* a quantised alpha / beta beyond the int range */
    double *inter = so_xcalloc(2 * (size_t)(I > 0 ? I : 1), sizeof(double));
    int *qaB = so_xcalloc((size_t)(I > 0 ? I : 1), sizeof(int)), *qaC = so_xcalloc((size_t)(I > 0 ? I : 1), sizeof(int));
#pragma omp parallel for schedule(static) reduction(| : oor)
    for (int id = 0; id < I; id++) {
        int bId = bIds[id], cId = cIds[id];
        double alpha, beta;
        int ty = so_lsi(bPoly[bId], bPoly[bId + 1], bPoly[bId + 2], bPoly[bId + 3], cPoly[cId],
                        cPoly[cId + 1], cPoly[cId + 2], cPoly[cId + 3], &alpha, &beta);
        if (!ty) continue;
        switch (ty) {
        case 1: {
            double w = 1.0 - alpha;
            double ax = w * bPoly[bId], ay = w * bPoly[bId + 1];
            double bx = alpha * bPoly[bId + 2], by = alpha * bPoly[bId + 3];
            inter[2 * id] = ax + bx;
            inter[2 * id + 1] = ay + by;
        } break;
        case 3: case 4: case 5: case 7: case 8:
            inter[2 * id] = cPoly[cId];
            inter[2 * id + 1] = cPoly[cId + 1];
            break;
        case 2: case 6:
            inter[2 * id] = bPoly[bId];
            inter[2 * id + 1] = bPoly[bId + 1];
            break;
        }
        types[id] = ty;
        qaB[id] = alpha_to_int(alpha, &oor);
        qaC[id] = alpha_to_int(beta, &oor);
    }
    if (oor) r->status |= 16;
    r->inter = inter;
    r->qa_b = qaB;
    r->qa_c = qaC;

    /* A13 */
    int *sB = so_xmalloc(sizeof(int) * (size_t)(I > 0 ? I : 1)), *sC = so_xmalloc(sizeof(int) * (size_t)(I > 0 ? I : 1));
    sort_at_edges(qaB, bSorted, 0, bAll, bNd, bU, nuB, I, sB, canonical);
    sort_at_edges(qaC, cSorted, cStart, cAll, cNd, cU, nuC, I, sC, canonical);

    /* A14: copyIntersectionsParallel, oacc:7508-7664 */
    double *augB = so_xcalloc(2 * (size_t)nbAug, sizeof(double)), *augC = so_xcalloc(2 * (size_t)ncAug, sizeof(double));
    int *avB = so_xmalloc(sizeof(int) * (size_t)nbAug), *avC = so_xmalloc(sizeof(int) * (size_t)ncAug);
    int *nbB = so_xmalloc(sizeof(int) * (size_t)nbAug), *nbC = so_xmalloc(sizeof(int) * (size_t)ncAug);
    for (int i = 0; i < nbAug; i++) avB[i] = nbB[i] = -100;
    for (int i = 0; i < ncAug; i++) avC[i] = nbC[i] = -100;
    for (int id = 0; id < bSize; id++) {
        int p = id + bNd[id];
        augB[2 * p] = bPoly[2 * id];
        augB[2 * p + 1] = bPoly[2 * id + 1];
        avB[p] = -100;
        nbB[p] = -100;
    }
    for (int id = 0; id < cSize; id++) { /* all overlay vertices (see header) */
        int c2 = 2 * (id + cStart);
        int p = id + cNd[id];
        augC[2 * p] = cPoly[c2];
        augC[2 * p + 1] = cPoly[c2 + 1];
        avC[p] = -100;
        nbC[p] = -100;
    }
    int *bNI = so_xmalloc(sizeof(int) * (size_t)(I > 0 ? I : 1)), *cNI = so_xmalloc(sizeof(int) * (size_t)(I > 0 ? I : 1));
    for (int i = 0; i < I; i++) bNI[i] = cNI[i] = -1;
    for (int id = 0; id < I; id++) {
        int blid = bSorted[sB[id]];
        int clid = cSorted[sC[id]];
        int bid = id, cid = id;
        int prevb = 0, prevc = 0;
        if (id > 0) {
            prevb = bSorted[sB[id - 1]];
            prevc = cSorted[sC[id - 1]];
        }
        if (id <= 0 || bIds[prevb] != bIds[blid]) {
            int ty = types[blid];
            if (ty == 2 || ty == 4 || ty == 6 || ty == 8) {
                int v = bIds[blid] / 2;
                v += bNd[v];
                avB[v] = qaB[blid];
                bNI[blid] = v;
                if (bid < I - 1) blid = bSorted[sB[++bid]];
            }
            int e = bIds[blid] / 2;
            int s0 = bNd[e], s1 = bNd[e + 1];
            for (int i = 0, nB = s0 + e + 1; i < s1 - s0; ++i, ++nB) {
                augB[2 * nB] = inter[2 * blid];
                augB[2 * nB + 1] = inter[2 * blid + 1];
                avB[nB] = qaB[blid];
                bNI[blid] = nB;
                if (bid < I - 1) blid = bSorted[sB[++bid]];
            }
        }
        if (id <= 0 || cIds[prevc] != cIds[clid]) {
            int ty = types[clid];
            if (ty == 3 || ty == 4 || ty == 5 || ty == 7 || ty == 8) {
                int v = cIds[clid] / 2 - cStart;
                v += cNd[v];
                avC[v] = qaC[clid];
                cNI[clid] = v;
                if (cid < I - 1) clid = cSorted[sC[++cid]];
            }
            int e = cIds[clid] / 2 - cStart;
            int s0 = cNd[e], s1 = cNd[e + 1];
            for (int i = 0, nC = s0 + e + 1; i < s1 - s0; i++, ++nC) {
                augC[2 * nC] = inter[2 * clid];
                augC[2 * nC + 1] = inter[2 * clid + 1];
                avC[nC] = qaC[clid];
                cNI[clid] = nC;
                if (cid < I - 1) clid = cSorted[sC[++cid]];
            }
        }
    }
    for (int i = 0; i < I; i++) {
        if (bNI[i] < 0 || cNI[i] < 0) {
            r->status |= 1;
            continue;
        }
        nbB[bNI[i]] = cNI[i] + 1;
        nbC[cNI[i]] = bNI[i] + 1;
    }
    free(bNI); free(cNI);

    /* A15: calculateInitLabelParallel, oacc:7925-7988 */
    int *labB = so_xcalloc((size_t)nbAug, sizeof(int)), *labC = so_xcalloc((size_t)ncAug, sizeof(int));
    int lab_undef = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : lab_undef)
    for (int bid = 0; bid < bSize; ++bid) {
        int start = bid + bNd[bid], end = bid + bNd[bid + 1] + 1;
        for (int i = start; i < end; i++) {
            labB[i] = 0;
            if (avB[i] == -100) continue;
            double cx = augB[2 * i], cy = augB[2 * i + 1];
            int tm = (i - 1 == -1) ? nbAug - 1 : (i - 1 == nbAug ? 0 : i - 1);
            double pMx = augB[2 * tm], pMy = augB[2 * tm + 1];
            int pMN = avB[tm] != -100 ? nbB[tm] - 1 : -100;
            int tp = (i + 1 == nbAug) ? 0 : i + 1;
            double pPx = augB[2 * tp], pPy = augB[2 * tp + 1];
            int pPN = avB[tp] != -100 ? nbB[tp] - 1 : -100;
            int nId = nbB[i] - 1;
            if (nId < 0) {
                lab_undef |= 1;
                continue;
            }
            int qm = (nId - 1 == -1) ? ncAug - 1 : (nId - 1 == ncAug ? 0 : nId - 1);
            int qp = (nId + 1 == ncAug) ? 0 : (nId + 1 == -1 ? ncAug - 1 : nId + 1);
            int types2[2];
            int qs[2] = {qm, qp};
            for (int s = 0; s < 2; s++) {
                int q = qs[s];
                double qx = augC[2 * q], qy = augC[2 * q + 1];
                int rt;
                if (pMN != -100 && pMN == q) rt = 2;
                else if (pPN != -100 && pPN == q) rt = 3;
                else {
                    double s1 = so_A(qx, qy, pMx, pMy, cx, cy);
                    double s2 = so_A(qx, qy, cx, cy, pPx, pPy);
                    double s3 = so_A(pMx, pMy, cx, cy, pPx, pPy);
                    if (s3 > 0) rt = (s1 > 0 && s2 > 0) ? 0 : 1;
                    else rt = (s1 < 0 && s2 < 0) ? 1 : 0;
                }
                types2[s] = rt;
            }
            int qM = types2[0], qP = types2[1], lab;
            /* getInitialLabel, oacc:7896-7920 */
            if ((qM == 0 && qP == 1) || (qM == 1 && qP == 0)) lab = 1;
            else if ((qM == 0 && qP == 0) || (qM == 1 && qP == 1)) lab = 2;
            else if ((qP == 3 && qM == 1) || (qM == 3 && qP == 1)) lab = 3;
            else if ((qP == 3 && qM == 0) || (qM == 3 && qP == 0)) lab = 4;
            else if ((qP == 3 && qM == 2) || (qM == 3 && qP == 2)) lab = 5;
            else if ((qM == 2 && qP == 1) || (qP == 2 && qM == 1)) lab = 6;
            else if ((qM == 2 && qP == 0) || (qP == 2 && qM == 0)) lab = 7;
            else lab = -102;
            labB[i] = lab;
            labC[nId] = lab;
        }
    }
    r->status |= lab_undef;
    r->nbaug = nbAug;
    r->ncaug = ncAug;
    r->aug_b = augB;
    r->aug_c = augC;
    r->nbr_b = nbB;
    r->nbr_c = nbC;
    r->av_b = avB;
    r->av_c = avC;
    r->lab_b = labB;
    r->lab_c = labC;

    /* A16/A17 */
    int st = 0;
    so_foster(augB, nbAug, avB, labB, augC, ncAug, nbC, labC, bPoly, bSize, cPoly + 2 * (size_t)cStart,
              cSize, uni, &st, &r->nrings, &r->ring_off, &r->ring_xy);
    r->status |= st;

    free(bSorted); free(cSorted); free(bAll); free(bNd); free(cAll); free(cNd); free(bU); free(cU);
    free(sB); free(sC);
    return r;
}

void so_result_free(so_result *r) {
    if (!r) return;
    free(r->pb); free(r->pc); free(r->pt); free(r->inter); free(r->qa_b); free(r->qa_c);
    free(r->aug_b); free(r->aug_c); free(r->nbr_b); free(r->nbr_c); free(r->av_b); free(r->av_c);
    free(r->lab_b); free(r->lab_c); free(r->ring_off); free(r->ring_xy);
    free(r);
}

int so_res_status(const so_result *r) { return r->status; }
int64_t so_res_num_candidates(const so_result *r) { return r->ncand; }
int so_res_num_pairs(const so_result *r) { return r->I; }
const int *so_res_pair_b(const so_result *r) { return r->pb; }
const int *so_res_pair_c(const so_result *r) { return r->pc; }
const int *so_res_pair_type(const so_result *r) { return r->pt; }
const double *so_res_intersections(const so_result *r) { return r->inter; }
const int *so_res_alpha_b(const so_result *r) { return r->qa_b; }
const int *so_res_alpha_c(const so_result *r) { return r->qa_c; }
int so_res_nb_aug(const so_result *r) { return r->nbaug; }
int so_res_nc_aug(const so_result *r) { return r->ncaug; }
const double *so_res_aug_b(const so_result *r) { return r->aug_b; }
const double *so_res_aug_c(const so_result *r) { return r->aug_c; }
const int *so_res_nbr_b(const so_result *r) { return r->nbr_b; }
const int *so_res_nbr_c(const so_result *r) { return r->nbr_c; }
const int *so_res_alphav_b(const so_result *r) { return r->av_b; }
const int *so_res_alphav_c(const so_result *r) { return r->av_c; }
const int *so_res_label_b(const so_result *r) { return r->lab_b; }
const int *so_res_label_c(const so_result *r) { return r->lab_c; }
int so_res_num_rings(const so_result *r) { return r->nrings; }
const int *so_res_ring_off(const so_result *r) { return r->ring_off; }
const double *so_res_ring_xy(const so_result *r) { return r->ring_xy; }

/* ------------------------------------------------------------------ */
/* brute force pair set (checker of the tree's completeness)            */
/* ------------------------------------------------------------------ */

int64_t so_brute_pairs(const so_tree *t, int k, int *b, int *c, int *ty, int64_t cap) {
    int64_t n = 0;
    int cStart = t->coff[k - 1], cs = t->csz[k - 1];
    const double *B = t->bpoly, *C = t->cpoly;
    for (int i = 0; i < t->nb; i++) {
        int bId = 2 * i;
        for (int j = 0; j < cs; j++) {
            int cId = 2 * (cStart + j);
            if (!so_lsmf8(B[bId], B[bId + 1], B[bId + 2], B[bId + 3], C[cId], C[cId + 1], C[cId + 2], C[cId + 3]))
                continue;
            int y = so_lsi(B[bId], B[bId + 1], B[bId + 2], B[bId + 3], C[cId], C[cId + 1], C[cId + 2],
                           C[cId + 3], NULL, NULL);
            if (!y) continue;
            if (n < cap) {
                b[n] = bId;
                c[n] = cId;
                ty[n] = y;
            }
            n++;
        }
    }
    return n;
}

double so_rings_area(const int *off, const double *xy, int nrings) {
    double a = 0.0;
    for (int r = 0; r < nrings; r++) {
        int s = off[r], e = off[r + 1];
        double s2 = 0.0;
        for (int i = s; i < e; i++) {
            int j = (i + 1 < e) ? i + 1 : s;
            s2 += xy[2 * i] * xy[2 * j + 1] - xy[2 * j] * xy[2 * i + 1];
        }
        a += 0.5 * s2;
    }
    return a;
}
