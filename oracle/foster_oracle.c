/*
 * foster_oracle.c -- CPU oracle, Foster/Greiner-Hormann stages (A16-A17).
 *
 * TEST INFRASTRUCTURE ONLY (see seg_oracle.h).  Restates, on index-linked
 * vertex records, the reference's copyToLinkedList (src/clip.cpp:606-695) and
 * labelIntersections2 / createResult2 / cleanUpResult2
 * (foster/polyclip.h:337-709, 717-759, 767-787), for the intersection
 * (UNION = false, polyclip.h:39) and, with uni set, the union, one base copy
 * per overlay.
 *
 * Intended semantics used where the reference is undefined:
 *  - cleanUpResult2 deletes the vertex its iterator stands on and then reads
 *    its next pointer (polyclip.h:778-782); the oracle continues from the
 *    removed vertex's old successor, which is what that read returns.
 *  - a traversal that cannot close (inconsistent labels) is cut after a step
 *    bound and reported through status bit 4.
 */
#include "so_internal.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { NONE = 0, CROSSING, BOUNCING, LEFT_ON, RIGHT_ON, ON_ON, ON_LEFT, ON_RIGHT, DELAYED_CROSSING, DELAYED_BOUNCING };
enum { EXIT = 0, ENTRY = 1, NEITHER = 2 };
enum { LEFT = 0, RIGHT = 1 };

typedef struct {
    double x, y;
    int prev, next, nbr;
    int source, inter, label, enex;
} fv;

typedef struct {
    fv *v;
    int n, cap;
} pool;

typedef struct {
    int root;
    int size;
} poly;

static int new_vertex(pool *p, double x, double y) {
    if (p->n == p->cap) {
        p->cap = p->cap ? 2 * p->cap : 1024;
        p->v = realloc(p->v, sizeof(fv) * (size_t)p->cap);
        if (!p->v) abort();
    }
    fv *V = &p->v[p->n];
    V->x = x;
    V->y = y;
    V->prev = V->next = V->nbr = -1;
    V->source = 0;
    V->inter = 0;
    V->label = NONE;
    V->enex = NEITHER;
    return p->n++;
}

/* polygon::newVertex, polygon.h:395-416 */
static void poly_append(pool *p, poly *P, double x, double y, int source) {
    int v = new_vertex(p, x, y);
    p->v[v].source = source;
    if (P->root < 0) {
        p->v[v].next = p->v[v].prev = v;
        P->root = v;
    } else {
        int r = P->root, last = p->v[r].prev;
        p->v[v].prev = last;
        p->v[v].next = r;
        p->v[last].next = v;
        p->v[r].prev = v;
    }
    P->size++;
}

/* polygon::removeVertex, polygon.h:421-429 */
static void poly_remove(pool *p, poly *P, int V) {
    if (P->root == V) {
        P->root = p->v[V].next;
        if (p->v[P->root].next == P->root) P->root = -1;
    }
    p->v[p->v[V].prev].next = p->v[V].next;
    p->v[p->v[V].next].prev = p->v[V].prev;
}

/* insertVertex(V, curr) without alpha, polygon.h:95-107: V goes after curr */
static void insert_after(pool *p, int V, int curr) {
    curr = p->v[curr].next;
    p->v[p->v[curr].prev].next = V;
    p->v[V].prev = p->v[curr].prev;
    p->v[V].next = curr;
    p->v[curr].prev = V;
}

/* insert V immediately before 'current' (copyToLinkedList, clip:624-631) */
static void insert_before(pool *p, int V, int current) {
    p->v[V].next = current;
    p->v[p->v[current].prev].next = V;
    p->v[V].prev = p->v[current].prev;
    p->v[current].prev = V;
}

static double A3(const pool *p, int a, int b, int c) {
    return so_A(p->v[a].x, p->v[a].y, p->v[b].x, p->v[b].y, p->v[c].x, p->v[c].y);
}

/* polygon::pointInPoly, polygon.h:435-457 */
static int point_in_poly(const pool *p, const poly *P, double rx, double ry) {
    int w = 0;
    if (P->root < 0) return 0;
    int one = P->root;
    do {
        int two = p->v[one].next;
        double p0x = p->v[one].x, p0y = p->v[one].y, p1x = p->v[two].x, p1y = p->v[two].y;
        if ((p0y < ry) != (p1y < ry)) {
            if (p0x >= rx) {
                if (p1x > rx) w = w + 2 * (p1y > p0y) - 1;
                else if ((so_A(p0x, p0y, p1x, p1y, rx, ry) > 0) == (p1y > p0y)) w = w + 2 * (p1y > p0y) - 1;
            } else if (p1x > rx) {
                if ((so_A(p0x, p0y, p1x, p1y, rx, ry) > 0) == (p1y > p0y)) w = w + 2 * (p1y > p0y) - 1;
            }
        }
        one = two;
    } while (one != P->root);
    return (w % 2) != 0;
}

/* polygon::noCrossingVertex(union_case), polygon.h:465-475 */
static int no_crossing_vertex(const pool *p, const poly *P, int uni) {
    int V = P->root;
    do {
        if (p->v[V].inter) {
            if (p->v[V].label == CROSSING || p->v[V].label == DELAYED_CROSSING) return 0;
            if (uni && (p->v[V].label == BOUNCING || p->v[V].label == DELAYED_BOUNCING)) return 0;
        }
        V = p->v[V].next;
    } while (V != P->root);
    return 1;
}

static int all_on_on(const pool *p, const poly *P) { /* polygon.h:462-467 */
    int V = P->root;
    do {
        if (p->v[V].label != ON_ON) return 0;
        V = p->v[V].next;
    } while (V != P->root);
    return 1;
}

/* getNonIntersectionPoint, polygon.h:489-502 */
static int non_intersection_point(const pool *p, const poly *P, double *x, double *y) {
    int V = P->root;
    do {
        if (!p->v[V].inter) {
            *x = p->v[V].x;
            *y = p->v[V].y;
            return 1;
        }
        V = p->v[V].next;
    } while (V != P->root);
    V = P->root;
    do {
        int nx = p->v[V].next, nb = p->v[V].nbr;
        if (nb >= 0 && p->v[nx].nbr != p->v[nb].prev && p->v[nx].nbr != p->v[nb].next) {
            *x = 0.5 * (p->v[V].x + p->v[nx].x);
            *y = 0.5 * (p->v[V].y + p->v[nx].y);
            return 1;
        }
        V = p->v[V].next;
    } while (V != P->root);
    return 0;
}

/* getNonIntersectionVertex, polygon.h:507-523 */
static int non_intersection_vertex(pool *p, const poly *P) {
    int V = P->root;
    do {
        if (!p->v[V].inter) return V;
        V = p->v[V].next;
    } while (V != P->root);
    V = P->root;
    do {
        int nx = p->v[V].next, nb = p->v[V].nbr;
        if (nb >= 0 && p->v[nx].nbr != p->v[nb].prev && p->v[nx].nbr != p->v[nb].next) {
            double x = 0.5 * (p->v[V].x + p->v[nx].x), y = 0.5 * (p->v[V].y + p->v[nx].y);
            int T = new_vertex(p, x, y);
            insert_after(p, T, V);
            return T;
        }
        V = p->v[V].next;
    } while (V != P->root);
    return -1;
}

static int toggle(int s) { return s == ENTRY ? EXIT : (s == EXIT ? ENTRY : s); }

int so_foster(const double *augB, int nbAug, const int *avB, const int *labB, const double *augC,
              int ncAug, const int *nbC, const int *labC, const double *bring, int nb,
              const double *cring, int nc, int uni, int *status, int *nrings, int **ring_off,
              double **ring_xy) {
    uni = uni ? 1 : 0; /* UNION (foster/polyclip.h:39): the operation's XOR switch */
    pool p = {0};
    poly P = {-1, 0}, Q = {-1, 0};
    for (int i = 0; i < nb; i++) poly_append(&p, &P, bring[2 * i], bring[2 * i + 1], 1);
    for (int i = 0; i < nc; i++) poly_append(&p, &Q, cring[2 * i], cring[2 * i + 1], 1);

    /* copyToLinkedList, clip:606-695 */
    int *bptr = malloc(sizeof(int) * (size_t)(nbAug > 0 ? nbAug : 1));
    {
        long i = 0, pi = 0, mx = 2L * nbAug;
        int V = P.root;
        for (int ii = 0; ii <= P.size; ii++) {
            int current = V;
            while (augB[i % mx] != p.v[V].x || augB[(i + 1) % mx] != p.v[V].y) {
                if (i + 1 >= mx || pi >= nbAug) { *status |= 2; goto bail; }
                int t = new_vertex(&p, augB[i], augB[i + 1]);
                p.v[t].label = labB[i / 2];
                p.v[t].inter = 1;
                insert_before(&p, t, current);
                bptr[pi++] = t;
                i += 2;
            }
            if (ii < P.size) {
                if (pi >= nbAug) { *status |= 2; goto bail; }
                bptr[pi++] = V;
                p.v[V].label = labB[i / 2];
                if (avB[i / 2] != -100) p.v[V].inter = 1;
            }
            i += 2;
            V = p.v[current].next;
        }
        if (pi != nbAug) { *status |= 2; goto bail; }
    }
    {
        long i = 0, mx = 2L * ncAug;
        int V = Q.root;
        for (int ii = 0; ii <= Q.size; ii++) {
            int current = V;
            while (augC[i % mx] != p.v[V].x || augC[(i + 1) % mx] != p.v[V].y) {
                if (i + 1 >= mx) { *status |= 2; goto bail; }
                int t = new_vertex(&p, augC[i], augC[i + 1]);
                p.v[t].label = labC[i / 2];
                p.v[t].inter = 1;
                insert_before(&p, t, current);
                int nbi = nbC[i / 2] - 1;
                if (nbi < 0 || nbi >= nbAug) { *status |= 1; goto bail; }
                p.v[t].nbr = bptr[nbi];
                p.v[bptr[nbi]].nbr = t;
                i += 2;
            }
            if (ii < Q.size) {
                p.v[V].label = labC[i / 2];
                if (nbC[i / 2] != -100) {
                    int nbi = nbC[i / 2] - 1;
                    if (nbi < 0 || nbi >= nbAug) { *status |= 1; goto bail; }
                    p.v[V].inter = 1;
                    p.v[V].nbr = bptr[nbi];
                    p.v[bptr[nbi]].nbr = V;
                }
            }
            i += 2;
            V = p.v[current].next;
        }
    }
    /* every intersection vertex must carry a neighbour */
    for (int i = 0; i < p.n; i++)
        if (p.v[i].inter && p.v[i].nbr < 0) { *status |= 1; goto bail; }

    /* ---- labelIntersections2, polyclip.h:337-709 ---- */
    {
        /* 2) intersection chains */
        int root = P.root, I = root;
        do {
            if (p.v[I].inter && (p.v[I].label == LEFT_ON || p.v[I].label == RIGHT_ON)) {
                int x = p.v[I].label == LEFT_ON ? LEFT : RIGHT;
                int X = I, J = I, guard = 0;
                do {
                    p.v[J].label = NONE;
                    if (p.v[J].nbr >= 0) p.v[p.v[J].nbr].label = NONE;
                    J = p.v[J].next;
                    if (++guard > p.n) { *status |= 4; goto bail; }
                } while (p.v[J].label == ON_ON);
                int y = p.v[J].label == ON_LEFT ? LEFT : RIGHT;
                int ct = x != y ? DELAYED_CROSSING : DELAYED_BOUNCING;
                p.v[X].label = ct;
                p.v[J].label = ct;
                if (p.v[X].nbr >= 0) p.v[p.v[X].nbr].label = ct;
                if (p.v[J].nbr >= 0) p.v[p.v[J].nbr].label = ct;
            }
            I = p.v[I].next;
        } while (I != root);
    }

    /* these are collected in the order the reference pushes them */
    int res_roots_cap = 16, nres = 0;
    int *res_roots = malloc(sizeof(int) * (size_t)res_roots_cap);
    poly R_list_tmp;
    (void)R_list_tmp;
#define PUSH_RES(r)                                                                      \
    do {                                                                                 \
        if (nres == res_roots_cap) {                                                     \
            res_roots_cap *= 2;                                                          \
            res_roots = realloc(res_roots, sizeof(int) * (size_t)res_roots_cap);        \
        }                                                                                \
        res_roots[nres++] = (r);                                                         \
    } while (0)

    int noInter[2] = {0, 0}, ident[2] = {0, 0};
    {
        /* 3.5) special cases */
        for (int s = 0; s < 2; s++) {
            poly *A = s == 0 ? &P : &Q, *B = s == 0 ? &Q : &P;
            if (no_crossing_vertex(&p, A, uni)) {
                noInter[s] = 1;
                if (all_on_on(&p, A)) ident[s] = 1;
                else {
                    double px = 0, py = 0;
                    if (!non_intersection_point(&p, A, &px, &py)) { *status |= 4; goto bail2; }
                    if (point_in_poly(&p, B, px, py) ^ uni) PUSH_RES(A->root); /* polyclip.h:470 */
                }
            }
        }
        if (ident[0] && ident[1]) {
            int nbv = p.v[P.root].nbr, V = Q.root, found = 0;
            do {
                if (V == nbv) found = 1;
                V = p.v[V].next;
            } while (V != Q.root);
            if (found) PUSH_RES(P.root);
        }
    }

    /* 4) entry/exit flags */
    int *split[2] = {NULL, NULL}, nsplit[2] = {0, 0};
    int *cross[2] = {NULL, NULL}, ncross[2] = {0, 0};
    split[0] = malloc(sizeof(int) * (size_t)p.n + 8);
    split[1] = malloc(sizeof(int) * (size_t)p.n + 8);
    cross[0] = malloc(sizeof(int) * (size_t)p.n + 8);
    cross[1] = malloc(sizeof(int) * (size_t)p.n + 8);
    for (int s = 0; s < 2; s++) {
        poly *A = s == 0 ? &P : &Q, *B = s == 0 ? &Q : &P;
        if (noInter[s]) continue;
        int V = non_intersection_vertex(&p, A);
        if (V < 0) { *status |= 4; goto bail3; }
        int st = ENTRY;
        if (point_in_poly(&p, B, p.v[V].x, p.v[V].y)) st = toggle(st);
        int first_chain = 1;
        int I = V;
        if (!p.v[I].inter) {
            do { I = p.v[I].next; } while (!p.v[I].inter && I != V);
        }
        if (p.v[I].inter) {
            int start = I;
            (void)start;
            for (;;) {
                int lab = p.v[I].label;
                if (lab == CROSSING) {
                    p.v[I].enex = st;
                    st = toggle(st);
                }
                if (lab == BOUNCING && ((st == EXIT) ^ uni)) split[s][nsplit[s]++] = I;
                if (lab == DELAYED_CROSSING) {
                    p.v[I].enex = st;
                    if (first_chain) {
                        if ((st == EXIT) ^ uni) p.v[I].label = CROSSING;
                        first_chain = 0;
                    } else {
                        if ((st == ENTRY) ^ uni) p.v[I].label = CROSSING;
                        first_chain = 1;
                        st = toggle(st);
                    }
                }
                if (lab == DELAYED_BOUNCING) {
                    p.v[I].enex = st;
                    if (first_chain) {
                        if ((st == EXIT) ^ uni) cross[s][ncross[s]++] = I;
                        first_chain = 0;
                    } else {
                        if ((st == ENTRY) ^ uni) cross[s][ncross[s]++] = I;
                        first_chain = 1;
                    }
                    st = toggle(st);
                }
                do { I = p.v[I].next; } while (!p.v[I].inter && I != V);
                if (I == V) break;
            }
        }
    }

    /* 5) split bouncing vertex pairs */
    for (int a = 0; a < nsplit[0]; a++) {
        int IP = split[0][a], IQ = p.v[IP].nbr, found = 0;
        for (int b = 0; b < nsplit[1]; b++)
            if (split[1][b] == IQ) found = 1;
        if (!found) continue;
        int VP = new_vertex(&p, p.v[IP].x, p.v[IP].y);
        int VQ = new_vertex(&p, p.v[IQ].x, p.v[IQ].y);
        double sP = A3(&p, p.v[IP].prev, IP, p.v[IP].next);
        double sQ = A3(&p, p.v[IQ].prev, IQ, p.v[IQ].next);
        if (sP * sQ > 0) {
            p.v[IP].nbr = VQ; p.v[VQ].nbr = IP; p.v[IP].inter = p.v[VQ].inter = 1;
            p.v[IQ].nbr = VP; p.v[VP].nbr = IQ; p.v[IQ].inter = p.v[VP].inter = 1;
        } else {
            p.v[VP].nbr = VQ; p.v[VQ].nbr = VP; p.v[VP].inter = p.v[VQ].inter = 1;
        }
        insert_after(&p, VP, IP);
        insert_after(&p, VQ, IQ);
        if (!uni) { p.v[IP].enex = EXIT; p.v[VP].enex = ENTRY; p.v[IQ].enex = EXIT; p.v[VQ].enex = ENTRY; }
        else { p.v[IP].enex = ENTRY; p.v[VP].enex = EXIT; p.v[IQ].enex = ENTRY; p.v[VQ].enex = EXIT; }
        p.v[IP].label = p.v[VP].label = p.v[IQ].label = p.v[VQ].label = CROSSING;
    }
    /* 6) CROSSING candidates */
    for (int a = 0; a < ncross[0]; a++) {
        int IP = cross[0][a], IQ = p.v[IP].nbr;
        for (int b = 0; b < ncross[1]; b++)
            if (cross[1][b] == IQ) {
                p.v[IP].label = CROSSING;
                p.v[IQ].label = CROSSING;
            }
    }

    /* ---- createResult2, polyclip.h:717-759 ---- */
    {
        int root = P.root;
        int I = root;
        int have = p.v[I].inter && p.v[I].label == CROSSING;
        if (!have) {
            do { I = p.v[I].next; } while (!(p.v[I].inter && p.v[I].label == CROSSING) && I != root);
            have = I != root;
        }
        long bound = 4L * p.n + 16;
        while (have) {
            poly R = {-1, 0};
            int V = I;
            p.v[V].inter = 0;
            long steps = 0;
            do {
                int st = toggle(p.v[V].enex);
                do {
                    V = ((st == EXIT) ^ uni) ? p.v[V].next : p.v[V].prev;
                    p.v[V].inter = 0;
                    poly_append(&p, &R, p.v[V].x, p.v[V].y, 0);
                    if (++steps > bound) { *status |= 4; goto bail3; }
                } while (!(p.v[V].enex == st) && V != I);
                if (V != I) {
                    V = p.v[V].nbr;
                    if (V < 0) { *status |= 4; goto bail3; }
                    p.v[V].inter = 0;
                }
            } while (V != I);
            PUSH_RES(R.root);
            do { I = p.v[I].next; } while (!(p.v[I].inter && p.v[I].label == CROSSING) && I != root);
            have = I != root;
        }
    }

    /* ---- cleanUpResult2, polyclip.h:767-787 ---- */
    for (int r = 0; r < nres; r++) {
        poly R = {res_roots[r], 0};
        while (R.root >= 0 && fabs(A3(&p, p.v[R.root].prev, R.root, p.v[R.root].next)) < SO_EPS)
            poly_remove(&p, &R, R.root);
        if (R.root >= 0) {
            int it_root = R.root, V = R.root;
            do {
                int nx = p.v[V].next;
                if (fabs(A3(&p, p.v[V].prev, V, p.v[V].next)) < SO_EPS) poly_remove(&p, &R, V);
                V = nx;
            } while (V != it_root);
        }
        res_roots[r] = R.root;
    }

    {
        int total = 0, nr = 0;
        for (int r = 0; r < nres; r++) {
            if (res_roots[r] < 0) continue;
            nr++;
            int V = res_roots[r];
            do { total++; V = p.v[V].next; } while (V != res_roots[r]);
        }
        *nrings = nr;
        *ring_off = malloc(sizeof(int) * (size_t)(nr + 1));
        *ring_xy = malloc(sizeof(double) * 2 * (size_t)(total > 0 ? total : 1));
        int q = 0, k = 0;
        (*ring_off)[0] = 0;
        for (int r = 0; r < nres; r++) {
            if (res_roots[r] < 0) continue;
            int V = res_roots[r];
            do {
                (*ring_xy)[2 * q] = p.v[V].x;
                (*ring_xy)[2 * q + 1] = p.v[V].y;
                q++;
                V = p.v[V].next;
            } while (V != res_roots[r]);
            (*ring_off)[++k] = q;
        }
    }
    free(split[0]); free(split[1]); free(cross[0]); free(cross[1]);
    free(res_roots);
    free(bptr);
    free(p.v);
    return 0;

bail3:
    free(split[0]); free(split[1]); free(cross[0]); free(cross[1]);
bail2:
    free(res_roots);
bail:
    free(bptr);
    free(p.v);
    *nrings = 0;
    *ring_off = calloc(1, sizeof(int));
    *ring_xy = calloc(2, sizeof(double));
    return -1;
#undef PUSH_RES
}
