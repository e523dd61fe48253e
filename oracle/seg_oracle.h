/*
 * seg_oracle.h -- CPU oracle for the segment-tree polygon clipping path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * multicore path (src/clip.cpp multicore_clipping, clip:1415-2115) run with a
 * single thread.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py may load it, and only as the checker or the timed CPU baseline.
 * The product (libsegclip.so) never links or calls it.
 *
 * Every stage cites the reference function it restates.  Where the reference
 * reads uninitialised or out-of-range memory the oracle implements the
 * intended semantics stated next to the stage; those spots are listed in
 * so_quirks() and in the oracle's header comment of seg_oracle.c.
 */
#ifndef SEG_ORACLE_H
#define SEG_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* OpenMP threads of the oracle's parallel loops (results do not depend on it) */
void so_set_threads(int n);
int so_get_threads(void);

/* one edge record, p_seg_tree/segment_tree/edge.h:141-202 */
typedef struct {
    int id;       /* 2 * vertex offset of the edge's first vertex in its poly array */
    int type;     /* 0 = base, k = overlay k (1-based) */
    double start; /* min x of the edge */
    double end;   /* max x of the edge */
    double y;     /* x of the first endpoint (field named y in the reference) */
} so_edge;

typedef struct so_tree so_tree;
typedef struct so_result so_result;

/* Build the tree over one base ring and K overlay rings (A1-A6).
 * bxy: nb*2 doubles (no closing vertex); cxy: sum(csz)*2 doubles, overlays
 * back to back; csz: K vertex counts.  Returns NULL on invalid input. */
so_tree *so_build(const double *bxy, int nb, const double *cxy, const int *csz, int K);
void so_free(so_tree *t);

/* A1/A2 */
int so_num_edges(const so_tree *t);
const so_edge *so_edges(const so_tree *t);
const double *so_bpoly(const so_tree *t);
const double *so_cpoly(const so_tree *t);
const int *so_coff(const so_tree *t);
/* A3 */
int so_epc(const so_tree *t);
const double *so_elementary(const so_tree *t);
/* A4 */
int so_tree_size(const so_tree *t);
const double *so_node_intervals(const so_tree *t);
/* A5: CSR over nodes [0,T), entries are edge indices in insertion order */
const int64_t *so_cover_off(const so_tree *t);
const int *so_cover_edges(const so_tree *t);
/* A6: CSR over nodes [0,T), entries are edge indices sorted by (id,type) */
const int64_t *so_end_off(const so_tree *t);
const int *so_end_edges(const so_tree *t);
/* per edge leaf-level insertion targets (node index, -1 = not inserted) */
const int *so_after_leaf(const so_tree *t);
const int *so_before_leaf(const so_tree *t);

/* Clip the base against overlay k (1-based), A8-A17.
 * canonical = 0: literal single-thread reference order (stab emission order,
 *               per-edge alpha sort only when an edge has >1 non-degenerate
 *               intersections, oacc:7466-7468).
 * canonical = 1: the order the GPU path defines (pairs in (bId,cId) order,
 *               every edge's intersections sorted by quantised alpha).
 * Returns NULL on failure. */
so_result *so_clip(so_tree *t, int k, int canonical);
/* so_clip for an operation: uni = 0 intersection, 1 union (the UNION switch of
 * foster/polyclip.h:39 and its XORs in labelIntersections2 / createResult2) */
so_result *so_clip_op(so_tree *t, int k, int canonical, int uni);
void so_result_free(so_result *r);

/* status: 0 ok; >0 = reference behaviour undefined on this input (bit set):
 *   1 an intersection was never assigned a slot (copyIntersectionsParallel),
 *   2 ring/array coordinate walk diverged (copyToLinkedList),
 *   4 a traversal did not close (createResult2),
 *   8 id radix sort truncated (ids wider than the digit count used),
 *  16 a quantised alpha or beta beyond the int range (alphaToInt) */
int so_res_status(const so_result *r);
int64_t so_res_num_candidates(const so_result *r);  /* K (with duplicates) */
int so_res_num_pairs(const so_result *r);            /* I */
const int *so_res_pair_b(const so_result *r);        /* bId per unique pair, dedupe order */
const int *so_res_pair_c(const so_result *r);
const int *so_res_pair_type(const so_result *r);
const double *so_res_intersections(const so_result *r); /* 2*I */
const int *so_res_alpha_b(const so_result *r);          /* quantised alpha, I */
const int *so_res_alpha_c(const so_result *r);
int so_res_nb_aug(const so_result *r);  /* bIntersectedCount */
int so_res_nc_aug(const so_result *r);  /* cIntersectedCount */
const double *so_res_aug_b(const so_result *r);     /* intersectedB */
const double *so_res_aug_c(const so_result *r);
const int *so_res_nbr_b(const so_result *r);        /* neighborsB (1-based, -100) */
const int *so_res_nbr_c(const so_result *r);
const int *so_res_alphav_b(const so_result *r);     /* alphaValuesB (-100 = plain) */
const int *so_res_alphav_c(const so_result *r);
const int *so_res_label_b(const so_result *r);      /* initial labels */
const int *so_res_label_c(const so_result *r);
/* output rings after cleanUpResult2: components, vertex counts, coordinates */
int so_res_num_rings(const so_result *r);
const int *so_res_ring_off(const so_result *r);     /* num_rings+1 */
const double *so_res_ring_xy(const so_result *r);   /* 2*total */

/* Brute force: all (b,c) with LSMF and a nonzero lineSegmentIntersection2
 * type over base x overlay k, sorted by (bId,cId).  Returns count; fills up to
 * cap triples. */
int64_t so_brute_pairs(const so_tree *t, int k, int *b, int *c, int *ty, int64_t cap);

/* FP64 predicates exactly as the reference evaluates them (no contraction) */
int so_lsmf(const double *p);            /* p = b1x b1y b2x b2y c1x c1y c2x c2y */
int so_classify(const double *p, double *alpha, double *beta);

/* Shoelace area of rings (sum over components, signed) */
double so_rings_area(const int *off, const double *xy, int nrings);

#ifdef __cplusplus
}
#endif
#endif
