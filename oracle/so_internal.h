/* so_internal.h -- shared internals of the CPU oracle (test infrastructure only). */
#ifndef SO_INTERNAL_H
#define SO_INTERNAL_H
#include "seg_oracle.h"
#include <stddef.h>

#define SO_EPS 0.000000001 /* p_seg_tree/util/constants.h:13 */
#define SO_EPS_POSITIONS 9 /* p_seg_tree/util/constants.h:14 */

struct so_tree {
    int nb, K, nct;      /* base vertices, overlays, total overlay vertices */
    int *csz, *coff;     /* overlay sizes; cSizes prefix (clip:186-190) */
    double *bpoly;       /* (nb+1)*2, closing vertex duplicated */
    double *cpoly;       /* (nct+K)*2 */
    int ne;              /* edges = nb + nct */
    so_edge *edges;
    double minv, maxv;
    /* A3 */
    double *ea;
    int epc, nleaves, T;
    /* A4 */
    double *niv;
    /* A5 */
    int64_t *cov_off;
    int *cov_e;
    /* A6 */
    int64_t *end_off;
    int *end_e;
    int *after_leaf, *before_leaf;
};

struct so_result {
    int status;
    int64_t ncand;
    int I;
    int *pb, *pc, *pt;          /* unique pairs in dedupe order */
    double *inter;              /* 2*I */
    int *qa_b, *qa_c;           /* quantised alpha */
    int nbaug, ncaug;
    double *aug_b, *aug_c;
    int *nbr_b, *nbr_c, *av_b, *av_c, *lab_b, *lab_c;
    int nrings;
    int *ring_off;
    double *ring_xy;
};

/* FP64 kernels of the reference, evaluated in the reference's operation order */
double so_A(double px, double py, double qx, double qy, double rx, double ry);
int so_lsmf8(double b1x, double b1y, double b2x, double b2y, double c1x, double c1y, double c2x,
             double c2y);
int so_lsi(double b1x, double b1y, double b2x, double b2y, double c1x, double c1y, double c2x,
           double c2y, double *alpha, double *beta);

/* Foster labelling / traversal / clean-up on the augmented arrays (foster_oracle.c) */
int so_foster(const double *aug_b, int nbaug, const int *av_b, const int *lab_b, const double *aug_c,
              int ncaug, const int *nbr_c, const int *lab_c, const double *bring, int nb,
              const double *cring, int nc, int uni, int *status, int *nrings, int **ring_off,
              double **ring_xy);

void *so_xmalloc(size_t n);
void *so_xcalloc(size_t n, size_t s);
#endif
