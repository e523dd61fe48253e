// sc_sparse.hpp -- the base ring of a one-to-many batch, kept sparse.
//
// In a one-to-many batch every job's augmented base ring is the same base
// ring with that job's intersections inserted.  Instead of one dense copy per
// job, only the intersection slots are stored: per job, its pairs in
// (edge, alpha) order ("positions"), each with its local slot in the virtual
// augmented ring (ascending within the job), its alpha value, neighbour link,
// label and point.  Any other slot q is a plain base vertex
// v = q - (non-degenerate slots before q).  A slot that holds a leading
// degenerate intersection is the source vertex's own slot (its point is the
// vertex), exactly as in the dense layout.
#pragma once
#include "sc_common.hpp"
#include "sc_fp64.hpp"

namespace sc {

struct SparseP {
    const int *P0;    // J+1 position offsets (the jobs' pair ranges)
    const int *AB;    // J+1 virtual slot offsets
    const int *nbv;   // per job: base vertices
    const int *bvoff; // per job: base ring vertex offset in bpoly
    const int *slot;  // per position: local slot (ascending within a job), -1 for none
    const int *ndx;   // per position: non-degenerate slots of the job before it
    const int *edge;  // per position: base edge within the ring
    const double *x, *y;
    const double *bpoly;
};

struct SlotInfo {
    int pos; // intersection position, or -1 for a plain vertex
    double x, y;
};

// first position of job j whose slot is >= q
__device__ inline int sp_lower(const SparseP &s, int j, int q) {
    int lo = s.P0[j], hi = s.P0[j + 1];
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (s.slot[mid] < q) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// position holding local slot q of job j, or -1
__device__ inline int sp_find(const SparseP &s, int j, int q) {
    int k = sp_lower(s, j, q);
    return (k < s.P0[j + 1] && s.slot[k] == q) ? k : -1;
}

// content of local slot q of job j
__device__ inline SlotInfo sp_slot(const SparseP &s, int j, int q) {
    SlotInfo r;
    const int k = sp_lower(s, j, q);
    if (k < s.P0[j + 1] && s.slot[k] == q) {
        r.pos = k;
        r.x = s.x[k];
        r.y = s.y[k];
        return r;
    }
    const int before = k < s.P0[j + 1] ? s.ndx[k] : (s.AB[j + 1] - s.AB[j]) - s.nbv[j];
    int v = q - before;
    if (v < 0) v = 0; // only for inconsistent (flagged) input
    if (v >= s.nbv[j]) v = s.nbv[j] - 1;
    r.pos = -1;
    const double *P = s.bpoly + 2 * (int64_t(s.bvoff[j]) + v);
    r.x = P[0];
    r.y = P[1];
    return r;
}

// the winding test of polygon::pointInPoly (foster/polygon.h:435-457) for
// one edge p0 -> p1 and the query point r
__device__ inline int wind_edge(double p0x, double p0y, double p1x, double p1y, double rx, double ry) {
    int w = 0;
    if ((p0y < ry) != (p1y < ry)) {
        const int sg = 2 * (p1y > p0y) - 1;
        if (p0x >= rx) {
            if (p1x > rx) w = sg;
            else if ((areaA(p0x, p0y, p1x, p1y, rx, ry) > 0) == (p1y > p0y)) w = sg;
        } else if (p1x > rx) {
            if ((areaA(p0x, p0y, p1x, p1y, rx, ry) > 0) == (p1y > p0y)) w = sg;
        }
    }
    return w;
}

} // namespace sc
