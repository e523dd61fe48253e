// sc_fp64.hpp -- FP64 geometric predicates of the clipping path.
//
// Every expression keeps the operation order of the reference
// (oacc_segment_tree.cpp:5282-5316 LSMF, 5334-5337 A(), 5520-5611
// lineSegmentIntersection2, 7247-7252 alphaToInt, 7847-7920 oracle /
// getInitialLabel) and the library is compiled with -ffp-contract=off, so
// the GPU reproduces the host's IEEE double results bit for bit: no fused
// multiply-add, correctly rounded division.
#pragma once
#include <hip/hip_runtime.h>

namespace sc {

constexpr double EPS = 0.000000001;

__host__ __device__ inline double areaA(double px, double py, double qx, double qy, double rx, double ry) {
    double t1 = (qx - px) * (ry - py);
    double t2 = (qy - py) * (rx - px);
    return t1 - t2;
}

__host__ __device__ inline bool lsmf(double l1p1x, double l1p1y, double l1p2x, double l1p2y, double l2p1x,
                                     double l2p1y, double l2p2x, double l2p2y) {
    double minPX = l1p1x, minPY = l1p1y, maxPX = l1p2x, maxPY = l1p2y;
    double minQX = l2p1x, minQY = l2p1y, maxQX = l2p2x, maxQY = l2p2y;
    if (minPX > l1p2x) { minPX = l1p2x; maxPX = l1p1x; }
    if (minPY > l1p2y) { minPY = l1p2y; maxPY = l1p1y; }
    if (minQX > l2p2x) { minQX = l2p2x; maxQX = l2p1x; }
    if (minQY > l2p2y) { minQY = l2p2y; maxQY = l2p1y; }
    if (minPX > maxQX || maxPX < minQX) return false;
    if (minPY > maxQY || maxPY < minQY) return false;
    return true;
}

// intersection type 0..8 of [b1,b2] x [c1,c2] and the parameters alpha (on b)
// and beta (on c); the collinear branch keeps the reference's beta_is_0 test
// on alpha (oacc:5595)
__host__ __device__ inline int classify(double b1x, double b1y, double b2x, double b2y, double c1x, double c1y,
                                        double c2x, double c2y, double &alpha, double &beta) {
    double AP1 = areaA(b1x, b1y, c1x, c1y, c2x, c2y);
    double AP2 = areaA(b2x, b2y, c1x, c1y, c2x, c2y);
    alpha = 0.0;
    beta = 0.0;
    if (fabs(AP1 - AP2) > EPS) {
        double AQ1 = areaA(c1x, c1y, b1x, b1y, b2x, b2y);
        double AQ2 = areaA(c2x, c2y, b1x, b1y, b2x, b2y);
        alpha = AP1 / (AP1 - AP2);
        beta = AQ1 / (AQ1 - AQ2);
        bool a0 = false, a01 = false, b0 = false, b01 = false;
        if (alpha > EPS && alpha < 1.0 - EPS) a01 = true;
        else if (fabs(alpha) <= EPS) a0 = true;
        if (beta > EPS && beta < 1.0 - EPS) b01 = true;
        else if (fabs(beta) <= EPS) b0 = true;
        if (a01 && b01) return 1;
        if (a0 && b01) return 2;
        if (b0 && a01) return 3;
        if (a0 && b0) return 4;
    } else if (fabs(AP1) < EPS) {
        double dPx = b2x - b1x, dPy = b2y - b1y;
        double dQx = c2x - c1x, dQy = c2y - c1y;
        double PQx = c1x - b1x, PQy = c1y - b1y;
        double m1 = PQx * dPx, m2 = PQy * dPy;
        double n1 = dPx * dPx, n2 = dPy * dPy;
        alpha = (m1 + m2) / (n1 + n2);
        double k1 = PQx * dQx, k2 = PQy * dQy;
        double l1 = dQx * dQx, l2 = dQy * dQy;
        beta = -(k1 + k2) / (l1 + l2);
        bool a0 = false, a01 = false, an = false, b0 = false, b01 = false, bn = false;
        if (alpha > EPS && alpha < 1.0 - EPS) a01 = true;
        else if (fabs(alpha) <= EPS) a0 = true;
        else an = true;
        if (beta > EPS && beta < 1.0 - EPS) b01 = true;
        else if (fabs(alpha) <= EPS) b0 = true;
        else bn = true;
        if (a01 && b01) return 5;
        if (an && b01) return 6;
        if (bn && a01) return 7;
        if (a0 && b0) return 8;
    }
    return 0;
}

// classify() with division-free early rejection, for the stab's many tests:
// in the non-parallel branch a pair can only get a non-zero type when
// alpha = AP1/(AP1-AP2) and beta = AQ1/(AQ1-AQ2) both lie in [-eps, 1-eps).
// A pair with alpha < -0.5 or alpha > 2 (likewise beta) is rejected before the
// divisions; the bounds are tested exactly (x0.5 and x2 are exact, |d| > eps
// excludes subnormals) and the division is monotone, so every rejected pair
// also gets type 0 from classify().  Accepted pairs run classify()'s exact
// arithmetic on the same values, hence the same type.
__host__ __device__ inline int classify_pre(double b1x, double b1y, double b2x, double b2y, double c1x,
                                            double c1y, double c2x, double c2y) {
    double AP1 = areaA(b1x, b1y, c1x, c1y, c2x, c2y);
    double AP2 = areaA(b2x, b2y, c1x, c1y, c2x, c2y);
    double d = AP1 - AP2;
    if (fabs(d) > EPS) {
        if (d > 0 ? (AP1 < -0.5 * d || AP1 > 2.0 * d) : (AP1 > -0.5 * d || AP1 < 2.0 * d)) return 0;
        double AQ1 = areaA(c1x, c1y, b1x, b1y, b2x, b2y);
        double AQ2 = areaA(c2x, c2y, b1x, b1y, b2x, b2y);
        double e = AQ1 - AQ2;
        if (e > 0 ? (AQ1 < -0.5 * e || AQ1 > 2.0 * e) : (AQ1 > -0.5 * e || AQ1 < 2.0 * e)) return 0;
        double alpha = AP1 / d;
        double beta = AQ1 / e;
        bool a0 = false, a01 = false, b0 = false, b01 = false;
        if (alpha > EPS && alpha < 1.0 - EPS) a01 = true;
        else if (fabs(alpha) <= EPS) a0 = true;
        if (beta > EPS && beta < 1.0 - EPS) b01 = true;
        else if (fabs(beta) <= EPS) b0 = true;
        if (a01 && b01) return 1;
        if (a0 && b01) return 2;
        if (b0 && a01) return 3;
        if (a0 && b0) return 4;
        return 0;
    }
    double al, be;
    return classify(b1x, b1y, b2x, b2y, c1x, c1y, c2x, c2y, al, be);
}

// point of an intersection of type t (saveIntersectionsParallel, oacc:7257-7341):
// type 1 at (1-alpha) b1 + alpha b2, types 3,4,5,7,8 at c1, types 2,6 at b1
__host__ __device__ inline void intersection_point(int t, double al, double b1x, double b1y, double b2x, double b2y,
                                                  double c1x, double c1y, double &x, double &y) {
    x = 0.0;
    y = 0.0;
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
}

// LSMF's box of one segment: min/max of its endpoints formed as LSMF forms them
struct Box {
    double minx, maxx, miny, maxy;
};
__device__ inline Box box_of(double ax, double ay, double bx, double by) {
    Box r;
    r.minx = ax;
    r.maxx = bx;
    if (r.minx > bx) {
        r.minx = bx;
        r.maxx = ax;
    }
    r.miny = ay;
    r.maxy = by;
    if (r.miny > by) {
        r.miny = by;
        r.maxy = ay;
    }
    return r;
}
__device__ inline bool box_overlap(const Box &p, const Box &q) {
    return !(p.minx > q.maxx || p.maxx < q.minx) && !(p.miny > q.maxy || p.maxy < q.miny);
}

// The box in single precision rounded outward: it contains the double box, so
// two such boxes that do not overlap come from double boxes that do not
// overlap.  The empty box (+inf, -inf, +inf, -inf) overlaps no finite box.
struct __align__(16) BoxF {
    float minx, maxx, miny, maxy;
};
__device__ inline float f_down(double d) {
    float f = float(d);
    if (double(f) > d) f = nextafterf(f, -INFINITY);
    return f;
}
__device__ inline float f_up(double d) {
    float f = float(d);
    if (double(f) < d) f = nextafterf(f, INFINITY);
    return f;
}
__device__ inline BoxF boxf_of(const Box &b) {
    BoxF r;
    r.minx = f_down(b.minx);
    r.maxx = f_up(b.maxx);
    r.miny = f_down(b.miny);
    r.maxy = f_up(b.maxy);
    return r;
}
__device__ inline BoxF boxf_empty() {
    BoxF r;
    r.minx = INFINITY;
    r.maxx = -INFINITY;
    r.miny = INFINITY;
    r.maxy = -INFINITY;
    return r;
}
// the same test as LSMF's, evaluated without branches
__device__ inline bool boxf_overlap(const BoxF &p, const BoxF &q) {
    return !(p.minx > q.maxx) & !(p.maxx < q.minx) & !(p.miny > q.maxy) & !(p.maxy < q.miny);
}

// The y-extent of LSMF's box rounded outward to single precision.  Two
// segments whose single-precision y-extents are disjoint have disjoint LSMF
// boxes; the empty range (+inf, -inf) meets no finite range.
struct __align__(8) YRangeF {
    float lo, hi;
};
__device__ inline YRangeF yrange_of(double ay, double by) {
    const Box b = box_of(0.0, ay, 0.0, by);
    YRangeF r;
    r.lo = f_down(b.miny);
    r.hi = f_up(b.maxy);
    return r;
}
__device__ inline YRangeF yrange_empty() {
    YRangeF r;
    r.lo = INFINITY;
    r.hi = -INFINITY;
    return r;
}

// alphaToInt (oacc:7247-7252).  The reference converts 1e9*a to int with no
// range check; beyond INT_MAX that is undefined behaviour (on x86 INT_MIN,
// which its 9-digit radix sort then uses as a negative digit index).  Here it
// saturates to INT_MAX and alpha_in_range() lets the caller report the job
// as undefined (status bit 16).
__host__ __device__ inline bool alpha_in_range(double a) { return a < EPS || 1000000000.0 * a < 2147483647.0; }
__host__ __device__ inline int alpha_to_int(double a) {
    if (a < EPS) return 0;
    const double q = 1000000000.0 * a;
    return q < 2147483647.0 ? int(q) : 2147483647;
}

// Initial Foster label of an intersection vertex (calculateInitLabelParallel,
// oracle and getInitialLabel, oacc:7847-7920): where the neighbours Q- and Q+
// of its partner lie relative to (P-, I, P+); pMN / pPN are the partners of
// P- / P+ when those are intersections (-100 otherwise).
__host__ __device__ inline int rel_pos(int pMN, int pPN, int q, double qx, double qy, double pmx, double pmy, double cx,
                              double cy, double ppx, double ppy) {
    if (pMN != -100 && pMN == q) return 2;
    if (pPN != -100 && pPN == q) return 3;
    double s1 = areaA(qx, qy, pmx, pmy, cx, cy);
    double s2 = areaA(qx, qy, cx, cy, ppx, ppy);
    double s3 = areaA(pmx, pmy, cx, cy, ppx, ppy);
    if (s3 > 0) return (s1 > 0 && s2 > 0) ? 0 : 1;
    return (s1 < 0 && s2 < 0) ? 1 : 0;
}

__host__ __device__ inline int label_of(int pMN, int pPN, int qm, int qp, double qmx, double qmy, double qpx, double qpy,
                               double pmx, double pmy, double X, double Y, double ppx, double ppy) {
    int a = rel_pos(pMN, pPN, qm, qmx, qmy, pmx, pmy, X, Y, ppx, ppy);
    int b = rel_pos(pMN, pPN, qp, qpx, qpy, pmx, pmy, X, Y, ppx, ppy);
    if ((a == 0 && b == 1) || (a == 1 && b == 0)) return 1;
    if ((a == 0 && b == 0) || (a == 1 && b == 1)) return 2;
    if ((b == 3 && a == 1) || (a == 3 && b == 1)) return 3;
    if ((b == 3 && a == 0) || (a == 3 && b == 0)) return 4;
    if ((b == 3 && a == 2) || (a == 3 && b == 2)) return 5;
    if ((a == 2 && b == 1) || (b == 2 && a == 1)) return 6;
    if ((a == 2 && b == 0) || (b == 2 && a == 0)) return 7;
    return -102;
}

__host__ __device__ inline bool nondeg_b(int t) { return t == 1 || t == 3 || t == 5 || t == 7; }
__host__ __device__ inline bool nondeg_c(int t) { return t == 1 || t == 2 || t == 6; }
__host__ __device__ inline bool degen_b_first(int t) { return t == 2 || t == 4 || t == 6 || t == 8; }
__host__ __device__ inline bool degen_c_first(int t) { return t == 3 || t == 4 || t == 5 || t == 7 || t == 8; }

} // namespace sc
