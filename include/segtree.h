/*
 * segtree.h -- segclip::SegTree, the reference's SegTree interface
 * (p_seg_tree/segment_tree/segment_tree.h:204-420, implemented upstream in
 * p_seg_tree/oacc_segment_tree.cpp) as a C++ facade over libsegclip's C ABI.
 *
 * Method names, argument order and argument meaning follow the reference, so
 * its drivers (src/clip.cpp multicore_clipping clip:1501-2062 and the tree
 * calls of hybrid_clipping clip:854-993) switch by including this header and
 * `using segclip::SegTree;`.  Differences a caller sees:
 *  - the tree is built on the GPU at the first call that carries polygon
 *    coordinates (copyInputDataToGPU or saveIntersectionsIdsMulticore): the
 *    Edge records hold only x-extents, so treeSize is valid from then on and
 *    the tree-stage calls before it only record the edges;
 *  - there is no `lock` array: the host locks of the reference's cover-list
 *    insertion do not exist, so the driver's omp_destroy_lock loop goes;
 *  - saveIntersectionsIdsMulticore (and countIntersectionsParallel +
 *    polyClipArrayGPUParallel) return every intersecting pair once, in
 *    (bId, cId) order; the driver's radix sort and duplicate removal keep it;
 *    iCount is a host array and holds the total, not per-node counts;
 *  - errors throw segclip::Error (code + message) instead of printing and
 *    exiting.
 * REAL is double.  Edge is any type with the reference Edge's fields
 * (`int id, type; double start, end`).
 */
#ifndef SEGCLIP_SEGTREE_H
#define SEGCLIP_SEGTREE_H

#include "segclip.h"

#include <stdexcept>
#include <string>
#include <vector>

namespace segclip {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

class SegTree {
  public:
    // public state read by the drivers (segment_tree.h:37-60)
    int numEdges = 0, intervalCount = 0;
    double minInterval = 0.0, maxInterval = 0.0;
    int treeSize = 0, numLeaves = 0, POLY_TYPE_COUNT = 0;

    SegTree();
    SegTree(int aNumEdges, int ic, double minVal, double maxVal);
    ~SegTree();
    SegTree(const SegTree &) = delete;
    SegTree &operator=(const SegTree &) = delete;

    // ---- tree construction, multicore driver (clip:1501-1611) ----
    template <class EdgeT> void constructElementaryIntervalsArrayMulticore(EdgeT *intervals) {
        record_edges(intervals);
    }
    void init() {}
    void constructTreeIntervalArrayMulticore() {}
    template <class EdgeT> void insertAllToEdgeListMulticore(EdgeT *, int) {}
    template <class EdgeT> void insertAllParallel(EdgeT *) {}

    // ---- tree construction, hybrid driver (clip:854-993) ----
    template <class EdgeT>
    void copyInputDataToGPU(double *bPoly, double *cPoly, int bSize, int cSize, int cPolysSize, int cPolyCount,
                            int ic, EdgeT *edges) {
        intervalCount = ic;
        record_edges(edges);
        (void)cSize;
        (void)cPolysSize;
        (void)cPolyCount;
        load(bPoly, cPoly, bSize);
    }
    void constructElementaryIntervalsArrayGPU() { build(); }
    void initHybrid() {}
    void constructTreeIntervalArray() {}
    void insertAllToEdgeListParallel(int) {}
    void constructCoverListGPU() {}
    void countEndlistCoverListSizes() {}
    void copyNodeCoverListToGPU(double *, double *, int, int) {}

    // ---- per overlay (clip:1713-1990) ----
    // pairs (bId, cId, type) of overlay cType against the base
    void saveIntersectionsIdsMulticore(double *bPoly, double *cPoly, int bSize, int cSize, int bType, int cType,
                                       std::vector<int> &bPolLineIds, std::vector<int> &cPolLineIds,
                                       std::vector<int> &intersectTypesDup);
    // hybrid driver (clip:1024-1090): iCount (treeSize ints, host) then the
    // driver's exclusive scan sizes the pair arrays; here iCount[0] carries
    // the pair total of overlay cType and every other entry is 0, so the
    // scanned iCount[treeSize-1] is that total, and polyClipArrayGPUParallel
    // writes every pair once, in (bId, cId) order
    void countIntersectionsParallel(double *bPoly, double *cPoly, int bSize, int cSize, int cPolyCount, int bType,
                                    int cType, int *iCount);
    void polyClipArrayGPUParallel(double *bPoly, double *cPoly, int bSize, int cSize, int bType, int cType,
                                  int *iCount, int *bPolLineIds, int *cPolLineIds, int *intersectTypes);
    void saveIntersectionsParallel(double *bPol, double *cPol, int *bPolyLineIds, int *cPolyLineIds, int polySize,
                                   double *intersections, int *intersectTypes, int *intersectAlphaValuesB,
                                   int *intersectAlphaValuesC);
    void sortIntersectionsAtEdgesParallel(int *intersectAlphaValues, int *sortedIndicies, int cStart,
                                          int *allIntersectCounts, int *nonDegenIntersectCounts,
                                          int *polLineUniqueIds, int uniqueIntersectCount, int intersectCount,
                                          int *sortedAlphaIndicies);
    void copyIntersectionsParallel(double *bPoly, double *cPoly, int bSize, int cSize, int cStart,
                                   int *bPolyLineIds, int *cPolyLineIds, int polySize,
                                   int *bPolLineIdsSortedIndex, int *cPolLineIdsSortedIndex,
                                   int *sortedBPolyIdIndicies, int *sortedCPolyIdIndicies, double *intersections,
                                   int *intersectTypes, int *bAllIntersectCounts, int *cAllIntersectCounts,
                                   int *bNonDegenIntersectCountPS, int *cNonDegenIntersectCountPS,
                                   int *intersectAlphaValuesB, int *intersectAlphaValuesC, double *intersectedB,
                                   double *intersectedC, int *neighborsB, int *neighborsC, int *alphaValuesB,
                                   int *alphaValuesC);
    void calculateInitLabelParallel(int bSize, int *bNonDegenIntersectCounts, double *intersectedB,
                                    double *intersectedC, int *alphaValuesB, int *neighborsB, int bIntersectedCount,
                                    int cIntersectedCount, int *bInitLabels, int *cInitLabels);

    segclip_ctx *context() { return ctx_; }

  private:
    template <class EdgeT> void record_edges(const EdgeT *e) {
        std::vector<int> counts;
        for (int i = 0; i < intervalCount; ++i) {
            int t = e[i].type;
            if (t < 0) throw Error(SEGCLIP_EINVAL, "edge with a negative type");
            if (int(counts.size()) <= t) counts.resize(size_t(t) + 1, 0);
            counts[t]++;
        }
        set_ring_sizes(counts);
    }
    void set_ring_sizes(const std::vector<int> &counts);
    int64_t pair_count(double *bPoly, double *cPoly, int bSize, int bType, int cType, const char *what);
    void fill_pairs(int cType, int64_t n, int *b, int *c, int *t, const char *what);
    void load(const double *bPoly, const double *cPoly, int bSize);
    void build();
    void check(int rc, const char *what);

    segclip_ctx *ctx_ = nullptr;
    std::vector<int> csz_;
    int nb_ = 0;
    bool loaded_ = false, built_ = false;
};

} // namespace segclip

#endif
