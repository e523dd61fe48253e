// sc_segtree.cpp -- segclip::SegTree (include/segtree.h): the reference's
// SegTree interface (p_seg_tree/segment_tree/segment_tree.h:204-420) over the
// C ABI.  Host code only: every stage runs in libsegclip's kernels.
#include "../../include/segtree.h"

#include <algorithm>

namespace segclip {

SegTree::SegTree() = default;

SegTree::SegTree(int aNumEdges, int ic, double minVal, double maxVal)
    : numEdges(aNumEdges), intervalCount(ic), minInterval(minVal), maxInterval(maxVal) {}

SegTree::~SegTree() {
    if (ctx_) segclip_destroy(ctx_);
}

void SegTree::check(int rc, const char *what) {
    if (rc == SEGCLIP_OK) return;
    std::string msg = std::string("SegTree::") + what + ": ";
    msg += ctx_ ? segclip_last_error(ctx_) : "no context";
    throw Error(rc, msg);
}

// base = type 0, overlay k = type k (edge.h:141-202)
void SegTree::set_ring_sizes(const std::vector<int> &counts) {
    if (counts.size() < 2 || counts[0] < 3) throw Error(SEGCLIP_EINVAL, "SegTree: need a base and an overlay");
    nb_ = counts[0];
    csz_.assign(counts.begin() + 1, counts.end());
    for (int s : csz_)
        if (s < 3) throw Error(SEGCLIP_EINVAL, "SegTree: an overlay has fewer than 3 edges");
    if (POLY_TYPE_COUNT == 0) POLY_TYPE_COUNT = int(counts.size());
}

// bPoly: bSize+1 vertices (closing vertex repeated); cPoly: every overlay with
// its closing vertex (clip:186-190, 285-294)
void SegTree::load(const double *bPoly, const double *cPoly, int bSize) {
    if (csz_.empty()) throw Error(SEGCLIP_ESTATE, "SegTree: edges not recorded");
    if (bSize != nb_) throw Error(SEGCLIP_EINVAL, "SegTree: bSize does not match the base edges");
    if (!ctx_) check(segclip_create(&ctx_, 0, nullptr), "create");
    std::vector<double> bxy(bPoly, bPoly + 2 * size_t(bSize));
    std::vector<double> cxy;
    size_t at = 0;
    for (int s : csz_) {
        cxy.insert(cxy.end(), cPoly + 2 * at, cPoly + 2 * (at + size_t(s)));
        at += size_t(s) + 1;
    }
    check(segclip_set_polygons(ctx_, bxy.data(), bSize, cxy.data(), csz_.data(), int(csz_.size())),
          "set_polygons");
    loaded_ = true;
    built_ = false;
}

void SegTree::build() {
    if (!loaded_) throw Error(SEGCLIP_ESTATE, "SegTree: polygons not loaded");
    if (built_) return;
    check(segclip_build_tree(ctx_), "build_tree");
    check(segclip_find_candidates(ctx_), "find_candidates");
    segclip_counts c;
    check(segclip_get_counts(ctx_, &c), "get_counts");
    treeSize = int(c.tree_size);
    numLeaves = treeSize / 2;
    built_ = true;
}

void SegTree::saveIntersectionsIdsMulticore(double *bPoly, double *cPoly, int bSize, int cSize, int bType,
                                            int cType, std::vector<int> &bPolLineIds,
                                            std::vector<int> &cPolLineIds, std::vector<int> &intersectTypesDup) {
    (void)cSize;
    const int64_t need = pair_count(bPoly, cPoly, bSize, bType, cType, "saveIntersectionsIdsMulticore");
    const size_t n = size_t(need);
    bPolLineIds.resize(n);
    cPolLineIds.resize(n);
    intersectTypesDup.resize(n);
    fill_pairs(cType, need, bPolLineIds.data(), cPolLineIds.data(), intersectTypesDup.data(),
               "saveIntersectionsIdsMulticore");
}

int64_t SegTree::pair_count(double *bPoly, double *cPoly, int bSize, int bType, int cType, const char *what) {
    if (bType != 0) throw Error(SEGCLIP_EINVAL, "SegTree: the base type is 0");
    if (!loaded_) load(bPoly, cPoly, bSize);
    build();
    int64_t need = 0;
    int rc = segclip_copy_pairs(ctx_, cType, nullptr, nullptr, nullptr, 0, &need);
    if (rc != SEGCLIP_OK && rc != SEGCLIP_ESPACE) check(rc, what);
    return need;
}

void SegTree::countIntersectionsParallel(double *bPoly, double *cPoly, int bSize, int cSize, int cPolyCount,
                                         int bType, int cType, int *iCount) {
    (void)cSize;
    (void)cPolyCount;
    const int64_t n = pair_count(bPoly, cPoly, bSize, bType, cType, "countIntersectionsParallel");
    if (n > int64_t(0x7fffffff)) throw Error(SEGCLIP_ERANGE, "SegTree: pair count exceeds int");
    std::fill(iCount, iCount + treeSize, 0);
    iCount[0] = int(n);
}

void SegTree::polyClipArrayGPUParallel(double *bPoly, double *cPoly, int bSize, int cSize, int bType, int cType,
                                       int *iCount, int *bPolLineIds, int *cPolLineIds, int *intersectTypes) {
    (void)cSize;
    const int64_t n = pair_count(bPoly, cPoly, bSize, bType, cType, "polyClipArrayGPUParallel");
    if (iCount && treeSize > 0 && int64_t(iCount[treeSize - 1]) != n)
        throw Error(SEGCLIP_EINVAL, "SegTree: iCount does not hold the scanned pair total");
    fill_pairs(cType, n, bPolLineIds, cPolLineIds, intersectTypes, "polyClipArrayGPUParallel");
}

// the n pairs of overlay cType with their types (lineSegmentIntersection, as
// the stab classified them)
void SegTree::fill_pairs(int cType, int64_t n, int *b, int *c, int *t, const char *what) {
    if (n <= 0) return;
    check(segclip_copy_pairs(ctx_, cType, b, c, nullptr, n, nullptr), what);
    std::vector<double> xy(2 * size_t(n));
    std::vector<int32_t> qa(static_cast<size_t>(n)), qb(static_cast<size_t>(n));
    check(segclip_save_intersections_ids(ctx_, b, c, int(n), xy.data(), t, qa.data(), qb.data()), what);
}

void SegTree::saveIntersectionsParallel(double *, double *, int *bPolyLineIds, int *cPolyLineIds, int polySize,
                                        double *intersections, int *intersectTypes, int *intersectAlphaValuesB,
                                        int *intersectAlphaValuesC) {
    if (!loaded_) throw Error(SEGCLIP_ESTATE, "SegTree: polygons not loaded");
    check(segclip_save_intersections_ids(ctx_, bPolyLineIds, cPolyLineIds, polySize, intersections, intersectTypes,
                                         intersectAlphaValuesB, intersectAlphaValuesC),
          "saveIntersectionsParallel");
}

void SegTree::sortIntersectionsAtEdgesParallel(int *intersectAlphaValues, int *sortedIndicies, int cStart,
                                               int *allIntersectCounts, int *nonDegenIntersectCounts,
                                               int *polLineUniqueIds, int uniqueIntersectCount, int intersectCount,
                                               int *sortedAlphaIndicies) {
    if (!loaded_) throw Error(SEGCLIP_ESTATE, "SegTree: polygons not loaded");
    // the count arrays cover every edge up to the last listed one
    int last = -1;
    for (int i = 0; i < uniqueIntersectCount; ++i) last = std::max(last, polLineUniqueIds[i] / 2 - cStart);
    check(segclip_sort_intersections_at_edges(ctx_, intersectAlphaValues, sortedIndicies, cStart, allIntersectCounts,
                                              nonDegenIntersectCounts, last + 2, polLineUniqueIds,
                                              uniqueIntersectCount, intersectCount, sortedAlphaIndicies),
          "sortIntersectionsAtEdgesParallel");
}

void SegTree::copyIntersectionsParallel(double *, double *, int bSize, int cSize, int cStart, int *bPolyLineIds,
                                        int *cPolyLineIds, int polySize, int *bPolLineIdsSortedIndex,
                                        int *cPolLineIdsSortedIndex, int *sortedBPolyIdIndicies,
                                        int *sortedCPolyIdIndicies, double *intersections, int *intersectTypes,
                                        int *, int *, int *bNonDegenIntersectCountPS, int *cNonDegenIntersectCountPS,
                                        int *intersectAlphaValuesB, int *intersectAlphaValuesC,
                                        double *intersectedB, double *intersectedC, int *neighborsB,
                                        int *neighborsC, int *alphaValuesB, int *alphaValuesC) {
    if (!loaded_) throw Error(SEGCLIP_ESTATE, "SegTree: polygons not loaded");
    check(segclip_copy_intersections_arrays(ctx_, bSize, cSize, cStart, bPolyLineIds, cPolyLineIds, polySize,
                                            bPolLineIdsSortedIndex, cPolLineIdsSortedIndex, sortedBPolyIdIndicies,
                                            sortedCPolyIdIndicies, intersections, intersectTypes,
                                            bNonDegenIntersectCountPS, cNonDegenIntersectCountPS,
                                            intersectAlphaValuesB, intersectAlphaValuesC, intersectedB,
                                            intersectedC, neighborsB, neighborsC, alphaValuesB, alphaValuesC),
          "copyIntersectionsParallel");
}

void SegTree::calculateInitLabelParallel(int, int *, double *intersectedB, double *intersectedC, int *alphaValuesB,
                                         int *neighborsB, int bIntersectedCount, int cIntersectedCount,
                                         int *bInitLabels, int *cInitLabels) {
    if (!ctx_) check(segclip_create(&ctx_, 0, nullptr), "create");
    check(segclip_init_labels_arrays(ctx_, bIntersectedCount, cIntersectedCount, intersectedB, intersectedC,
                                     alphaValuesB, neighborsB, bInitLabels, cInitLabels),
          "calculateInitLabelParallel");
}

} // namespace segclip
