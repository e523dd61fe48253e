// segtree_check.cpp -- test driver of the segclip::SegTree facade.
//
// Replays the per-overlay call sequence of the reference's multicore driver
// (src/clip.cpp multicore_clipping, clip:1501-1990) through the facade: the
// tree-stage calls on Edge records, then per overlay
// saveIntersectionsIdsMulticore, the driver's own stable sorts and edge
// counts (countEdgeIntersections, clip:522-601, restated here as driver
// code), saveIntersectionsParallel, sortIntersectionsAtEdgesParallel on both
// rings, copyIntersectionsParallel and calculateInitLabelParallel.  Writes the
// augmented rings and labels of every overlay for comparison with the oracle.
//
//   segtree_check <in.bin> <out.bin>
//   in:  int32 nb, int32 K, int32 csz[K], f64 base[2nb], f64 overlays[2*sum(csz)]
//   out: per overlay int32 nba, nca, f64 xyB[2nba], f64 xyC[2nca],
//        int32 nbrB[nba], nbrC[nca], alphaB[nba], alphaC[nca], labB[nba], labC[nca]
#include "../../include/segtree.h"

#include <algorithm>
#include <cstdio>
#include <numeric>
#include <vector>

struct Edge {  // the fields of the reference's Edge (edge.h:141-202) the facade reads
    int id, type;
    double start, end, y;
};

static bool nondeg(char side, int t) {
    return side == 'b' ? (t == 1 || t == 3 || t == 5 || t == 7) : (t == 1 || t == 2 || t == 6);
}

// per-edge counts over the sorted pair order, the list of intersected edge ids
static void count_edges(const std::vector<int> &ids, const std::vector<int> &sorted, int cStart,
                        const std::vector<int> &types, char side, int size, std::vector<int> &all,
                        std::vector<int> &nd, std::vector<int> &uniq) {
    all.assign(size + 1, 0);
    nd.assign(size + 1, 0);
    uniq.clear();
    for (size_t i = 0; i < sorted.size(); ++i) {
        int p = sorted[i];
        int e = ids[p] / 2 - cStart;
        if (uniq.empty() || uniq.back() != ids[p]) uniq.push_back(ids[p]);
        all[e]++;
        if (nondeg(side, types[p])) nd[e]++;
    }
    int a = 0, b = 0;
    for (int i = 0; i <= size; ++i) {
        int x = all[i], y = nd[i];
        all[i] = a;
        nd[i] = b;
        a += x;
        b += y;
    }
}

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: segtree_check in.bin out.bin\n");
        return 2;
    }
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int nb = 0, K = 0;
    if (std::fread(&nb, 4, 1, f) != 1 || std::fread(&K, 4, 1, f) != 1 || nb < 3 || K < 1) return 2;
    std::vector<int> csz(K);
    if (std::fread(csz.data(), 4, K, f) != size_t(K)) return 2;
    int nct = std::accumulate(csz.begin(), csz.end(), 0);
    std::vector<double> bxy(2 * size_t(nb)), cxy(2 * size_t(nct));
    if (std::fread(bxy.data(), 8, bxy.size(), f) != bxy.size() || std::fread(cxy.data(), 8, cxy.size(), f) != cxy.size())
        return 2;
    std::fclose(f);
    // reference layout: closing vertex repeated per ring (clip:186-190, 285-294)
    std::vector<double> bPoly(bxy), cPoly;
    bPoly.push_back(bxy[0]);
    bPoly.push_back(bxy[1]);
    std::vector<int> cSizes(K + 1, 0);  // vertex offsets with closing vertices
    for (int k = 0, at = 0; k < K; ++k) {
        cPoly.insert(cPoly.end(), cxy.begin() + 2 * at, cxy.begin() + 2 * (at + csz[k]));
        cPoly.push_back(cxy[2 * at]);
        cPoly.push_back(cxy[2 * at + 1]);
        at += csz[k];
        cSizes[k + 1] = cSizes[k] + csz[k] + 1;
    }
    std::vector<Edge> edges;
    double mn = 1e300, mx = -1e300;
    auto add = [&](const double *P, int n, int type, int off) {
        for (int i = 0; i < n; ++i) {
            double a = P[2 * i], c = P[2 * i + 2];
            edges.push_back({2 * (off + i), type, std::min(a, c), std::max(a, c), a});
            mn = std::min(mn, a);
            mx = std::max(mx, a);
        }
    };
    add(bPoly.data(), nb, 0, 0);
    for (int k = 0; k < K; ++k) add(cPoly.data() + 2 * size_t(cSizes[k]), csz[k], k + 1, cSizes[k]);
    const int ic = int(edges.size());
    FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    try {
        segclip::SegTree sTree(ic, ic, mn, mx);
        sTree.constructElementaryIntervalsArrayMulticore(edges.data());
        sTree.init();
        sTree.POLY_TYPE_COUNT = K + 1;
        sTree.constructTreeIntervalArrayMulticore();
        int maxId = std::max(nb, nct + K);
        sTree.insertAllToEdgeListMulticore(edges.data(), maxId);
        sTree.insertAllParallel(edges.data());
        for (int cType = 1; cType <= K; ++cType) {
            const int cSize = csz[cType - 1], cStart = cSizes[cType - 1];
            std::vector<int> bIds, cIds, types;
            sTree.saveIntersectionsIdsMulticore(bPoly.data(), cPoly.data(), nb, cSize, 0, cType, bIds, cIds, types);
            const int n = int(bIds.size());
            {   // the hybrid driver's sequence (clip:1024-1090) gives the same pairs
                std::vector<int> iCount(sTree.treeSize);
                sTree.countIntersectionsParallel(bPoly.data(), cPoly.data(), nb, cSize, K, 0, cType, iCount.data());
                int run = 0;
                for (int &v : iCount) {  // the driver's exclusive scan
                    const int x = v;
                    v = run;
                    run += x;
                }
                const int dup = iCount.back();
                std::vector<int> hb(std::max(dup, 1)), hc(std::max(dup, 1)), ht(std::max(dup, 1));
                sTree.polyClipArrayGPUParallel(bPoly.data(), cPoly.data(), nb, cSize, 0, cType, iCount.data(),
                                               hb.data(), hc.data(), ht.data());
                bool same = dup == n;
                for (int i = 0; same && i < n; ++i)
                    same = hb[i] == bIds[i] && hc[i] == cIds[i] && ht[i] == types[i];
                if (!same) {
                    std::fprintf(stderr, "segtree_check: hybrid pairs differ from the multicore pairs (overlay %d)\n",
                                 cType);
                    std::fclose(o);
                    return 1;
                }
            }
            std::vector<int> bSorted(n), cSorted(n);
            std::iota(bSorted.begin(), bSorted.end(), 0);
            std::iota(cSorted.begin(), cSorted.end(), 0);
            std::stable_sort(bSorted.begin(), bSorted.end(), [&](int a, int b) { return bIds[a] < bIds[b]; });
            std::stable_sort(cSorted.begin(), cSorted.end(), [&](int a, int b) { return cIds[a] < cIds[b]; });
            std::vector<int> bAll, bNd, bUniq, cAll, cNd, cUniq;
            count_edges(bIds, bSorted, 0, types, 'b', nb, bAll, bNd, bUniq);
            count_edges(cIds, cSorted, cStart, types, 'c', cSize, cAll, cNd, cUniq);
            const int nba = nb + bNd[nb], nca = cSize + cNd[cSize];
            std::vector<double> inter(2 * size_t(std::max(n, 1)));
            std::vector<int> qa(std::max(n, 1)), qb(std::max(n, 1)), sB(std::max(n, 1)), sC(std::max(n, 1));
            std::vector<double> xyB(2 * size_t(nba)), xyC(2 * size_t(nca));
            std::vector<int> nbrB(nba), nbrC(nca), aB(nba), aC(nca), lB(nba), lC(nca);
            if (n > 0) {
                sTree.saveIntersectionsParallel(bPoly.data(), cPoly.data(), bIds.data(), cIds.data(), n, inter.data(),
                                                types.data(), qa.data(), qb.data());
                sTree.sortIntersectionsAtEdgesParallel(qa.data(), bSorted.data(), 0, bAll.data(), bNd.data(),
                                                       bUniq.data(), int(bUniq.size()), n, sB.data());
                sTree.sortIntersectionsAtEdgesParallel(qb.data(), cSorted.data(), cStart, cAll.data(), cNd.data(),
                                                       cUniq.data(), int(cUniq.size()), n, sC.data());
            }
            sTree.copyIntersectionsParallel(bPoly.data(), cPoly.data(), nb, cSize, cStart, bIds.data(), cIds.data(),
                                            n, bSorted.data(), cSorted.data(), sB.data(), sC.data(), inter.data(),
                                            types.data(), bAll.data(), cAll.data(), bNd.data(), cNd.data(), qa.data(),
                                            qb.data(), xyB.data(), xyC.data(), nbrB.data(), nbrC.data(), aB.data(),
                                            aC.data());
            sTree.calculateInitLabelParallel(nb, bNd.data(), xyB.data(), xyC.data(), aB.data(), nbrB.data(), nba, nca,
                                             lB.data(), lC.data());
            std::fwrite(&nba, 4, 1, o);
            std::fwrite(&nca, 4, 1, o);
            std::fwrite(xyB.data(), 8, xyB.size(), o);
            std::fwrite(xyC.data(), 8, xyC.size(), o);
            for (auto *v : {&nbrB, &nbrC, &aB, &aC, &lB, &lC}) std::fwrite(v->data(), 4, v->size(), o);
        }
        std::printf("treeSize %d\n", sTree.treeSize);
    } catch (const segclip::Error &e) {
        std::fprintf(stderr, "segtree_check: %s (code %d)\n", e.what(), e.code);
        std::fclose(o);
        return 1;
    }
    std::fclose(o);
    return 0;
}
