// seg_cli.cpp -- bin/seg, the reference's command-line contract on libsegclip.
//
//   seg [-union] <base polygon> <overlay polygon> [<overlay polygon> ...]
//
// -union computes the union instead of the intersection (the UNION switch of
// foster/polyclip.h:39; the reference's sequential clipper takes the same flag,
// lib/optimizedFostersAlgorithm/polyclip.cpp).
// Input files: first line = vertex count n, then n lines "x,y" without the
// repeated closing vertex (src/clip.cpp:172-311).  Every overlay is clipped
// against the base (one tree over all rings, as the reference's one-to-many
// mode, clip:1026-1297, 1702-2062).  Like the reference built with -DTIME it
// prints " skeleton endlist coverlist candidate labeling total" in
// microseconds (clip:2103-2111, file I/O excluded); with SEGCLIP_SAVE=1 the
// intersection of overlay k is written to ../results/Fig-def-R<k>.txt as
// "x y," lines with ';' after each component (foster/polygon.h:572-581),
// SEGCLIP_PRECISION digits (default 6, the reference's stream default).
#include "../../include/segclip.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

static bool read_ring(const char *path, std::vector<double> &xy) {
    FILE *f = std::fopen(path, "r");
    if (!f) return false;
    int n = 0;
    if (std::fscanf(f, "%d", &n) != 1 || n < 3) {
        std::fclose(f);
        return false;
    }
    xy.resize(2 * size_t(n));
    for (int i = 0; i < n; ++i)
        if (std::fscanf(f, "%lf,%lf", &xy[2 * i], &xy[2 * i + 1]) != 2) {
            std::fclose(f);
            return false;
        }
    std::fclose(f);
    return true;
}

static int fail(segclip_ctx *c, const char *what) {
    std::fprintf(stderr, "seg: %s failed: %s\n", what, c ? segclip_last_error(c) : "");
    return 2;
}

int main(int argc, char **argv) {
    std::printf("MI355X algorithm\n");
    bool uni = false;
    if (argc > 1 && std::string(argv[1]) == "-union") {
        uni = true;
        --argc;
        ++argv;
    }
    if (argc < 3) {
        std::printf("3 parameters required\n./seg <base polygon path> <subject polygon path>\n");
        return 0;
    }
    std::vector<double> base, ov;
    std::vector<int> sizes;
    if (!read_ring(argv[1], base)) {
        std::fprintf(stderr, "seg: cannot read polygon %s\n", argv[1]);
        return 1;
    }
    for (int i = 2; i < argc; ++i) {
        std::vector<double> r;
        if (!read_ring(argv[i], r)) {
            std::fprintf(stderr, "seg: cannot read polygon %s\n", argv[i]);
            return 1;
        }
        sizes.push_back(int(r.size() / 2));
        ov.insert(ov.end(), r.begin(), r.end());
    }
    const int K = int(sizes.size());
    segclip_ctx *c = nullptr;
    if (segclip_create(&c, 0, nullptr) != SEGCLIP_OK) return fail(nullptr, "segclip_create");
    segclip_set_timing(c, 1);
    if (uni && segclip_set_operation(c, SEGCLIP_OP_UNION) != SEGCLIP_OK) return fail(c, "segclip_set_operation");
    if (segclip_set_polygons(c, base.data(), int(base.size() / 2), ov.data(), sizes.data(), K) != SEGCLIP_OK)
        return fail(c, "segclip_set_polygons");
    const bool save = std::getenv("SEGCLIP_SAVE") && std::string(std::getenv("SEGCLIP_SAVE")) != "0";
    const int prec = std::getenv("SEGCLIP_PRECISION") ? std::atoi(std::getenv("SEGCLIP_PRECISION")) : 6;
    double t_skel = 0, t_end = 0, t_cov = 0, t_cand = 0, t_lab = 0;
    auto t0 = std::chrono::high_resolution_clock::now();
    for (int k = 1; k <= K; ++k) {
        if (segclip_clip(c, k) != SEGCLIP_OK) return fail(c, "segclip_clip");
        float ms[8];
        segclip_get_stage_times(c, ms);
        t_skel += 1e3 * (ms[0] + ms[1]);
        t_end += 1e3 * ms[2];
        t_cov += 1e3 * ms[3];
        t_cand += 1e3 * (ms[4] + ms[6]);
        t_lab += 1e3 * ms[7];
        if (save) {
            int64_t need[2] = {0, 0};
            segclip_copy_rings(c, nullptr, 0, nullptr, 0, need);
            std::vector<int32_t> off(static_cast<size_t>(need[0]));
            std::vector<double> xy(2 * size_t(need[1] > 0 ? need[1] : 1));
            if (segclip_copy_rings(c, off.data(), need[0], xy.data(), need[1], need) != SEGCLIP_OK)
                return fail(c, "segclip_copy_rings");
            std::string path = "../results/Fig-def-R" + std::to_string(k) + ".txt";
            FILE *f = std::fopen(path.c_str(), "w");
            if (!f) {
                std::fprintf(stderr, "seg: cannot write %s\n", path.c_str());
                return 1;
            }
            for (size_t r = 0; r + 1 < off.size(); ++r) {
                for (int32_t i = off[r]; i < off[r + 1]; ++i) {
                    std::fprintf(f, "%.*g %.*g", prec, xy[2 * size_t(i)], prec, xy[2 * size_t(i) + 1]);
                    std::fprintf(f, i + 1 < off[r + 1] ? ",\n" : ";\n");
                }
            }
            std::fclose(f);
            std::printf("R ");
        }
    }
    auto t1 = std::chrono::high_resolution_clock::now();
    double total = std::chrono::duration<double, std::micro>(t1 - t0).count();
    std::printf(" %.0f %.0f %.0f %.0f %.0f %.0f\n", t_skel, t_end, t_cov, t_cand, t_lab, total);
    segclip_destroy(c);
    return 0;
}
