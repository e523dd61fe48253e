// sc_foster.hip -- Foster/Greiner-Hormann labelling and traversal (rows A16-A17).
#include "sc_ctx.hpp"
#include "sc_fp64.hpp"

namespace sc {

void label_and_traverse(Ctx &c) {
    SC_REQUIRE(c.cur_k >= 1, SC_ESTATE, "segclip_intersections must run first");
    throw Error(SC_ESTATE, "labelling and traversal are not available in this build");
}

void clip_all(Ctx &c, int k) {
    SC_REQUIRE(k >= 1 && k <= c.K, SC_EINVAL, "overlay index out of range");
    if (c.T == 0) {
        build_edges(c);
        build_elementary(c);
        build_node_intervals(c);
        build_leaf_end_entries(c);
        build_cover_lists(c);
    }
    if (int(c.pair_off.size()) != c.K + 1) find_candidates(c);
    post_overlay(c, k);
}

} // namespace sc
