// sc_capi.hip -- the extern "C" boundary of libsegclip (declared in include/segclip.h).
#include "../../include/segclip.h"
#include "sc_ctx.hpp"

#include <algorithm>
#include <cstring>

struct segclip_ctx {
    sc::Ctx c;
};

namespace sc {
Ctx::~Ctx() {
    pairs_graph_reset(*this);
    if (ev_ready)
        for (auto &e : ev) (void)hipEventDestroy(e);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
}
namespace {
struct PinnedStage {
    void *p = nullptr;
    char *dev = nullptr; // the same memory as seen by the device
    size_t cap = 0;
};
thread_local PinnedStage tl_pin; // never freed: a read-back has finished with it when it returns

// A read-back's items copied into the pinned stage by one launch (the
// device's stores reach host memory before the stream reports completion).
struct GatherArgs {
    const uint32_t *src[RB_MAX];
    uint32_t *dst[RB_MAX];
    uint32_t words[RB_MAX];
    int n;
};
__global__ void k_readback(GatherArgs a) {
    for (int q = 0; q < a.n; ++q)
        for (uint32_t i = threadIdx.x; i < a.words[q]; i += blockDim.x) a.dst[q][i] = a.src[q][i];
}
constexpr size_t RB_GATHER_MAX = 65536; // larger items go through the copy engine path
} // namespace
void *Ctx::pinned(size_t bytes) {
    PinnedStage &ps = tl_pin;
    if (bytes > ps.cap) {
        if (ps.p) SC_HIP(hipHostFree(ps.p));
        ps.p = nullptr;
        ps.cap = 0;
        const size_t want = bytes < 65536 ? 65536 : bytes + bytes / 4;
        SC_HIP(hipHostMalloc(&ps.p, want, hipHostMallocDefault));
        void *dp = nullptr;
        SC_HIP(hipHostGetDevicePointer(&dp, ps.p, 0));
        ps.dev = static_cast<char *>(dp);
        ps.cap = want;
    }
    return ps.p;
}
void Ctx::wait() {
    for (;;) {
        const hipError_t e = hipStreamQuery(stream);
        if (e == hipSuccess) return;
        if (e != hipErrorNotReady) SC_HIP(e);
    }
}
void ReadBack::run() {
    size_t tot = 0;
    for (int i = 0; i < n; ++i) tot += (bytes[i] + 15) & ~size_t(15);
    char *st = static_cast<char *>(c.pinned(tot));
    char *sd = tl_pin.dev;
    GatherArgs g;
    g.n = 0;
    size_t at = 0;
    for (int i = 0; i < n; ++i) {
        if (bytes[i] && bytes[i] % 4 == 0 && bytes[i] <= RB_GATHER_MAX && (uintptr_t(src[i]) & 3) == 0) {
            g.src[g.n] = static_cast<const uint32_t *>(src[i]);
            g.dst[g.n] = reinterpret_cast<uint32_t *>(sd + at);
            g.words[g.n] = uint32_t(bytes[i] / 4);
            ++g.n;
        } else if (bytes[i]) {
            SC_HIP(hipMemcpyAsync(st + at, src[i], bytes[i], hipMemcpyDeviceToHost, c.stream));
        }
        at += (bytes[i] + 15) & ~size_t(15);
    }
    if (g.n) {
        k_readback<<<1, 256, 0, c.stream>>>(g);
        SC_HIP(hipGetLastError());
    }
    c.wait();
    at = 0;
    for (int i = 0; i < n; ++i) {
        if (bytes[i]) std::memcpy(dst[i], st + at, bytes[i]);
        at += (bytes[i] + 15) & ~size_t(15);
    }
    n = 0;
}
void Ctx::ensure_side_stream() {
    if (stream2) return;
    SC_HIP(hipStreamCreate(&stream2));
    SC_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    SC_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
}
void Ctx::ensure_events() {
    if (ev_ready) return;
    for (auto &e : ev) SC_HIP(hipEventCreate(&e));
    ev_ready = true;
}
float Ctx::elapsed(int a, int b) {
    float ms = 0;
    SC_HIP(hipEventElapsedTime(&ms, ev[a], ev[b]));
    return ms;
}

void release_derived(Ctx &c) {
    SC_HIP(hipStreamSynchronize(c.stream));
    fetch_flags(c); // the batch status stays readable after a release
    for (DBuf<double> *b : {&c.ex0, &c.ey0, &c.ex1, &c.ey1, &c.estart, &c.eend, &c.uxs, &c.uxe, &c.eymin, &c.eymax,
                            &c.ea, &c.niv, &c.p_x, &c.p_y, &c.aug_bx, &c.aug_by, &c.aug_cx, &c.aug_cy})
        b->release();
    for (DBuf<int> *b : {&c.sp_slot, &c.sp_ndx, &c.sp_edge, &c.sp_av, &c.sp_nbr, &c.sp_lab, &c.sp_pos, &c.sp_ndf})
        b->release();
    for (DBuf<YRangeF> *b : {&c.eyr, &c.lend_yr, &c.cov_yr}) b->release();
    c.lend_bl.release();
    c.ep_lohi.release();
    c.exy.release();
    c.near.release();
    c.near_cnt.release();
    c.srow_off.release();
    c.srow_node.release();
    c.sp_x.release();
    c.sp_y.release();
    for (DBuf<int> *b : {&c.etype, &c.eid, &c.after_leaf, &c.before_leaf, &c.p_j, &c.p_b, &c.p_c, &c.p_type,
                         &c.p_qa, &c.p_qb, &c.ordB, &c.ordC, &c.nd_b, &c.nd_c, &c.all_b, &c.all_c, &c.nbr_b,
                         &c.nbr_c, &c.av_b, &c.av_c, &c.lab_b, &c.lab_c, &c.slot_b, &c.slot_c, &c.sv_b, &c.sp_b,
                         &c.sv_c, &c.sp_c, &c.arr_b, &c.arr_c})
        b->release();
    for (DBuf<uint32_t> *b : {&c.lend_off, &c.lend_e, &c.mend_off, &c.mend_e, &c.cov_off, &c.cov_e, &c.cov_split,
                              &c.d_pair_off, &c.u32a, &c.u32b, &c.u32c, &c.u32d, &c.lend_e_alt,
                              &c.ord_val, &c.ord_val_alt, &c.seg_wb, &c.seg_wc, &c.seg_bb,
                              &c.seg_bc, &c.seg_ctl})
        b->release();
    for (DBuf<uint64_t> *b : {&c.lend_key, &c.lend_key_alt, &c.cov_key, &c.cov_key_alt, &c.unit_desc, &c.cand,
                              &c.cand_alt, &c.pairs, &c.k64a, &c.k64b, &c.u64a, &c.u64b, &c.ord_key, &c.ord_key_alt,
                              &c.seg_kb, &c.seg_kc})
        b->release();
    c.rs.global_hist.release();
    c.rs.tile_status.release();
    c.rs.keys_alt.release();
    c.rs.vals_alt.release();
    release_foster_scratch();
    c.epc = c.nleaves = c.T = 0;
    c.n_lend = c.n_cov = c.n_cand = c.n_pairs = c.n_mend = 0;
    c.cur_k = 0;
    c.I = 0;
    c.pair_off.clear();
}
} // namespace sc

namespace {

template <class F> int guarded(segclip_ctx *h, F &&f) {
    if (!h) return SEGCLIP_EINVAL;
    try {
        SC_HIP(hipSetDevice(h->c.device));
        f(h->c);
        h->c.last_error.clear();
        return SEGCLIP_OK;
    } catch (const sc::Error &e) {
        h->c.last_error = e.what();
        return e.code;
    } catch (const std::exception &e) {
        h->c.last_error = e.what();
        return SEGCLIP_EHIP;
    }
}

int need_space(int64_t n, int64_t cap, int64_t *need) {
    if (need) *need = n;
    return n > cap ? SEGCLIP_ESPACE : SEGCLIP_OK;
}

template <class T> void d2h(sc::Ctx &c, T *dst, const T *src, int64_t n) {
    if (n <= 0) return;
    SC_HIP(hipMemcpyAsync(dst, src, sizeof(T) * size_t(n), hipMemcpyDeviceToHost, c.stream));
}

void tree_all(sc::Ctx &c) {
    sc::build_edges(c);
    sc::build_elementary(c);
    sc::build_node_intervals(c);
    sc::build_leaf_end_entries(c);
    sc::build_cover_lists(c);
}

} // namespace

extern "C" {

const char *segclip_version(void) { return "segclip 0.1 (gfx950)"; }

int segclip_create(segclip_ctx **out, int device, void *stream) {
    if (!out) return SEGCLIP_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return SEGCLIP_EHIP;
    }
    if (device < 0 || device >= n) return SEGCLIP_EINVAL;
    segclip_ctx *h = new (std::nothrow) segclip_ctx();
    if (!h) return SEGCLIP_ENOMEM;
    h->c.device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete h;
        return SEGCLIP_EHIP;
    }
    if (stream) {
        h->c.stream = static_cast<hipStream_t>(stream);
    } else {
        if (hipStreamCreateWithFlags(&h->c.stream, hipStreamNonBlocking) != hipSuccess) {
            delete h;
            return SEGCLIP_EHIP;
        }
        h->c.own_stream = true;
    }
    *out = h;
    return SEGCLIP_OK;
}

void segclip_destroy(segclip_ctx *h) {
    if (!h) return;
    (void)hipSetDevice(h->c.device);
    (void)hipStreamSynchronize(h->c.stream);
    delete h;
}

const char *segclip_last_error(const segclip_ctx *h) { return h ? h->c.last_error.c_str() : "null context"; }

int segclip_set_polygons(segclip_ctx *h, const double *b, int nb, const double *c, const int *csz, int K) {
    return guarded(h, [&](sc::Ctx &x) { sc::ctx_set_polygons(x, b, nb, c, csz, K, false); });
}
int segclip_set_polygons_device(segclip_ctx *h, const double *b, int nb, const double *c, const int *csz, int K) {
    return guarded(h, [&](sc::Ctx &x) { sc::ctx_set_polygons(x, b, nb, c, csz, K, true); });
}
int segclip_set_pairs(segclip_ctx *h, const double *b, const int *bsz, const double *c, const int *csz, int P) {
    return guarded(h, [&](sc::Ctx &x) { sc::ctx_set_pairs(x, b, bsz, c, csz, P, false); });
}
int segclip_set_pairs_device(segclip_ctx *h, const double *b, const int *bsz, const double *c, const int *csz,
                             int P) {
    return guarded(h, [&](sc::Ctx &x) { sc::ctx_set_pairs(x, b, bsz, c, csz, P, true); });
}
int segclip_build_elementary(segclip_ctx *h) {
    return guarded(h, [&](sc::Ctx &x) {
        sc::build_edges(x);
        sc::build_elementary(x);
    });
}
int segclip_build_node_intervals(segclip_ctx *h) {
    return guarded(h, [&](sc::Ctx &x) { sc::build_node_intervals(x); });
}
int segclip_build_end_lists(segclip_ctx *h) {
    return guarded(h, [&](sc::Ctx &x) { sc::build_leaf_end_entries(x); });
}
int segclip_build_cover_lists(segclip_ctx *h) {
    return guarded(h, [&](sc::Ctx &x) { sc::build_cover_lists(x); });
}
int segclip_materialize_end_lists(segclip_ctx *h) {
    return guarded(h, [&](sc::Ctx &x) { sc::materialize_end_lists(x); });
}
int segclip_build_tree(segclip_ctx *h) { return guarded(h, [&](sc::Ctx &x) { tree_all(x); }); }
int segclip_find_candidates(segclip_ctx *h) {
    return guarded(h, [&](sc::Ctx &x) { sc::find_candidates(x); });
}
int segclip_intersections(segclip_ctx *h, int k) {
    int rc = guarded(h, [&](sc::Ctx &x) { sc::post_overlay(x, k); });
    return rc;
}
int segclip_label_traverse(segclip_ctx *h) {
    return guarded(h, [&](sc::Ctx &x) { sc::label_and_traverse(x); });
}
int segclip_clip(segclip_ctx *h, int k) {
    return guarded(h, [&](sc::Ctx &x) { sc::clip_all(x, k); });
}
int segclip_set_provenance(segclip_ctx *h, int on) {
    return guarded(h, [&](sc::Ctx &x) { x.prov = on != 0; });
}
int segclip_release(segclip_ctx *h) {
    return guarded(h, [&](sc::Ctx &x) { sc::release_derived(x); });
}
int segclip_set_allocator(segclip_alloc_fn alloc, segclip_free_fn free_fn, void *user) {
    if ((alloc == nullptr) != (free_fn == nullptr)) return SEGCLIP_EINVAL;
    sc::AllocHooks &h = sc::alloc_hooks();
    h.alloc = alloc;
    h.free = free_fn;
    h.user = user;
    return SEGCLIP_OK;
}
int segclip_memory(int64_t *current, int64_t *peak, int reset_peak) {
    if (current) *current = sc::dev_bytes().load();
    if (peak) *peak = sc::dev_peak().load();
    if (reset_peak) sc::dev_peak().store(sc::dev_bytes().load());
    return SEGCLIP_OK;
}
int segclip_batch_area(segclip_ctx *h, double *out4, int64_t cap, int64_t *need) {
    if (!h) return SEGCLIP_EINVAL;
    const int J = h->c.jobs.J;
    if (J < 1) return SEGCLIP_ESTATE;
    int rc = need_space(4 * int64_t(J), cap, need);
    if (rc || !out4) return rc;
    return guarded(h, [&](sc::Ctx &x) {
        sc::DBuf<double> d;
        d.ensure(4 * size_t(J));
        sc::batch_area(x, d.p);
        d2h(x, out4, d.p, 4 * int64_t(J));
        SC_HIP(hipStreamSynchronize(x.stream));
    });
}
int segclip_batch_area_device(segclip_ctx *h, double *out4) {
    if (!h || !out4) return SEGCLIP_EINVAL;
    return guarded(h, [&](sc::Ctx &x) {
        sc::batch_area(x, out4);
        SC_HIP(hipStreamSynchronize(x.stream));
    });
}
int segclip_batch_area_backward_device(segclip_ctx *h, const double *g_area, const double *g_iou, double *grad_b,
                                       double *grad_c) {
    if (!h || !grad_b || !grad_c) return SEGCLIP_EINVAL;
    return guarded(h, [&](sc::Ctx &x) { sc::batch_area_backward(x, g_area, g_iou, grad_b, grad_c); });
}
int segclip_save_intersections_ids(segclip_ctx *h, const int32_t *b_id, const int32_t *c_id, int n, double *xy,
                                   int32_t *type, int32_t *alpha_b, int32_t *alpha_c) {
    return guarded(h, [&](sc::Ctx &x) { sc::arr_save_intersections(x, b_id, c_id, n, xy, type, alpha_b, alpha_c); });
}
int segclip_sort_intersections_at_edges(segclip_ctx *h, const int32_t *alpha, const int32_t *sorted_index,
                                        int c_start, const int32_t *all_counts, const int32_t *nd_counts,
                                        int ncounts, const int32_t *unique_ids, int unique_count, int n,
                                        int32_t *sorted_alpha_index) {
    return guarded(h, [&](sc::Ctx &x) {
        sc::arr_sort_at_edges(x, alpha, sorted_index, c_start, all_counts, nd_counts, ncounts, unique_ids,
                              unique_count, n, sorted_alpha_index);
    });
}
int segclip_copy_intersections_arrays(segclip_ctx *h, int b_size, int c_size, int c_start, const int32_t *b_ids,
                                      const int32_t *c_ids, int n, const int32_t *b_sorted_index,
                                      const int32_t *c_sorted_index, const int32_t *sorted_b,
                                      const int32_t *sorted_c, const double *intersections, const int32_t *types,
                                      const int32_t *b_nd_ps, const int32_t *c_nd_ps, const int32_t *alpha_b,
                                      const int32_t *alpha_c, double *intersected_b, double *intersected_c,
                                      int32_t *neighbors_b, int32_t *neighbors_c, int32_t *alpha_values_b,
                                      int32_t *alpha_values_c) {
    return guarded(h, [&](sc::Ctx &x) {
        sc::arr_copy_intersections(x, b_size, c_size, c_start, b_ids, c_ids, n, b_sorted_index, c_sorted_index,
                                   sorted_b, sorted_c, intersections, types, b_nd_ps, c_nd_ps, alpha_b, alpha_c,
                                   intersected_b, intersected_c, neighbors_b, neighbors_c, alpha_values_b,
                                   alpha_values_c);
    });
}
int segclip_init_labels_arrays(segclip_ctx *h, int nb_aug, int nc_aug, const double *intersected_b,
                               const double *intersected_c, const int32_t *alpha_values_b,
                               const int32_t *neighbors_b, int32_t *labels_b, int32_t *labels_c) {
    return guarded(h, [&](sc::Ctx &x) {
        sc::arr_init_labels(x, nb_aug, nc_aug, intersected_b, intersected_c, alpha_values_b, neighbors_b, labels_b,
                            labels_c);
    });
}
int segclip_clip_batch(segclip_ctx *h, int k0, int k1) {
    return guarded(h, [&](sc::Ctx &x) { sc::clip_jobs(x, k0, k1); });
}
int segclip_set_operation(segclip_ctx *h, int op) {
    return guarded(h, [&](sc::Ctx &x) {
        SC_REQUIRE(op == SEGCLIP_OP_INTERSECTION || op == SEGCLIP_OP_UNION, SEGCLIP_EINVAL, "unknown operation");
        if (x.uni != (op == SEGCLIP_OP_UNION ? 1 : 0)) {
            x.uni = op == SEGCLIP_OP_UNION ? 1 : 0;
            x.jobs.J = 0; // results of the other operation are gone
            x.jobs.ring_off.clear();
        }
    });
}
int segclip_pairs_area_graph(segclip_ctx *h, const double *base_xy, const double *ov_xy, double *out4) {
    return guarded(h, [&](sc::Ctx &x) { sc::pairs_area_graph(x, base_xy, ov_xy, out4); });
}
int segclip_reset(segclip_ctx *h) {
    return guarded(h, [&](sc::Ctx &x) {
        x.epc = x.nleaves = x.T = 0;
        x.n_lend = x.n_cov = x.n_cand = x.n_pairs = x.n_mend = x.n_near = 0;
        x.cur_k = 0;
        x.pair_off.clear();
        x.jobs.J = 0;
        x.jobs.AB.clear();
        x.jobs.ring_off.clear();
        x.flags_pending = false;
    });
}
int segclip_copy_batch_rings(segclip_ctx *h, int32_t *job_ring_off, int64_t cap, int64_t *need) {
    if (!h) return SEGCLIP_EINVAL;
    const sc::JobSet &js = h->c.jobs;
    if (int(js.ring_off.size()) != js.J + 1 || js.J < 1) return SEGCLIP_ESTATE;
    int rc = need_space(js.J + 1, cap, need);
    if (rc || !job_ring_off) return rc;
    std::memcpy(job_ring_off, js.ring_off.data(), sizeof(int32_t) * (js.J + 1));
    return SEGCLIP_OK;
}
int segclip_copy_batch_routes(segclip_ctx *h, int32_t *route, int32_t *reason, int64_t cap, int64_t *need) {
    if (!h) return SEGCLIP_EINVAL;
    const sc::JobSet &js = h->c.jobs;
    if (int(js.route.size()) != js.J || js.J < 1) return SEGCLIP_ESTATE;
    int rc = need_space(js.J, cap, need);
    if (rc || !route || !reason) return rc;
    std::memcpy(route, js.route.data(), sizeof(int32_t) * js.J);
    std::memcpy(reason, js.reason.data(), sizeof(int32_t) * js.J);
    return SEGCLIP_OK;
}
int segclip_copy_batch_status(segclip_ctx *h, int32_t *status, int64_t cap, int64_t *need) {
    if (!h) return SEGCLIP_EINVAL;
    int frc = guarded(h, [&](sc::Ctx &x) { sc::fetch_flags(x); });
    if (frc != SEGCLIP_OK) return frc;
    const sc::JobSet &js = h->c.jobs;
    if (int(js.flags.size()) != js.J || js.J < 1) return SEGCLIP_ESTATE;
    int rc = need_space(js.J, cap, need);
    if (rc || !status) return rc;
    std::memcpy(status, js.flags.data(), sizeof(int32_t) * js.J);
    return SEGCLIP_OK;
}

int segclip_get_counts(segclip_ctx *h, segclip_counts *o) {
    if (!h || !o) return SEGCLIP_EINVAL;
    int frc = guarded(h, [&](sc::Ctx &x) { sc::fetch_flags(x); });
    if (frc != SEGCLIP_OK) return frc;
    const sc::Ctx &c = h->c;
    std::memset(o, 0, sizeof(*o));
    o->num_edges = c.ne;
    o->epc = c.epc;
    o->tree_size = c.T;
    o->cover_entries = c.n_cov;
    o->leaf_end_entries = c.n_lend;
    o->end_entries = c.n_mend;
    o->tests = c.n_tests;
    o->candidates = c.n_cand;
    o->pairs = c.n_pairs;
    o->overlay = c.cur_k;
    o->overlay_pairs = c.I;
    o->nb_aug = c.nbaug;
    o->nc_aug = c.ncaug;
    o->rings = c.nrings;
    o->undefined = c.undefined;
    o->out_vertices = c.n_out;
    o->stab_units = c.n_units;
    o->near_pairs = c.n_near;
    return SEGCLIP_OK;
}

int segclip_copy_elementary(segclip_ctx *h, double *out, int64_t cap, int64_t *need) {
    if (!h) return SEGCLIP_EINVAL;
    int rc = need_space(h->c.epc, cap, need);
    if (rc || !out) return rc;
    return guarded(h, [&](sc::Ctx &c) {
        d2h(c, out, c.ea.p, c.epc);
        SC_HIP(hipStreamSynchronize(c.stream));
    });
}

int segclip_copy_node_intervals(segclip_ctx *h, double *out, int64_t cap, int64_t *need) {
    if (!h) return SEGCLIP_EINVAL;
    int rc = need_space(2 * int64_t(h->c.T), cap, need);
    if (rc || !out) return rc;
    return guarded(h, [&](sc::Ctx &c) {
        d2h(c, out, c.niv.p, 2 * int64_t(c.T));
        SC_HIP(hipStreamSynchronize(c.stream));
    });
}

int segclip_copy_cover_lists(segclip_ctx *h, uint32_t *off, uint32_t *edges, int64_t cap, int64_t *need) {
    if (!h) return SEGCLIP_EINVAL;
    int rc = need_space(h->c.n_cov, cap, need);
    if (rc || !off || !edges) return rc;
    return guarded(h, [&](sc::Ctx &c) {
        d2h(c, off, c.cov_off.p, int64_t(c.T) + 1);
        d2h(c, edges, c.cov_e.p, c.n_cov);
        SC_HIP(hipStreamSynchronize(c.stream));
    });
}

int segclip_copy_end_lists(segclip_ctx *h, uint32_t *off, uint32_t *edges, int64_t cap, int64_t *need) {
    if (!h) return SEGCLIP_EINVAL;
    int rc = need_space(h->c.n_mend, cap, need);
    if (rc || !off || !edges) return rc;
    return guarded(h, [&](sc::Ctx &c) {
        d2h(c, off, c.mend_off.p, int64_t(c.T) + 1);
        d2h(c, edges, c.mend_e.p, c.n_mend);
        SC_HIP(hipStreamSynchronize(c.stream));
    });
}

int segclip_copy_leaf_targets(segclip_ctx *h, int32_t *after, int32_t *before, int64_t cap, int64_t *need) {
    if (!h) return SEGCLIP_EINVAL;
    int rc = need_space(h->c.ne, cap, need);
    if (rc || !after || !before) return rc;
    return guarded(h, [&](sc::Ctx &c) {
        d2h(c, after, c.after_leaf.p, c.ne);
        d2h(c, before, c.before_leaf.p, c.ne);
        SC_HIP(hipStreamSynchronize(c.stream));
    });
}

int segclip_copy_pairs(segclip_ctx *h, int k, int32_t *b_id, int32_t *c_id, int32_t *type, int64_t cap,
                       int64_t *need) {
    if (!h) return SEGCLIP_EINVAL;
    sc::Ctx &c = h->c;
    if (int(c.pair_off.size()) != c.K + 1 || k < 1 || k > c.K) return SEGCLIP_ESTATE;
    int64_t n = c.pair_off[k] - c.pair_off[k - 1];
    int rc = need_space(n, cap, need);
    if (rc || !b_id || !c_id) return rc;
    return guarded(h, [&](sc::Ctx &x) {
        std::vector<uint64_t> keys(size_t(n > 0 ? n : 1));
        d2h(x, keys.data(), x.pairs.p + x.pair_off[k - 1], n);
        SC_HIP(hipStreamSynchronize(x.stream));
        // one-to-many: the reference's ids (base ring first, overlay k at its
        // vertex offset); paired: ids within the job's own two rings, as if
        // the pair were clipped alone
        const int br = sc::job_base(x, k);
        const int cStart = x.paired ? 0 : x.coff[k - 1];
        const int bFirst = x.beoff[br];
        for (int64_t i = 0; i < n; ++i) {
            int b = int((keys[i] >> x.bits_c) & ((1ull << x.bits_b) - 1ull)) - bFirst;
            int cl = int(keys[i] & ((1ull << x.bits_c) - 1ull));
            b_id[i] = 2 * b;
            c_id[i] = 2 * (cStart + cl);
            if (type) type[i] = -1; // filled by segclip_intersections order; see copy_intersections
        }
        if (type) {
            // type recomputed on the device during segclip_intersections; copy if it matches this overlay
            if (x.cur_k == k && x.I == n) d2h(x, type, x.p_type.p, n);
            SC_HIP(hipStreamSynchronize(x.stream));
        }
    });
}

int segclip_copy_intersections(segclip_ctx *h, double *xy, int32_t *alpha_b, int32_t *alpha_c, int64_t cap,
                               int64_t *need) {
    if (!h) return SEGCLIP_EINVAL;
    int rc = need_space(h->c.I, cap, need);
    if (rc || !xy) return rc;
    return guarded(h, [&](sc::Ctx &c) {
        int n = c.I;
        std::vector<double> x(size_t(n > 0 ? n : 1)), y(size_t(n > 0 ? n : 1));
        d2h(c, x.data(), c.p_x.p, n);
        d2h(c, y.data(), c.p_y.p, n);
        if (alpha_b) d2h(c, alpha_b, c.p_qa.p, n);
        if (alpha_c) d2h(c, alpha_c, c.p_qb.p, n);
        SC_HIP(hipStreamSynchronize(c.stream));
        for (int i = 0; i < n; ++i) {
            xy[2 * i] = x[i];
            xy[2 * i + 1] = y[i];
        }
    });
}

int segclip_copy_augmented(segclip_ctx *h, int side, double *xy, int32_t *nbr, int32_t *alpha, int32_t *label,
                           int64_t cap, int64_t *need) {
    if (!h) return SEGCLIP_EINVAL;
    if (side == 0 && h->c.sparseP) {
        h->c.last_error = "the base rings of a one-to-many batch are kept sparse; copy them per overlay";
        return SEGCLIP_ESTATE;
    }
    int64_t n = side == 0 ? h->c.nbaug : h->c.ncaug;
    int rc = need_space(n, cap, need);
    if (rc || !xy) return rc;
    return guarded(h, [&](sc::Ctx &c) {
        std::vector<double> x(size_t(n > 0 ? n : 1)), y(size_t(n > 0 ? n : 1));
        d2h(c, x.data(), side == 0 ? c.aug_bx.p : c.aug_cx.p, n);
        d2h(c, y.data(), side == 0 ? c.aug_by.p : c.aug_cy.p, n);
        if (nbr) d2h(c, nbr, side == 0 ? c.nbr_b.p : c.nbr_c.p, n);
        if (alpha) d2h(c, alpha, side == 0 ? c.av_b.p : c.av_c.p, n);
        if (label) d2h(c, label, side == 0 ? c.lab_b.p : c.lab_c.p, n);
        SC_HIP(hipStreamSynchronize(c.stream));
        for (int64_t i = 0; i < n; ++i) {
            xy[2 * i] = x[i];
            xy[2 * i + 1] = y[i];
        }
    });
}

int segclip_copy_rings(segclip_ctx *h, int32_t *ring_off, int64_t off_cap, double *xy, int64_t xy_cap,
                       int64_t *need) {
    if (!h) return SEGCLIP_EINVAL;
    const sc::Ctx &c0 = h->c;
    if (need) {
        need[0] = int64_t(c0.nrings) + 1;
        need[1] = c0.n_out;
    }
    if (int64_t(c0.nrings) + 1 > off_cap || c0.n_out > xy_cap) return SEGCLIP_ESPACE;
    if (!ring_off || !xy) return SEGCLIP_EINVAL;
    return guarded(h, [&](sc::Ctx &c) {
        int64_t n = c.n_out;
        std::vector<double> x(size_t(n > 0 ? n : 1)), y(size_t(n > 0 ? n : 1));
        d2h(c, ring_off, c.d_ring_off.p, int64_t(c.nrings) + 1);
        d2h(c, x.data(), c.d_ring_x.p, n);
        d2h(c, y.data(), c.d_ring_y.p, n);
        SC_HIP(hipStreamSynchronize(c.stream));
        for (int64_t i = 0; i < n; ++i) {
            xy[2 * i] = x[i];
            xy[2 * i + 1] = y[i];
        }
    });
}

int segclip_set_timing(segclip_ctx *h, int on) {
    return guarded(h, [&](sc::Ctx &c) {
        if (on) c.ensure_events();
        c.timing = on != 0;
    });
}

int segclip_set_tuning(segclip_ctx *h, const char *name, int value) {
    return guarded(h, [&](sc::Ctx &c) {
        SC_REQUIRE(name != nullptr, SEGCLIP_EINVAL, "tuning name is null");
        if (std::strcmp(name, "stab_tiles") == 0) {
            SC_REQUIRE(value >= 1 && value <= 16, SEGCLIP_EINVAL, "stab_tiles must be in 1..16");
            c.stab_tiles = value;
        } else if (std::strcmp(name, "stab_units_per_wave") == 0) {
            SC_REQUIRE(value >= 1 && value <= 32, SEGCLIP_EINVAL, "stab_units_per_wave must be in 1..32");
            c.stab_upw = value;
        } else if (std::strcmp(name, "small_fused") == 0) {
            SC_REQUIRE(value == 0 || value == 1, SEGCLIP_EINVAL, "small_fused must be 0 or 1");
            c.small_fused = value;
        } else if (std::strcmp(name, "seq_parallel_walks") == 0) {
            SC_REQUIRE(value == 0 || value == 1, SEGCLIP_EINVAL, "seq_parallel_walks must be 0 or 1");
            c.seq_par = value;
        } else {
            SC_REQUIRE(false, SEGCLIP_EINVAL, std::string("unknown tuning setting: ") + name);
        }
    });
}

int segclip_get_tuning(segclip_ctx *h, const char *name, int *value) {
    return guarded(h, [&](sc::Ctx &c) {
        SC_REQUIRE(name != nullptr && value != nullptr, SEGCLIP_EINVAL, "null argument");
        if (std::strcmp(name, "stab_tiles") == 0) {
            *value = c.stab_tiles;
        } else if (std::strcmp(name, "stab_units_per_wave") == 0) {
            *value = c.stab_upw;
        } else if (std::strcmp(name, "small_fused") == 0) {
            *value = c.small_fused;
        } else if (std::strcmp(name, "seq_parallel_walks") == 0) {
            *value = c.seq_par;
        } else if (std::strcmp(name, "small_fallback") == 0) {
            *value = c.small_fail;
        } else {
            SC_REQUIRE(false, SEGCLIP_EINVAL, std::string("unknown tuning setting: ") + name);
        }
    });
}

int segclip_get_stage_times(segclip_ctx *h, float *ms8) {
    if (!h || !ms8) return SEGCLIP_EINVAL;
    const sc::StageTimes &t = h->c.times;
    float v[8] = {t.elementary, t.tree, t.endlist, t.cover, t.candidates, t.stab_kernel, t.post, t.label};
    std::memcpy(ms8, v, sizeof(v));
    return SEGCLIP_OK;
}

} // extern "C"
