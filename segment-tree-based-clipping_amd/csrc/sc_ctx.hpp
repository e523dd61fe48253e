// sc_ctx.hpp -- device-resident state of one clipping problem (one base ring,
// K overlay rings) and the pipeline stage entry points.
#pragma once
#include "sc_common.hpp"
#include "sc_sparse.hpp"

namespace sc {

// Per-stage timings filled when Ctx::timing is on (milliseconds, HIP events).
struct StageTimes {
    float elementary = 0, tree = 0, endlist = 0, cover = 0, candidates = 0, stab_kernel = 0, post = 0, label = 0;
};

// Jobs of one post-processing batch: job j clips base ring job_base(k0+j)
// against overlay k0+j.  Host vectors and their device copies; per-edge count
// arrays hold n+1 slots per job (JB/JC offsets), augmented rings are
// concatenated job after job (AB/AC offsets, J+1 entries each).
struct JobSet {
    int k0 = 1, J = 0;
    std::vector<int> p0, bvoff, beoff, nb, nc, cvoff, JB, JC, AB, AC, flags;
    std::vector<int> ring_off; // J+1: rings of job j are [ring_off[j], ring_off[j+1]) of the output
    std::vector<int> cin;      // first input vertex (no closing vertices) of each job's overlay
    std::vector<int> route, reason; // labelling route per job (0 crossing, 1 no crossing, 2 sequential), why sequential
    DBuf<int> d_AB, d_AC, d_ring;
    // device copies of p0, bvoff, beoff, nb, nc, cvoff, JB, JC, cin: views into
    // one buffer filled by one host-to-device copy
    DBuf<int> d_pack;
    std::vector<int> h_pack;
    int *d_p0 = nullptr, *d_bvoff = nullptr, *d_beoff = nullptr, *d_nb = nullptr, *d_nc = nullptr,
        *d_cvoff = nullptr, *d_JB = nullptr, *d_JC = nullptr, *d_cin = nullptr;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string last_error;

    // ---- input (reference layout of clip:172-311) ----
    // Base rings: one shared base (one-to-many, the reference's mode) or one
    // base per overlay ("paired": job k clips base k against overlay k).  nb
    // counts the vertices (= edges) of all base rings together.
    int nb = 0, K = 0, nct = 0, ne = 0, maxcsz = 0, maxbsz = 0, nbr = 1;
    bool paired = false;
    std::vector<int> bsz, bvoff, beoff; // base sizes; vertex offsets with closing vertices; edge offsets
    std::vector<int> csz, coff;     // overlay sizes; vertex offsets with closing vertices
    std::vector<int> eoff;          // overlay edge offsets in the edge index space (K+1), eoff[0] = nb
    DBuf<double> bpoly, cpoly;      // (nb+nbr)*2, (nct+K)*2
    DBuf<int> d_coff, d_eoff, d_bvoff, d_beoff;
    DBuf<double> d_shift;           // per base/overlay ring x shift of the tree keys (paired mode)
    // per-edge SoA (edge index e in [0,ne): base edges, then overlay 1, 2, ...)
    DBuf<double> ex0, ey0, ex1, ey1; // endpoints in ring order
    DBuf<double> estart, eend;       // x-extent (tree keys; shifted per pair in paired mode)
    DBuf<double> uxs, uxe;           // paired mode: input x-extent for the bounding-box test
    DBuf<double> eymin, eymax;       // y-extent
    DBuf<int> etype, eid;
    DBuf<YRangeF> eyr;               // per edge: y-extent of LSMF's box, rounded outward to single precision
    DBuf<double4> exy;               // per edge: (x0, y0, x1, y1), one 32-byte record for gathers

    // ---- A3 elementary intervals ----
    DBuf<double> ea;
    int epc = 0, nleaves = 0, T = 0, depth = 0;
    int ea_count = 0;                    // elementary points before padding
    DBuf<uint2> ep_lohi;                 // per endpoint: its elementary index range, or NIL (search)
    // ---- A4 ----
    DBuf<double> niv;
    // ---- A6 leaf end entries, split by class (0 base, 1 overlay) ----
    DBuf<int> after_leaf, before_leaf;   // node index or -1
    DBuf<uint32_t> lend_off;             // 2*(nleaves+1): class-major leaf offsets
    DBuf<uint32_t> lend_e;               // entries: edge index | 0x80000000 for the "after" entry,
                                         // sorted by (class, leaf), edge order inside a leaf
    DBuf<uint64_t> lend_key;             // (class, leaf) key of each entry
    DBuf<uint64_t> lend_key_alt;         // sort partners of lend_key / lend_e
    DBuf<uint32_t> lend_e_alt;
    DBuf<int> lend_bl;                   // the "before" leaf of each entry's edge
    DBuf<YRangeF> lend_yr;               // the edge's single-precision y-extent, beside each entry
    int64_t n_lend = 0;
    // materialised end lists (parity / drop-in only)
    DBuf<uint32_t> mend_off;
    DBuf<uint32_t> mend_e;
    int64_t n_mend = 0;
    // ---- A5 cover lists ----
    DBuf<uint32_t> cov_off;              // T+1
    DBuf<uint32_t> cov_e;                // edge indices: base run then overlay run, order within a run unspecified
    DBuf<uint32_t> cov_split;            // per node: count of base entries
    DBuf<YRangeF> cov_yr;                // the edge's single-precision y-extent, beside each cover entry
    DBuf<uint64_t> cov_key, cov_key_alt; // (node, class) << 32 | edge sort keys, and their sort partner
    int64_t n_cov = 0;
    // ---- A8/A10 candidates and unique pairs ----
    DBuf<uint64_t> unit_desc;            // stab work units: 32-byte records
    DBuf<uint64_t> cand;                 // packed (k, b, c) keys
    DBuf<uint64_t> cand_alt;             // its sort partner, same capacity
    int64_t n_cand = 0, cand_cap = 0;
    DBuf<uint2> near;                    // stab: pairs whose boxes may meet, (row edge | swap bit, column edge),
                                         // in NEAR_SEGS segments of near_seg_cap entries
    DBuf<uint64_t> near_cnt;             // per segment: its counter (one per 128-byte line), then 3 totals
    DBuf<uint32_t> srow_off, srow_node;  // stab rows of small rectangles: per-node offsets, row -> node
    int64_t n_near = 0, near_seg_cap = 0;
    int64_t n_tests = 0;                 // W
    int64_t n_units = 0;                 // stab work units (waves)
    DBuf<uint64_t> pairs;                // unique packed keys, sorted
    int64_t n_pairs = 0;
    int bits_b = 0, bits_c = 0, bits_k = 0;
    std::vector<int64_t> pair_off;       // per overlay [K+1] offsets into pairs (host)
    DBuf<uint32_t> d_pair_off;

    // ---- post-processing (A11-A17) of the last batch of jobs ----
    JobSet jobs;
    int cur_k = 0;                       // overlay of a one-job batch, else 0
    int I = 0;                           // unique pairs of the batch
    DBuf<int> p_j, p_b, p_c, p_type;     // job, base / overlay edge within its ring, type
    DBuf<double> p_x, p_y;               // intersection point
    DBuf<int> p_qa, p_qb;                // quantised alpha (B side), beta (C side)
    DBuf<int> ordB, ordC;                // per-edge alpha order (positions into the pair list)
    DBuf<uint64_t> ord_key, ord_key_alt; // (edge slot, alpha) sort keys and their sort partners
    DBuf<uint32_t> ord_val, ord_val_alt;
    DBuf<int> arr_b, arr_c;              // arrival of each pair inside its edge group (-1: not counted)
    DBuf<uint64_t> seg_kb, seg_kc;       // (alpha, pair) keys at their edge groups' positions
    DBuf<uint32_t> seg_wb, seg_wc, seg_bb, seg_bc, seg_ctl; // groups sorted by a wave / a block; list counts
    DBuf<int> nd_b, nd_c;                // non-degenerate counts -> exclusive scans (n+1 per job)
    DBuf<int> all_b, all_c;
    int nbaug = 0, ncaug = 0;
    DBuf<double> aug_bx, aug_by, aug_cx, aug_cy;
    DBuf<int> nbr_b, nbr_c, av_b, av_c, lab_b, lab_c;
    DBuf<int> slot_b, slot_c;            // per pair: augmented slot on each ring
    bool prov = false;                   // record slot provenance (for the area gradient)
    // sparse base ring (one-to-many batches without provenance): per position
    // (pair in (job, base edge, alpha) order) its virtual slot and slot data
    bool sparseP = false;
    DBuf<int> sp_slot, sp_ndx, sp_edge, sp_av, sp_nbr, sp_lab, sp_pos, sp_ndf;
    DBuf<double> sp_x, sp_y;
    DBuf<int> sv_b, sp_b, sv_c, sp_c;    // per slot: source vertex (local) or pair index, else -1
    DBuf<int> flags;                     // device status words
    int nrings = 0;
    std::vector<int> ring_off;           // host copy of the result
    std::vector<double> ring_xy;
    DBuf<int> d_ring_off;
    DBuf<double> d_ring_x, d_ring_y;
    DBuf<int> d_ring_src;                // provenance of each output vertex (P slot g, Q slot ~g)
    int64_t n_out = 0;
    int undefined = 0;                   // bitmask like the oracle's status

    // scratch
    RadixScratch rs;
    DBuf<uint64_t> k64a, k64b;
    DBuf<uint32_t> u32a, u32b, u32c, u32d;
    DBuf<uint64_t> u64a, u64b;
    DBuf<uint32_t> d_u32_total;
    DBuf<uint64_t> d_u64_total;

    // second stream (default priority) for work that runs beside the main one
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    void ensure_side_stream();

    // tuning: 64-column tiles per stab work unit (1..16); results do not depend on it
    int stab_tiles = 2;
    // tuning: consecutive stab work units per wave (1..32); results do not depend on it
    int stab_upw = 2;
    // independent pairs of up to 64 + 64 vertices clipped by the persistent
    // small-pair kernel (1) or the staged path (0); results do not depend on it
    int uni = 0; // operation: 0 intersection, 1 union (UNION of foster/polyclip.h:39)
    int small_fused = 1;
    // sequential-route rings walked one per lane (1) or by the single walk
    // (0); results do not depend on it
    int seq_par = 1;
    int small_fail = 0; // fallback reasons of the last small-pair batch (0: it ran fused)
    // the small-pair forward captured as a hipGraph (segclip_pairs_area_graph):
    // the buffers it reads and writes belong to it, so a replay never sees a
    // reallocation; destroyed when the polygons change
    struct PairsGraph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        const double *bxy = nullptr, *cxy = nullptr;
        double *out4 = nullptr;
        DBuf<int> tables;      // job table and the rings' input / output offsets
        DBuf<uint32_t> words;  // queue counter, fallback word
        DBuf<int> fail, route, nv, nr, olen;
        DBuf<double> ox, oy;
    } pg;

    bool timing = false;
    StageTimes times;
    hipEvent_t ev[10] = {};
    bool ev_ready = false;
    void ensure_events();
    bool flags_pending = false; // jobs.flags / undefined still on the device
    // pinned host staging for read-backs (one buffer per host thread, kept
    // for the process: contexts come and go, pinned allocations are slow),
    // and a wait that spins on the stream (no sleep between the last kernel
    // and the next launch)
    static void *pinned(size_t bytes);
    void wait();
    void mark(int i) {
        if (timing) SC_HIP(hipEventRecord(ev[i], stream));
    }
    float elapsed(int a, int b);

    ~Ctx();
};

// stage entry points (throw sc::Error)
void ctx_set_polygons(Ctx &c, const double *bxy, int nb, const double *cxy, const int *csz, int K, bool device_ptrs);
void ctx_set_pairs(Ctx &c, const double *bxy, const int *bsz, const double *cxy, const int *csz, int P,
                   bool device_ptrs);
inline int job_base(const Ctx &c, int k) { return c.paired ? k - 1 : 0; } // base ring of overlay k (1-based)
void build_edges(Ctx &c);
void build_elementary(Ctx &c);
void build_node_intervals(Ctx &c);
void build_leaf_end_entries(Ctx &c);
void materialize_end_lists(Ctx &c);
void build_cover_lists(Ctx &c);
void find_candidates(Ctx &c);          // all overlays at once, then unique pairs
void post_jobs(Ctx &c, int k0, int k1); // A11-A15 for overlays k0..k1 in one batch
// the batch's per-job flags (jobs.flags, undefined) from the device, when the
// post stage left them there (labelling fetches them with its first read-back)
void fetch_flags(Ctx &c);
SparseP sparse_view(Ctx &c);           // the sparse base ring of the last batch
void post_overlay(Ctx &c, int k);      // A11-A15 for overlay k
void label_and_traverse(Ctx &c);       // A16-A17 for the jobs of the last post_jobs
void clip_all(Ctx &c, int k);          // full pipeline for one overlay
void clip_jobs(Ctx &c, int k0, int k1); // full pipeline for overlays k0..k1 as one batch
// paired batches of small rings in one persistent launch: whether batch k0..k1
// qualifies, and the clip (false: some pair needs the staged path; nothing
// of the batch's results was written)
bool small_batch_eligible(const Ctx &c, int k0, int k1);
bool clip_small_batch(Ctx &c, int k0, int k1);
// forward areas of every pair through the captured graph (captured on first
// use and when the pointers change); a pair the kernel does not cover makes the
// call rerun the batch on the staged path
void pairs_area_graph(Ctx &c, const double *bxy, const double *cxy, double *out4);
void pairs_graph_reset(Ctx &c);
// A19: per job of the last batch, areas (intersection, base ring, overlay ring)
// and IoU into out4 (device, 4*J doubles); gradient of sum(g_area*area +
// g_iou*iou) w.r.t. every input vertex into grad_b (2*nb) and grad_c (2*nct)
void batch_area(Ctx &c, double *out4);
void batch_area_backward(Ctx &c, const double *g_area, const double *g_iou, double *grad_b, double *grad_c);
void release_derived(Ctx &c);          // free everything but the resident input rings
// the reference's per-overlay array stages on caller (host) arrays
void arr_save_intersections(Ctx &c, const int *bid, const int *cid, int n, double *xy, int *type, int *qa, int *qb);
void arr_sort_at_edges(Ctx &c, const int *alpha, const int *sorted_index, int c_start, const int *all_counts,
                       const int *nd_counts, int ncounts, const int *unique_ids, int unique_count, int n, int *out);
void arr_copy_intersections(Ctx &c, int b_size, int c_size, int c_start, const int *b_ids, const int *c_ids, int n,
                            const int *b_sorted_index, const int *c_sorted_index, const int *sorted_b,
                            const int *sorted_c, const double *inters, const int *types, const int *b_nd_ps,
                            const int *c_nd_ps, const int *qa, const int *qb, double *inter_b, double *inter_c,
                            int *nbr_b, int *nbr_c, int *alpha_b, int *alpha_c);
void arr_init_labels(Ctx &c, int nba, int nca, const double *inter_b, const double *inter_c, const int *alpha_b,
                     const int *nbr_b, int *lab_b, int *lab_c);
// calculateInitLabelParallel on one pair of augmented rings given as device arrays
void init_labels_one(Ctx &c, int nba, int nca, const double *bx, const double *by, const int *avb, const int *nbrb,
                     const double *cx, const double *cy, int *labb, int *labc, int *flags);
void release_foster_scratch();

// Device-to-host read-backs through the context's pinned staging, one wait
// for all of them: add() up to RB_MAX pieces, then run().
constexpr int RB_MAX = 4;
struct ReadBack {
    Ctx &c;
    int n = 0;
    void *dst[RB_MAX];
    const void *src[RB_MAX];
    size_t bytes[RB_MAX];
    explicit ReadBack(Ctx &ctx) : c(ctx) {}
    template <class T> void add(T *h, const T *d, size_t count) {
        dst[n] = h;
        src[n] = d;
        bytes[n] = sizeof(T) * count;
        ++n;
    }
    void run();
};
uint32_t read_u32(Ctx &c, const uint32_t *d);
uint64_t read_u64(Ctx &c, const uint64_t *d);

} // namespace sc
