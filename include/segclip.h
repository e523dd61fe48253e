/*
 * segclip.h -- C ABI of libsegclip, the MI355X (gfx950) segment-tree polygon
 * clipping path.
 *
 * All entry points take plain pointers and sizes, return an int status
 * (SEGCLIP_OK = 0) and never exit the process; segclip_last_error() gives
 * the message of the last failure on a context.  One context holds one
 * problem (a base ring and K overlay rings) resident on one GPU and runs its
 * kernels in order on one HIP stream.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   segclip_set_polygons        gpc_read_multi_polygon (src/clip.cpp:172-311) and
 *                               SegTree::copyInputDataToGPU (oacc_segment_tree.cpp:2670-2738)
 *   segclip_build_elementary    SegTree::constructElementaryIntervalsArrayGPU
 *                               (oacc:3997-4510; multicore twin oacc:4513-4692)
 *   segclip_build_node_intervals SegTree::initHybrid + constructTreeIntervalArray
 *                               (segment_tree.h:183-199, oacc:4865-4943)
 *   segclip_build_end_lists     SegTree::insertAllToEdgeListParallel (oacc:1339-1709)
 *   segclip_materialize_end_lists  the per-level end-list array of buildInternalEndListsGPU
 *                               (oacc:1404-1616), for parity checks and drop-in callers
 *   segclip_build_cover_lists   SegTree::constructCoverListGPU (oacc:2958-3128)
 *   segclip_find_candidates     countIntersectionsParallel + thrustDeviceExclusiveScan +
 *                               polyClipArrayGPUParallel (oacc:6250-6681, clip:1032-1089)
 *                               and radixsort/removeDuplicatesAtEdge (clip:368-493)
 *   segclip_intersections       countEdgeIntersections (clip:522-601), saveIntersectionsParallel,
 *                               sortIntersectionsAtEdgesParallel, copyIntersectionsParallel,
 *                               calculateInitLabelParallel (oacc:7257-7988)
 *   segclip_label_traverse      copyToLinkedList (clip:606-695), labelIntersections2,
 *                               createResult2, cleanUpResult2 (foster/polyclip.h:337-787)
 *   segclip_clip                the per-overlay body of hybrid_clipping (clip:1026-1297)
 *   segclip_clip_batch          the per-overlay loop of the one-to-many mode (clip:1702-2062)
 *                               as one batch of launches
 *   segclip_stable_sort_by_key_* / segclip_exclusive_scan_*
 *                               thrust_func.h:4-12 (p_seg_tree/thrust_func.cu:19-42)
 */
#ifndef SEGCLIP_H
#define SEGCLIP_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define SEGCLIP_OK 0
#define SEGCLIP_EINVAL 1  /* bad argument or input polygon */
#define SEGCLIP_EHIP 2    /* HIP runtime failure */
#define SEGCLIP_ENOMEM 3  /* device allocation failed */
#define SEGCLIP_ESTATE 4  /* stage called before the one it depends on */
#define SEGCLIP_ERANGE 5  /* sizes beyond the supported range */
#define SEGCLIP_EUNDEF 6  /* the reference's result is undefined on this input */
#define SEGCLIP_ESPACE 7  /* output buffer too small; the needed size was stored */

typedef struct segclip_ctx segclip_ctx;

/* Create a context on HIP device 'device'.  'stream' is a hipStream_t or
 * NULL (the context then creates its own default-priority stream). */
int segclip_create(segclip_ctx **out, int device, void *stream);
void segclip_destroy(segclip_ctx *ctx);
const char *segclip_last_error(const segclip_ctx *ctx);
const char *segclip_version(void);

/* Input rings as x,y pairs without repeated closing vertex.  ov_xy holds the
 * K overlays back to back with ov_sizes[k] vertices each.  Host pointers
 * (set_polygons) or device pointers (set_polygons_device). */
int segclip_set_polygons(segclip_ctx *ctx, const double *base_xy, int nb, const double *ov_xy,
                         const int *ov_sizes, int K);
int segclip_set_polygons_device(segclip_ctx *ctx, const double *base_xy, int nb, const double *ov_xy,
                                const int *ov_sizes, int K);
/* P independent pairs: pair k clips base ring k (base_sizes[k] vertices, the
 * rings back to back in base_xy) against overlay ring k.  One tree holds all
 * pairs, their x keys moved apart so that no tree node mixes two pairs; every
 * pair's results equal those of clipping it alone (bin/seg run once per pair,
 * multicore_clipping clip:1415-2115).  Pair ids in segclip_copy_pairs are
 * then local to the pair's two rings. */
int segclip_set_pairs(segclip_ctx *ctx, const double *base_xy, const int *base_sizes, const double *ov_xy,
                      const int *ov_sizes, int P);
int segclip_set_pairs_device(segclip_ctx *ctx, const double *base_xy, const int *base_sizes,
                             const double *ov_xy, const int *ov_sizes, int P);

/* The operation of the following clips: SEGCLIP_OP_INTERSECTION (default, the
 * reference's bin/seg) or SEGCLIP_OP_UNION (the UNION switch of
 * foster/polyclip.h:39 with its XORs in labelIntersections2 / createResult2,
 * polyclip.h:451-470,571-618,669-681,740; the clipper's -union,
 * lib/optimizedFostersAlgorithm/polyclip.cpp).  Union jobs take the
 * sequential labelling route; a union output may hold holes, so the
 * area / IoU entry points refuse it (SEGCLIP_ESTATE). */
#define SEGCLIP_OP_INTERSECTION 0
#define SEGCLIP_OP_UNION 1
int segclip_set_operation(segclip_ctx *ctx, int op);

/* Tree stages.  segclip_build_tree runs elementary, node intervals, end lists
 * and cover lists in that order. */
int segclip_build_elementary(segclip_ctx *ctx);
int segclip_build_node_intervals(segclip_ctx *ctx);
int segclip_build_end_lists(segclip_ctx *ctx);
int segclip_build_cover_lists(segclip_ctx *ctx);
int segclip_materialize_end_lists(segclip_ctx *ctx);
int segclip_build_tree(segclip_ctx *ctx);

/* Candidate pairs of every overlay in one pass (stab + dedupe). */
int segclip_find_candidates(segclip_ctx *ctx);
/* Per-overlay bookkeeping (k is 1-based) and labelling / traversal. */
int segclip_intersections(segclip_ctx *ctx, int k);
int segclip_label_traverse(segclip_ctx *ctx);
/* Whole path for overlay k: the tree and candidates are built on first use. */
int segclip_clip(segclip_ctx *ctx, int k);
/* Whole path for overlays (or pairs) k0..k1 as one batch: intersections,
 * labelling, traversal and clean-up of all of them in the same launches.
 * Results equal those of segclip_clip(k) for each k; the output rings of job
 * k are rings [job_ring_off[k-k0], job_ring_off[k-k0+1]) of segclip_copy_rings. */
int segclip_clip_batch(segclip_ctx *ctx, int k0, int k1);
/* The reference's per-overlay array stages on the caller's host arrays, with
 * the reference's argument meaning (pair ids = 2 * vertex offset into bPoly /
 * cPoly with closing vertices; the context's resident rings are used).  They
 * back the SegTree facade (segtree.h) for drivers that keep their own arrays.
 *   segclip_save_intersections_ids       SegTree::saveIntersectionsParallel (oacc:7257-7341)
 *   segclip_sort_intersections_at_edges  SegTree::sortIntersectionsAtEdgesParallel (oacc:7436-7473);
 *                                        ncounts = entries of all_counts / nd_counts
 *   segclip_copy_intersections_arrays    SegTree::copyIntersectionsParallel (oacc:7508-7664);
 *                                        all overlay source vertices are copied
 *   segclip_init_labels_arrays           SegTree::calculateInitLabelParallel (oacc:7925-7988) */
int segclip_save_intersections_ids(segclip_ctx *ctx, const int32_t *b_id, const int32_t *c_id, int n, double *xy,
                                   int32_t *type, int32_t *alpha_b, int32_t *alpha_c);
int segclip_sort_intersections_at_edges(segclip_ctx *ctx, const int32_t *alpha, const int32_t *sorted_index,
                                        int c_start, const int32_t *all_counts, const int32_t *nd_counts,
                                        int ncounts, const int32_t *unique_ids, int unique_count, int n,
                                        int32_t *sorted_alpha_index);
int segclip_copy_intersections_arrays(segclip_ctx *ctx, int b_size, int c_size, int c_start, const int32_t *b_ids,
                                      const int32_t *c_ids, int n, const int32_t *b_sorted_index,
                                      const int32_t *c_sorted_index, const int32_t *sorted_b,
                                      const int32_t *sorted_c, const double *intersections, const int32_t *types,
                                      const int32_t *b_nd_ps, const int32_t *c_nd_ps, const int32_t *alpha_b,
                                      const int32_t *alpha_c, double *intersected_b, double *intersected_c,
                                      int32_t *neighbors_b, int32_t *neighbors_c, int32_t *alpha_values_b,
                                      int32_t *alpha_values_c);
int segclip_init_labels_arrays(segclip_ctx *ctx, int nb_aug, int nc_aug, const double *intersected_b,
                               const double *intersected_c, const int32_t *alpha_values_b,
                               const int32_t *neighbors_b, int32_t *labels_b, int32_t *labels_c);

/* Differentiable area / IoU (no counterpart in the reference, which only
 * writes the rings).  With provenance on, the intersection stage records
 * which input vertex or intersecting pair every augmented slot came from.
 * segclip_batch_area gives, per job of the last batch, 4*jobs doubles:
 * intersection areas, base ring areas, overlay ring areas, IoU (shoelace,
 * FP64).  The backward fills grad_b (2*nb doubles, the layout of the base
 * input) and grad_c (2*sum(ov_sizes), the layout of the overlay input) with
 * d/d(vertex) of sum_j g_area[j]*area[j] + g_iou[j]*iou[j], the output
 * combinatorics held fixed; g_area or g_iou may be null (zero).  Device
 * pointers, on the context's device; the call returns after its work is done. */
int segclip_set_provenance(segclip_ctx *ctx, int on);
int segclip_batch_area(segclip_ctx *ctx, double *out4, int64_t cap, int64_t *need);
int segclip_batch_area_device(segclip_ctx *ctx, double *out4);
int segclip_batch_area_backward_device(segclip_ctx *ctx, const double *g_area, const double *g_iou, double *grad_b,
                                       double *grad_c);
/* Forward areas / IoU (the segclip_batch_area layout, 4*P doubles in out4) of
 * every pair of a paired context whose rings all have at most 64 vertices,
 * replayed as one hipGraph on the context's stream: re-layout of the caller's
 * rings, queue reset and the persistent small-pair kernel with the areas
 * computed in place.  base_xy / ov_xy are device pointers with the layout and
 * sizes of the last segclip_set_pairs* call (rings back to back, no closing
 * vertex); they are read at every replay, so the caller may update them in
 * place between calls.  The graph is captured on the first call, again when
 * a pointer changes, and dropped by segclip_set_pairs* / set_polygons*.
 * After the replay one word is read back: when some pair needs the staged
 * path, the batch is clipped on it (areas identical either way) and the next
 * call captures again.  Returns when out4 is written.  No counterpart in the
 * reference (its per-pair loop launches every stage, clip:1026-1297). */
int segclip_pairs_area_graph(segclip_ctx *ctx, const double *base_xy, const double *ov_xy, double *out4);
/* Free every device buffer of the context but the resident input rings, the
 * last results (output rings and status words, so they stay readable) and a
 * captured pairs graph's own buffers; the next call rebuilds what it needs.
 * segclip_memory reports the device bytes held by the library (all contexts)
 * and their peak since the last reset. */
int segclip_release(segclip_ctx *ctx);
int segclip_memory(int64_t *current, int64_t *peak, int reset_peak);
/* Device allocator for the library's buffers (process-wide), for programs
 * that pool device memory (a framework's caching allocator).  Buffers
 * allocated while set come from alloc (null: out of memory) and go back
 * through free, which is called only after the library's queued device work
 * has finished.  Null alloc restores hipMalloc/hipFree for later buffers;
 * buffers keep the allocator they came from, so the hooks must stay callable
 * until those are released.  No counterpart in the reference (its buffers
 * are host arrays and OpenACC copies). */
typedef void *(*segclip_alloc_fn)(size_t bytes, void *user);
typedef void (*segclip_free_fn)(void *ptr, void *user);
int segclip_set_allocator(segclip_alloc_fn alloc, segclip_free_fn free_fn, void *user);
/* per job of the last batch (or the single job of segclip_clip): ring
 * offsets (jobs+1 entries) and status bits (the 'undefined' bits per job) */
int segclip_copy_batch_rings(segclip_ctx *ctx, int32_t *job_ring_off, int64_t cap, int64_t *need);
int segclip_copy_batch_status(segclip_ctx *ctx, int32_t *status, int64_t cap, int64_t *need);
/* per job of the last batch: labelling route (0 parallel crossing traversal,
 * 1 no crossing vertex, 2 the reference's sequential algorithm) and, for
 * route 2, why (bits: 1 a label other than CROSSING/BOUNCING, 2 a ring without
 * a plain vertex, 4 crossings not alternating, 8 a bouncing pair to split,
 * 16 a neighbour link that is not a crossing, 32 the union) */
int segclip_copy_batch_routes(segclip_ctx *ctx, int32_t *route, int32_t *reason, int64_t cap, int64_t *need);
/* Drop everything derived from the polygons (tree, candidates, results) so
 * that the next segclip_clip recomputes it; the resident polygons stay. */
int segclip_reset(segclip_ctx *ctx);

typedef struct {
    int64_t num_edges;        /* N */
    int64_t epc;              /* elementary points (2^k + 1) */
    int64_t tree_size;        /* T */
    int64_t cover_entries;    /* L */
    int64_t leaf_end_entries; /* leaf-level end-list entries */
    int64_t end_entries;      /* E, after segclip_materialize_end_lists */
    int64_t tests;            /* W, pair tests of the stab */
    int64_t candidates;       /* K, emitted (with duplicates) */
    int64_t pairs;            /* unique pairs, all overlays */
    int32_t overlay;          /* overlay of the last segclip_intersections */
    int32_t overlay_pairs;    /* I of that overlay */
    int32_t nb_aug, nc_aug;   /* augmented ring sizes */
    int32_t rings;            /* output components after clean-up */
    int32_t undefined;        /* bit 1: an intersection got no slot; 2: ring walk; 4: traversal;
                                 16: a quantised alpha or beta beyond the int range */
    int64_t out_vertices;
    int64_t stab_units;       /* work units (waves) of the stab */
    int64_t near_pairs;       /* stab pairs whose single-precision y-extents meet (classified in FP64) */
} segclip_counts;
int segclip_get_counts(segclip_ctx *ctx, segclip_counts *out);

/* Copy-out (device -> host).  Each returns SEGCLIP_ESPACE and stores the
 * needed element count in *need when 'cap' is too small. */
int segclip_copy_elementary(segclip_ctx *ctx, double *out, int64_t cap, int64_t *need);
int segclip_copy_node_intervals(segclip_ctx *ctx, double *out, int64_t cap, int64_t *need);
/* cover lists: off has tree_size+1 entries; edges are edge indices (base
 * edges first, then overlays in order), unordered within a node */
int segclip_copy_cover_lists(segclip_ctx *ctx, uint32_t *off, uint32_t *edges, int64_t cap, int64_t *need);
/* end lists per node sorted by (id,type); needs segclip_materialize_end_lists */
int segclip_copy_end_lists(segclip_ctx *ctx, uint32_t *off, uint32_t *edges, int64_t cap, int64_t *need);
/* per edge "after"/"before" leaf node of its leaf end entries (-1 = none) */
int segclip_copy_leaf_targets(segclip_ctx *ctx, int32_t *after, int32_t *before, int64_t cap, int64_t *need);
/* unique pairs of overlay k as reference ids (2*vertex offset) and types;
 * the types are those of segclip_clip(k) when k is the overlay clipped last,
 * else -1 (a batch keeps no per-pair types) */
int segclip_copy_pairs(segclip_ctx *ctx, int k, int32_t *b_id, int32_t *c_id, int32_t *type, int64_t cap,
                       int64_t *need);
/* intersections of the current overlay, in pair order */
int segclip_copy_intersections(segclip_ctx *ctx, double *xy, int32_t *alpha_b, int32_t *alpha_c, int64_t cap,
                               int64_t *need);
/* augmented rings: xy (2*n), neighbours (1-based, -100), alpha values (-100 = plain), initial labels */
int segclip_copy_augmented(segclip_ctx *ctx, int side, double *xy, int32_t *nbr, int32_t *alpha, int32_t *label,
                           int64_t cap, int64_t *need);
/* output rings: ring_off has rings+1 entries, xy 2*out_vertices */
int segclip_copy_rings(segclip_ctx *ctx, int32_t *ring_off, int64_t off_cap, double *xy, int64_t xy_cap,
                       int64_t *need);

/* thrust_func.h primitives on device pointers (stable, in place) */
int segclip_stable_sort_by_key_f64(double *keys, int32_t *vals, int64_t n, void *stream);
int segclip_stable_sort_by_key_i32(int32_t *keys, int32_t *vals, int64_t n, void *stream);
int segclip_stable_sort_by_key_i64(int64_t *keys, int32_t *vals, int64_t n, void *stream);
int segclip_exclusive_scan_i32(int32_t *data, int64_t n, void *stream);
int segclip_exclusive_scan_i64(int64_t *data, int64_t n, void *stream);

/* Candidate pairs of two polygon layers by bounding-box overlap, the filter in
 * front of many-to-many clipping (the reference joins layers through an R-tree
 * of polygon MBRs: lib/optimizedFostersAlgorithm/readShapefile.cpp:326-355,
 * foster/Rtree-Topdown-version; getCMBR's closed test, polyclip_time.h:95-124).
 * a_xy / b_xy: device pointers to rings back to back (x,y doubles, no closing
 * vertex), a_sizes / b_sizes: host vertex counts.  pairs (host, 2*cap ints)
 * receives every (a, b) whose MBRs overlap (closed intervals), sorted by
 * (a, b); *need = their number (SEGCLIP_ESPACE when cap is smaller).  Returns
 * after the work on 'stream' is done. */
int segclip_mbr_join(const double *a_xy, const int *a_sizes, int na, const double *b_xy, const int *b_sizes,
                     int nb, int32_t *pairs, int64_t cap, int64_t *need, void *stream);

/* segclip_stable_sort_by_key_f64 without any device allocation.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same result as segclip_stable_sort_by_key_f64 (stable,
 * in place, -0.0 == 0.0), for n <= 1,048,576 keys; a larger n returns
 * SEGCLIP_ERANGE.  Its scratch is one fixed device array of about 38 MB per
 * GPU, held until the process ends and shared by every call on that GPU:
 * calls that overlap on one GPU race.  Asynchronous on 'stream' (no host
 * wait); clears its scratch with kernels, not memset nodes. */
int segclip_stable_sort_by_key_f64_noalloc(double *keys, int32_t *vals, int64_t n, void *stream);

/* Seeded synthetic rings (host code; see the workload definitions) */
int segclip_synth_star(double *xy, int n, uint64_t seed, double cx, double cy, double rmin, double rmax,
                       double rot);
int segclip_synth_harmonic(double *xy, int n, uint64_t seed, double cx, double cy, double scale, double rot);
int segclip_synth_convex(double *xy, int n, uint64_t seed, double cx, double cy, double scale, double rot);
void segclip_synth_uniform(double *out, int n, uint64_t seed);
int segclip_count_vertical(const double *xy, int n);

/* Per-stage device times in ms of the last segclip_clip with timing enabled,
 * from HIP events on the context's stream: elementary (edges + sort + unique),
 * node intervals, leaf end lists, cover lists, candidates (stab + dedupe),
 * the stab kernel alone, intersections (A11-A15), labelling + traversal. */
int segclip_set_timing(segclip_ctx *ctx, int on);
int segclip_get_stage_times(segclip_ctx *ctx, float *ms8);

/* Tuning settings of a context, by name.  No counterpart in the reference
 * (its stab runs one tree node per parallel-loop iteration, countIntersectionsParallel,
 * oacc_segment_tree.cpp:6250-6526); they serve
 * the autotuner, which times candidate values on a workload and persists the
 * fastest as JSON (segclip.tune).  Results never depend on a setting.
 *   "stab_tiles"  64-column tiles per stab work unit, 1..16, default 2.
 *   "stab_units_per_wave"  consecutive stab work units one wave runs (near
 *                 pairs staged across them), 1..32, default 2.
 *   "small_fused" 1 (default): a batch of independent pairs whose rings all
 *                 have at most 64 vertices, without provenance, runs as one
 *                 persistent launch, each pair clipped whole in LDS; 0: the
 *                 staged path.  A batch holding a pair the persistent kernel
 *                 does not cover reruns on the staged path.
 *   "seq_parallel_walks"  1 (default): the traversal of a job on the
 *                 sequential route walks its rings one per lane and checks
 *                 that they are the ones the reference's single walk finds
 *                 (taking the single walk when they may not be); 0: the
 *                 single walk always.
 * segclip_get_tuning also reads "small_fallback": the fallback reasons of the
 * last small-pair batch (0 when it ran as one launch).
 * An unknown name or a value out of range returns SEGCLIP_EINVAL. */
int segclip_set_tuning(segclip_ctx *ctx, const char *name, int value);
int segclip_get_tuning(segclip_ctx *ctx, const char *name, int *value);

#ifdef __cplusplus
}
#endif
#endif
