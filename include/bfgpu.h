/*
 * bfgpu.h — C ABI of the MI355X K-lane batched PHAST engine.
 *
 * A drop-in for butterfly-route's K-lane batched PHAST over the customized
 * edge-based CCH (route/src/matrix/batched_phast.rs) and for the seeded,
 * bounded single-source PHAST behind isochrones and PHAST matrix rows
 * (route/src/server/isochrone_handler.rs:338-491).  The reference has no FFI
 * for this path; each entry point below names the Rust signature it replaces.
 *
 * Conventions (mirroring dl/include/butterfly.h:44-54,70-76 of the reference):
 *   - every call returns a bfg_result; 0 is success; bfg_last_error() gives the
 *     message of the last failure on the calling thread;
 *   - every buffer is caller-owned; inputs are copied, outputs are written in
 *     place; nothing returned needs freeing except the context;
 *   - nodes are CCH ranks (rank-aligned CCH: node id == rank), distances are
 *     u32 seconds, 0xFFFFFFFF (BFG_INF) means unreachable / beyond the limit;
 *   - a context takes calls from any number of threads: up to its concurrency
 *     (bfg_ctx_set_concurrency, default 4 or the environment variable
 *     BFG_STREAMS) run at once on the device, each on its own HIP stream with
 *     its own label arena; further calls wait for one to finish; the
 *     two-channel calls run one at a time per context.
 * Results are byte-identical to the reference CPU PHAST on the same CCH:
 * exact shortest up-down distances with saturating u32 adds; for bounded
 * calls every value <= the limit is exact and every other value is BFG_INF.
 */
#ifndef BFGPU_H
#define BFGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BFG_INF 0xFFFFFFFFu

typedef enum bfg_result {
    BFG_SUCCESS = 0,
    BFG_INVALID_PARAMETER = 1, /* bad pointer, size or rank >= n_nodes        */
    BFG_OUT_OF_MEMORY = 2,     /* device allocation failed                     */
    BFG_DEVICE_ERROR = 3,      /* HIP runtime or kernel failure                */
    BFG_NOT_SUPPORTED = 4,     /* valid request outside this build's limits    */
    BFG_INTERNAL_ERROR = 5,
    BFG_BUFFER_TOO_SMALL = 6   /* an output buffer holds less than the result; the sizes are reported */
} bfg_result;

/* The customized CCH as the reference holds it: CchTopo up/down CSR
 * (route/src/formats/cch_topo.rs:250-286) with CchWeights decoded to u32
 * (route/src/formats/cch_weights.rs:171-187, sentinel -> BFG_INF).
 * UP arcs at tail r have targets > r, DOWN arcs at tail r have targets < r.
 * INF-weight arcs are dropped on upload (bucket_ch.rs:191-424). */
typedef struct bfg_graph_desc {
    uint32_t n_nodes;
    const uint64_t *up_offsets;   /* n_nodes + 1 */
    const uint32_t *up_targets;
    const uint32_t *up_weights;
    const uint64_t *down_offsets; /* n_nodes + 1 */
    const uint32_t *down_targets;
    const uint32_t *down_weights;
} bfg_graph_desc;

typedef struct bfg_ctx bfg_ctx;

typedef struct bfg_ctx_info {
    uint32_t n_nodes;
    uint64_t m_up, m_down;          /* finite arcs after INF filtering           */
    uint32_t etree_depth;           /* deepest node of the elimination tree      */
    uint32_t n_top;                 /* nodes in the top region (separator chains) */
    uint32_t n_segments;            /* chain segments of the top region          */
    uint32_t n_segment_levels;      /* dependent stages of the top region        */
    uint32_t n_bottom_levels;       /* height levels of the bottom region        */
    uint32_t lanes;                 /* K: source lanes interleaved per group     */
    uint32_t lane_capacity;         /* sources per device batch                  */
    uint64_t device_bytes;          /* resident CCH + arenas                     */
} bfg_ctx_info;

/* Per-phase device time of the last call (HIP events on the context stream). */
typedef struct bfg_timings {
    float up_ms, sweep_ms, output_ms, total_ms;
    uint32_t batches, sweep_launches;
    uint64_t lanes_run;
    /* nodes and finite DOWN arcs one lane group's downward sweep covered: the
     * whole CCH, or for a matrix the top region plus the bottom paths above
     * its targets (a target's label depends on its ancestors alone) */
    uint64_t nodes_swept, arcs_swept;
    /* the dominant kernel (DOWN rows of the top region): summed durations of
     * its launches (HIP events on the context stream), launch count, and the
     * algorithmic bytes those launches move: per 64-lane wavefront group and
     * segment level, 8 per DOWN arc from outside the level's segments plus
     * 8*64 per node of the level (its labels written once, read once) */
    float dom_ms;
    uint32_t dom_launches;
    uint64_t dom_bytes;
    /* the block gate of bounded calls (0 without one), summed over batches:
     * (segment, 64-lane wave group) units of the top region and how many the
     * gate left open, and the same for the bottom trees */
    uint64_t gate_seg_units, gate_seg_open, gate_bt_units, gate_bt_open;
    /* bfg_table_bucket: the join kernel's duration (HIP events around its
     * launches) and the label positions it visited, summed over its 64 x 64
     * (source, target) tiles: each visit is 4096 min-plus cell updates */
    float join_ms;
    uint64_t join_visits;
    /* bfg_table_bucket: the label width the join ran at (16: packed u16
     * cells, labels capped at 0x7FFF, for 64-lane groups joined one wavefront
     * a tile; 32 otherwise) and the tiles with a cell at or past the cap that
     * the u32 join recomputed */
    uint32_t join_label_bits;
    uint64_t join_tiles_u32;
} bfg_timings;

/* BatchedPhastEngine::new (batched_phast.rs:89): copies the CCH to `device`,
 * derives the device layout and keeps it resident until destroy.  The
 * environment variable BFG_BOTTOM_LIMIT (a subtree size, default 256) tunes
 * where the layout switches from register tiles to per-node bottom sweeps. */
int bfg_ctx_create(const bfg_graph_desc *graph, int device, bfg_ctx **out);
void bfg_ctx_destroy(bfg_ctx *ctx);
const char *bfg_last_error(void);
int bfg_ctx_get_info(bfg_ctx *ctx, bfg_ctx_info *info);
int bfg_ctx_last_timings(bfg_ctx *ctx, bfg_timings *t);
/* Lane width K in {16, 32, 64} and the batch capacity in lanes (0 = auto;
 * rounded up to a multiple of 64).
 * Reallocates the arenas; used by the autotuner. */
int bfg_ctx_configure(bfg_ctx *ctx, uint32_t lanes, uint32_t lane_capacity);

/* Arithmetic width of the downward sweep's dense rows (no counterpart in the
 * reference): 16 (the default) relaxes two 64-lane groups per instruction in
 * packed 16-bit saturating arithmetic wherever every finite candidate of a
 * unit stays below 0xFFFF, and falls back to 32 bits elsewhere; 32 always
 * uses 32 bits.  Results are bit-identical either way; set by the autotuner. */
int bfg_ctx_set_label_width(bfg_ctx *ctx, uint32_t bits);

/* Calls of one context that may run on the device at once (1..16), each on
 * its own stream and label arena (no counterpart in the reference, whose
 * engines are per-thread; the serving loop overlaps its rounds with it). */
int bfg_ctx_set_concurrency(bfg_ctx *ctx, uint32_t max_calls);

/* Tuning of the downward sweep (no counterpart in the reference; set by the
 * autotuner, bfgpu/autotune.py): plan_rows = the most rows one DOWN unit
 * takes (>= 64; 0 = default 256 or BFG_PLAN_MAX_ROWS), tri_scalar_segs =
 * 1 + the most segments a DOWN level may have and still use the vector-loaded
 * intra-segment blocks (0 = default 64 or BFG_TRI_SCALAR_SEGS). */
int bfg_ctx_tune(bfg_ctx *ctx, uint32_t plan_rows, uint32_t tri_scalar_segs);

/* Host-only check of a CCH before upload (no device is touched): derives the
 * device layout, verifies that every finite arc joins an elimination-tree
 * ancestor/descendant pair and that the sweep schedule respects every arc,
 * and reports the layout figures (the role of the reference's validate-cch /
 * validate-phast CLI checks, route/src/cli.rs:1040-1090). */
int bfg_validate_graph(const bfg_graph_desc *graph, bfg_ctx_info *info);

/* BatchedPhastEngine::compute_matrix_flat (batched_phast.rs:667-706):
 * out[s * n_tgt + t] = d(sources[s], targets[t]), row-major, host memory. */
int bfg_matrix(bfg_ctx *ctx, const uint32_t *sources, uint32_t n_src,
               const uint32_t *targets, uint32_t n_tgt, uint32_t *out);

/* BatchedPhastEngine::compute_matrix_block_gated (batched_phast.rs:1223-1265):
 * as bfg_matrix, cells > threshold are BFG_INF. */
int bfg_matrix_bounded(bfg_ctx *ctx, const uint32_t *sources, uint32_t n_src,
                       const uint32_t *targets, uint32_t n_tgt, uint32_t threshold,
                       uint32_t *out);

/* Device-resident variant of the two above (no counterpart in the reference;
 * it is the same computation with HIP device pointers for callers that keep
 * queries and results on the GPU).  threshold = BFG_INF for unbounded. */
int bfg_matrix_device(bfg_ctx *ctx, const uint32_t *d_sources, uint32_t n_src,
                      const uint32_t *d_targets, uint32_t n_tgt, uint32_t threshold,
                      uint32_t *d_out);

/* Several compute_matrix_block_gated requests (batched_phast.rs:1223-1265) as
 * one lane batch (no counterpart in the reference, which serves each request
 * with its own engine call): request r has sources sources[src_offsets[r] ..
 * src_offsets[r+1]), targets targets[tgt_offsets[r] .. tgt_offsets[r+1]) and
 * threshold thresholds[r] (BFG_INF: unbounded); its S_r x T_r row-major
 * block is written to out from sum_{q < r} S_q * T_q on.  Every block equals
 * the request's own bfg_matrix_bounded result: the sources of all requests
 * share one downward sweep, and each request gathers only its own columns. */
int bfg_matrix_multi(bfg_ctx *ctx, uint32_t n_req, const uint32_t *src_offsets, const uint32_t *sources,
                     const uint32_t *tgt_offsets, const uint32_t *targets, const uint32_t *thresholds,
                     uint32_t *out);

/* bfg_matrix_device for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The call allocates nothing:
 * its device state lives in one static array of 512 MiB per GPU, which every
 * process using this library holds on each GPU it runs on until it exits.
 * Limits: at most 64 sources and 262144 targets per call, and the graph's
 * top region plus the bottom paths above the targets must fit the array
 * (64 lanes x about 2M labels); larger requests get BFG_NOT_SUPPORTED.
 * Calls on one GPU are serialised inside the library (they share the
 * array, whose contents each call overwrites; calls from other processes on
 * the same GPU have arrays of their own). */
int bfg_matrix_device_noalloc(bfg_ctx *ctx, const uint32_t *d_sources, uint32_t n_src, const uint32_t *d_targets,
                              uint32_t n_tgt, uint32_t threshold, uint32_t *d_out);

/* BatchedPhastEngine::query_batch / query_batch_soa (batched_phast.rs:131-260,
 * :470-609) for any number of sources: dist[i * n_nodes + v] = d(sources[i], v).
 * With threshold != BFG_INF this is query_batch_bounded / query_batch_block_gated
 * (:271-445, :946-1218): values > threshold are BFG_INF. */
int bfg_query_batch(bfg_ctx *ctx, const uint32_t *sources, uint32_t n_src,
                    uint32_t threshold, uint32_t *dist);

/* Block-gated isochrone sweeps with one limit per lane (the K-lane caller of
 * range/batched_isochrone.rs:121-226 generalised to uneven limits).
 * reach_bits[i * ceil(n/64) + v/64] bit v%64 set iff d(origins[i], v) <= limits[i];
 * times (nullable) [i * n_nodes + v] = that distance, BFG_INF otherwise. */
int bfg_isochrone(bfg_ctx *ctx, const uint32_t *origins, const uint32_t *limits,
                  uint32_t n_lanes, uint64_t *reach_bits, uint32_t *times);

/* run_phast_bounded_fast_seeded (isochrone_handler.rs:338-491) for n_lanes
 * independent queries, with the reference's output: lane i's settled nodes
 * (rank, dist <= thresholds[i]) in ascending rank (:453-471), as pairs
 * pairs[2k] = rank, pairs[2k+1] = dist for k in [list_start[i],
 * list_start[i] + list_len[i]).  Seeds as in bfg_phast_seeded.  reach_bits
 * (nullable) as in bfg_isochrone.  `capacity` counts pairs; *total (nullable)
 * receives the pairs of all lanes.  When they exceed capacity the call
 * returns BFG_BUFFER_TOO_SMALL with list_start, list_len and *total filled
 * and writes no pair past capacity; a second call with capacity >= *total
 * succeeds.  The downward sweep is block-gated: a (segment or bottom tree,
 * 64-lane wave group) whose lanes cannot settle a node within their limits is
 * skipped. */
int bfg_isochrone_lists(bfg_ctx *ctx, const uint32_t *seed_offsets, const uint32_t *seed_ranks,
                        const uint32_t *seed_costs, uint32_t n_lanes, const uint32_t *thresholds,
                        uint64_t *reach_bits, uint64_t *list_start, uint32_t *list_len, uint32_t *pairs,
                        uint64_t capacity, uint64_t *total);

/* Device-resident variant of bfg_isochrone_lists (no counterpart in the
 * reference): d_reach_bits (n_lanes * ceil(n/64) words) and d_pairs
 * (capacity pairs) are HIP device pointers; seeds, thresholds, list_start,
 * list_len and total stay in host memory. */
int bfg_isochrone_lists_device(bfg_ctx *ctx, const uint32_t *seed_offsets, const uint32_t *seed_ranks,
                               const uint32_t *seed_costs, uint32_t n_lanes, const uint32_t *thresholds,
                               uint64_t *d_reach_bits, uint64_t *list_start, uint32_t *list_len, uint32_t *d_pairs,
                               uint64_t capacity, uint64_t *total);

/* run_phast_bounded_fast_seeded (isochrone_handler.rs:338-491) for n_lanes
 * independent queries: lane i has seeds (seed_ranks[k], seed_costs[k]) for
 * k in [seed_offsets[i], seed_offsets[i+1]) (at most 64 per lane) and limit
 * thresholds[i].  dist[i * n_nodes + v] = bounded distance, BFG_INF beyond. */
int bfg_phast_seeded(bfg_ctx *ctx, const uint32_t *seed_offsets, const uint32_t *seed_ranks,
                     const uint32_t *seed_costs, uint32_t n_lanes, const uint32_t *thresholds,
                     uint32_t *dist);

/* Seeded S x T table through one seeded PHAST field per source: the
 * table_phast_lopsided path (matrix/bucket_ch.rs:6717-6853) that
 * table_seeded_bounded_routed (:6535-6574) takes when phast_wins (:6677-6711).
 * Source i has seeds (src_ranks[k], src_parts[k]) for k in
 * [src_offsets[i], src_offsets[i+1]) (at most 64), target j the same in
 * tgt_*.  The field of source i is run_phast_bounded_fast_seeded of its
 * seeds, bounded by threshold + its largest part + the largest target part
 * (saturating; unbounded when threshold == BFG_INF, :6738-6752);
 * out[i * n_tgt + j] = min over target seeds (r, p) with a field value F[r]
 * of F[r] - p (saturating), BFG_INF above threshold (:6785-6811).  A pair
 * whose source and target share a seed rank is the reference's same-edge
 * conflict, which it recomputes with its bucket engine (:6789-6793,
 * :6828-6845): its cell is BFG_INF here and conflicts[i * n_tgt + j] (when
 * conflicts is not null) is 1, else 0. */
int bfg_table_seeded(bfg_ctx *ctx, const uint32_t *src_offsets, const uint32_t *src_ranks,
                     const uint32_t *src_parts, uint32_t n_src, const uint32_t *tgt_offsets,
                     const uint32_t *tgt_ranks, const uint32_t *tgt_parts, uint32_t n_tgt,
                     uint32_t threshold, uint32_t *out, uint8_t *conflicts);

/* The bucket side of the matrix router (table_seeded_bounded_routed sends
 * matrices to it unless phast_wins, matrix/bucket_ch.rs:6535-6574).
 * bfg_ctx_pair_reverse gives ctx the context over the transposed graph
 * (UP arcs = this graph's DOWN arcs reversed, DOWN arcs = its UP arcs
 * reversed: the reference's DownReverseAdjFlat, bucket_ch.rs:191-424), whose
 * upward phase is the bucket engine's backward search.  Both contexts must
 * stay alive while paired; destroying either ends the pair, and rev == NULL
 * unpairs.  A context paired as a reverse cannot itself be paired.  Returns
 * BFG_INVALID_PARAMETER when rev is not over the transposed graph (different
 * positions or arc counts) or is on another device. */
int bfg_ctx_pair_reverse(bfg_ctx *ctx, bfg_ctx *rev);

/* Diagnostics (no counterpart in the reference): the number of device
 * allocations (hipMalloc) the library has made in this process, so a caller
 * can check that a call allocated nothing. */
int bfg_device_allocations(uint64_t *count);

/* bfg_ctx_create over the transposed graph of `graph` (the reverse metric's
 * CCH, whose forward PHAST is the reference's reverse PHAST,
 * range/phast.rs:879-1108, and whose upward phase is the bucket engine's
 * backward search over DownReverseAdjFlat, bucket_ch.rs:191-424): the
 * library transposes the CSR itself (per-tail targets ascending). */
int bfg_ctx_create_reverse(const bfg_graph_desc *graph, int device, bfg_ctx **out);

/* table_bucket_parallel_bounded(n_nodes, up_adj_flat, down_rev_flat, sources,
 * targets, threshold) (matrix/bucket_ch.rs:3436-3520): the forward UP search
 * of every source and the backward search of every target meet at their
 * common elimination-tree ancestors; out[i * n_tgt + j] = min over those of
 * fw + bw (saturating), BFG_INF when none or above threshold.  The cells equal
 * bfg_matrix_bounded's.  Needs a paired reverse context (BFG_NOT_SUPPORTED
 * otherwise); sources lease a Work of ctx, targets one of the reverse
 * context. */
int bfg_table_bucket(bfg_ctx *ctx, const uint32_t *sources, uint32_t n_src, const uint32_t *targets,
                     uint32_t n_tgt, uint32_t threshold, uint32_t *out);

/* bfg_table_bucket with sources, targets and the S x T result in HIP device
 * memory. */
int bfg_table_bucket_device(bfg_ctx *ctx, const uint32_t *d_sources, uint32_t n_src, const uint32_t *d_targets,
                            uint32_t n_tgt, uint32_t threshold, uint32_t *d_out);

/* Length-along-time weights for the two-channel calls: the CchWeights the
 * reference loads beside the time weights (cch.lat.<mode>.u32,
 * server/state.rs:2880-2891) and reads through flats that share the time
 * flats' topology (state.rs:157-165, isochrone_handler.rs:493-499).
 * up_len / down_len hold one weight per arc of `graph`, which must describe
 * the CCH the context was created from (checked arc by arc against the
 * context's layout: BFG_INVALID_PARAMETER otherwise); the two-channel calls
 * use the times of `graph` as given here.  Arcs of INF time are dropped with
 * their length.  Replaces any lengths attached before. */
int bfg_ctx_attach_length(bfg_ctx *ctx, const bfg_graph_desc *graph, const uint32_t *up_len,
                          const uint32_t *down_len);

/* run_phast_bounded_fast_seeded_2ch (isochrone_handler.rs:500-625) for n_lanes
 * independent queries: lane i has seeds (seed_ranks[k], seed_times[k],
 * seed_lens[k]) for k in [seed_offsets[i], seed_offsets[i+1]) (at most 64
 * distinct ranks) and time limit thresholds[i].  times[i * n_nodes + v] is the bounded time of the PHAST
 * above and lens[i * n_nodes + v] the length along it: the least length over
 * the fastest up-down paths, with per-rank seeds taken as the reference takes
 * them (the first of the fastest, :522-526); both BFG_INF where the time
 * exceeds the limit.  Needs bfg_ctx_attach_length.  Lanes run in batches of
 * 64-lane groups sized to free HBM; the environment variable BFG_T2_GROUPS
 * caps the groups per batch (a tuning override, as for the table below). */
int bfg_phast_seeded_2ch(bfg_ctx *ctx, const uint32_t *seed_offsets, const uint32_t *seed_ranks,
                         const uint32_t *seed_times, const uint32_t *seed_lens, uint32_t n_lanes,
                         const uint32_t *thresholds, uint32_t *times, uint32_t *lens);

/* Two-channel seeded S x T table: table_phast_lopsided_2ch
 * (matrix/bucket_ch.rs:6381-6520), the path table_seeded_bounded_routed_2ch
 * (:6329-6379) takes when phast_wins.  Seeds are (rank, part_time, part_len),
 * at most 64 distinct ranks per source.
 * The field of source i is bfg_phast_seeded_2ch of its seeds, bounded by
 * threshold + its largest part_time + the largest target part_time
 * (:6400-6415).  times/lens[i * n_tgt + j]: over target j's seeds in order,
 * the first seed of least field time - part_time (saturating) among the ranks
 * the field settled, with that seed's field length - part_len; BFG_INF in
 * both above threshold (:6454-6481).  Same-edge conflicts (a seed rank shared
 * by source and target) are BFG_INF and flagged in conflicts (nullable), as
 * in bfg_table_seeded.  Needs bfg_ctx_attach_length. */
int bfg_table_seeded_2ch(bfg_ctx *ctx, const uint32_t *src_offsets, const uint32_t *src_ranks,
                         const uint32_t *src_part_times, const uint32_t *src_part_lens, uint32_t n_src,
                         const uint32_t *tgt_offsets, const uint32_t *tgt_ranks, const uint32_t *tgt_part_times,
                         const uint32_t *tgt_part_lens, uint32_t n_tgt, uint32_t threshold, uint32_t *times,
                         uint32_t *lens, uint8_t *conflicts);

/* ---- Customization: triangle relaxation of the CCH ----------------------
 * The topology is the whole CchTopo CSR (route/src/formats/cch_topo.rs:
 * 250-286), INF arcs included; its weights field holds each arc's weight
 * before relaxation (the reference's bottom-up weights: the metric weight of
 * an original arc, the sum over its contraction middle for a shortcut, or
 * BFG_INF for a shortcut not yet weighted).  No context is needed: a call
 * uploads the topology to `device`, relaxes it and downloads the result. */
typedef struct bfg_relax_stats {
    uint64_t triangles;     /* (x -> m, m -> y) pairs evaluated with both halves finite   */
    uint64_t improvements;  /* atomic mins that lowered an arc (depends on the schedule,
                               like the reference's per-pass update counts)                */
    uint32_t levels;        /* dependency levels of apexes (one launch each when non-empty) */
    uint32_t launches;      /* relaxation kernel launches                                   */
    uint32_t passes;        /* 1: one pass in level order reaches the fixpoint             */
    float kernel_ms;        /* device time of the relaxation launches (HIP events)          */
    float total_ms;         /* wall time of the call: checks, levels, copies, kernels       */
} bfg_relax_stats;

/* triangle_relax_parallel (route/src/customization.rs:1083-1200): for every
 * apex m, DOWN arc x -> m and UP arc m -> y (y != x) with both weights finite,
 * the packed word (weight, middle) of the arc x -> y takes the least of
 * itself and (w(x -> m) + w(m -> y) saturating, m), to the fixpoint.
 * up_middle / down_middle (nullable: all u32::MAX, the reference's mark of
 * an original arc) are the middles before relaxation.  Outputs: the relaxed
 * weights, and the middles (nullable).  Equal to the reference's result bit
 * for bit (weights and middles; ties go to the smallest apex).  UP and DOWN
 * targets must ascend per tail (find_edge_index, :1507-1520); an arc x -> y
 * missing from the topology is skipped, as there. */
int bfg_triangle_relax(const bfg_graph_desc *topo, int device, const uint32_t *up_middle,
                       const uint32_t *down_middle, uint32_t *up_weights_out, uint32_t *down_weights_out,
                       uint32_t *up_middle_out, uint32_t *down_middle_out, bfg_relax_stats *stats);

/* triangle_relax_lex_parallel (route/src/customization.rs:1253-1440): the
 * same relaxation over the packed word (time, length), the topology's weights
 * being the times and up_length / down_length the lengths along them; then
 * each shortcut arc (bit i % 64 of word i / 64 of up_is_shortcut /
 * down_is_shortcut set) gets as middle the least apex whose two halves sum to
 * exactly its converged word, or its given middle (nullable: u32::MAX) when
 * none does; other arcs keep their given middle.  Outputs the relaxed times
 * and the middles. */
int bfg_triangle_relax_lex(const bfg_graph_desc *topo, int device, const uint32_t *up_length,
                           const uint32_t *down_length, const uint64_t *up_is_shortcut,
                           const uint64_t *down_is_shortcut, const uint32_t *up_middle, const uint32_t *down_middle,
                           uint32_t *up_times_out, uint32_t *down_times_out, uint32_t *up_middle_out,
                           uint32_t *down_middle_out, bfg_relax_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* BFGPU_H */
