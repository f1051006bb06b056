/*
 * cchbuild.h — synthetic road network -> customized edge-based CCH (host only).
 *
 * This is the input producer for the K-lane PHAST path: it restates, for a
 * seeded synthetic grid instead of an OSM extract, the reference pipeline's
 * steps that turn a road graph into the rank-aligned CCH the path consumes:
 *   - EBG nodes = directed road segments, arcs = legal turns
 *     (route/src/ebg/mod.rs:360-440, U-turn only at dead ends :380-422)
 *   - node weight = round-half-even(length_m*1000 / speed_mmps), min 1
 *     (route/src/weights.rs:14-40, :365-396); arc weight = w[head] + turn
 *     penalty (route/src/customization.rs:1443-1465)
 *   - car turn penalty sigmoid (route/src/ebg/turn_penalty.rs:234-288,
 *     TurnPenaltyConfig::car)
 *   - inertial nested dissection with greedy vertex-cover separators and
 *     minimum-degree leaves (route/src/ordering.rs:69-176, :480-774)
 *   - directed all-fill-in contraction (route/src/contraction.rs:247-400)
 *   - basic customization to the lower-triangle fixpoint
 *     (route/src/customization.rs:1005-1200)
 * Output arrays are rank-aligned (node id == rank), CSR per tail node, targets
 * ascending, u32 weights with 0xFFFFFFFF = no arc.
 */
#ifndef BFC_CCHBUILD_H
#define BFC_CCHBUILD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bfc_params {
    int32_t rows;            /* grid rows (NBG)                                  */
    int32_t cols;            /* grid columns (NBG)                               */
    uint64_t seed;           /* graph seed                                       */
    double spacing_m;        /* grid spacing in metres                           */
    double jitter_frac;      /* node jitter, fraction of spacing (uniform +-)    */
    double delete_frac;      /* fraction of grid links deleted                   */
    double oneway_frac;      /* fraction of links made one-way                   */
    double chord_frac;       /* long motorway chords per NBG node                */
    double ban_frac;         /* fraction of turns banned (turn restrictions)     */
    int32_t leaf_threshold;  /* nested-dissection leaf size (reference: 8192)    */
    int32_t n_threads;       /* 0 = OpenMP default                               */
    int32_t verbose;         /* progress lines on stderr                         */
} bfc_params;

typedef struct bfc_cch {
    /* rank-aligned CCH */
    uint32_t n;
    uint64_t m_up, m_down;
    uint64_t *up_off;  uint32_t *up_tgt;  uint32_t *up_w;
    uint64_t *dn_off;  uint32_t *dn_tgt;  uint32_t *dn_w;
    /* per-rank geometry (segment midpoint, metres) and identity maps */
    float *x, *y;
    uint32_t *rank_to_filtered;
    /* edge-based graph in rank space: arc u->v has weight w[v] + penalty */
    uint64_t m_ebg;
    uint64_t *ebg_off; uint32_t *ebg_head; uint32_t *ebg_w;
    /* bookkeeping */
    uint32_t nbg_nodes, nbg_edges, n_components;
    uint64_t n_shortcuts;
    double t_graph_s, t_order_s, t_contract_s;
} bfc_cch;

/* Build; returns NULL on failure (message via bfc_last_error). */
bfc_cch *bfc_build_grid(const bfc_params *p);
void bfc_free(bfc_cch *c);
const char *bfc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
