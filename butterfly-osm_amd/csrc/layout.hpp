// layout.hpp — host-side derivation of the device-resident CCH layout.
//
// Input: the rank-aligned CCH the reference engine holds (CchTopo up/down CSR
// + CchWeights decoded to u32, route/src/formats/cch_topo.rs:250-286,
// route/src/formats/cch_weights.rs:171-187).  Output: everything the kernels
// read, derived once per context.
//
// Elimination tree.  The tree of the symmetrised finite-arc pattern (Liu) has
// every arc between an ancestor and a descendant, so (a) the up search space
// of s lies on the root path of s, and (b) nodes that are not ancestor-related
// share no arc and can be relaxed concurrently.
//
// Regions.  Nodes whose subtree holds more than `bottom_limit` nodes form the
// top region T (the separators of the nested dissection and the upper part of
// its leaves); the rest is the bottom region B, a forest of small subtrees.
//   * T is cut into chains (maximal single-child paths) and the chains into
//     segments of at most `segment_limit` (64) nodes.  A segment depends on
//     the segment holding its head's parent; segments of one segment level
//     are independent.  A segment is one register tile: arcs into it from
//     outside are dense rows (source, 64 weights) relaxed in parallel, arcs
//     inside it 8x8 blocks relaxed in order within each thread.
//   * B is swept level-synchronously by subtree height (descending for the
//     downward sweep): nodes of equal height are never ancestor-related.
//
// Positions.  Nodes are renumbered: T first (segment levels ascending, the
// segments of a level in pre-order of their heads, a segment's nodes in chain
// order, parent first), then B by descending height (pre-order within a
// height).  Label arenas are indexed by position.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace bfg {

constexpr uint32_t INF = 0xFFFFFFFFu;
constexpr uint32_t NONE = 0xFFFFFFFFu;

struct CchInput {
    uint32_t n = 0;
    const uint64_t *up_off = nullptr;
    const uint32_t *up_tgt = nullptr, *up_w = nullptr;
    const uint64_t *dn_off = nullptr;
    const uint32_t *dn_tgt = nullptr, *dn_w = nullptr;
};

struct LayoutParams {
    uint32_t bottom_limit = 1024;   // subtree size at or below which a node is bottom
    uint32_t segment_limit = 64;    // max nodes of one chain segment (one register tile, <= 64)
    uint32_t block_limit = 256;     // max nodes of one bottom block (LDS resident)
};

struct Layout {
    uint32_t n = 0;
    uint32_t n_top = 0;             // positions [0, n_top) are the top region
    uint32_t max_depth = 0;         // etree depth of the deepest node (root = 0)
    uint32_t max_height = 0;        // largest bottom-subtree height
    // per rank
    std::vector<uint32_t> parent, depth, pre, subtree;
    std::vector<uint32_t> pos_of_rank, rank_of_pos;
    // per position
    std::vector<uint32_t> pos_parent;          // parent position (NONE at roots)
    std::vector<uint32_t> pos_pre, pos_subtree, pos_depth;
    // DOWN in-arcs by head position.  Top head p: [in_off[p], in_split[p]) come
    // from outside p's segment, [in_split[p], in_off[p+1]) from earlier nodes of
    // the same segment.  Bottom head: see the bottom chunks below.
    std::vector<uint32_t> in_off, in_split, in_src, in_w;
    // UP flat by tail position: arcs [up_off[p], up_split[p]) leave p's segment
    // (all arcs for bottom nodes), [up_split[p], up_off[p+1]) stay inside it
    // (target as offset from the segment start); targets are positions
    std::vector<uint32_t> up_off, up_split, up_tgt, up_w;
    // schedule
    std::vector<uint32_t> seg_pos;        // n_seg+1: segment s covers [seg_pos[s], seg_pos[s+1])
    std::vector<uint32_t> seglvl_seg;     // n_seglvl+1: level l covers segments [seglvl_seg[l], seglvl_seg[l+1])
    std::vector<uint32_t> botlvl_pos;     // n_botlvl+1: bottom level l covers positions [botlvl_pos[l], botlvl_pos[l+1])
    // Bottom chunks: bottom nodes in pre-order cut into chunks of at most
    // block_limit nodes (members of one chunk share their top ancestors).
    // Chunk c holds positions ch_pos[ch_off[c] .. ch_off[c+1]); its DOWN arcs
    // from top-region sources are grouped by source in cs_* / ca_* as above.
    // For a bottom position p, in-arcs [in_off[p], in_split[p]) come from
    // bottom sources, [in_split[p], in_off[p+1]) from top sources.
    std::vector<uint32_t> ch_off, ch_pos, cs_off, cs_src, cs_aoff, ca_w;
    std::vector<uint16_t> ca_v;
    // For a bottom head p, arc a in [in_off[p], in_split[p]) comes from a
    // bottom ancestor in_dd[a - in_off[n_top]] levels above p.
    std::vector<uint16_t> in_dd;
    // Dense tiles of the top region (segments of at most 64 nodes; a thread of
    // the segment kernels holds one lane's labels of a whole segment).  A row
    // is a key (a position) and W weights, the key's arcs to the first W
    // columns of a block (INF where there is none).
    //   DOWN arcs entering segment s from outside: rows [xr_off[s],
    //   xr_off[s+1]), key xr_src[r] (source), ordered by width: row i of the
    //   segment (i = r - xr_off[s]) is in run c (0..7) for xr_run[8s + c - 1]
    //   <= i < xr_run[8s + c] (xr_run[8s - 1] taken as 0) and holds W = 8(8 - c)
    //   weights, the smallest multiple of 8 that covers its last arc; a
    //   segment's rows are stored one after the other from xw_off[s], in key
    //   order within a run.
    //   Intra-segment arcs: 8x8 blocks W[i][j] (tail i of octet a, head j of
    //   octet o) at tb_w[64k ..], k = tb_dn[64s + 8a + o] (a <= o) or
    //   tb_up[64s + 8a + o] (a >= o), NONE where the block has no arc.
    //   UP arcs leaving segment s: its targets (positions) pt_tgt[64k .. 64k+64)
    //   for target blocks k in [pt_blk[s], pt_blk[s+1]) (NONE pads a block);
    //   rows [pr_off[k], pr_off[k+1]) of width pr_wid[k], key pr_x[r] (tail
    //   position), weights pr_w[pw_off[k] + (r - pr_off[k]) * W ..] to the
    //   block's targets.
    std::vector<uint32_t> xr_off, xr_run, xw_off, xr_src, xr_w;
    std::vector<uint32_t> tb_dn, tb_up, tb_w;
    std::vector<uint32_t> pt_blk, pt_tgt, pr_off, pr_wid, pw_off, pr_x, pr_w;
    // pk_seg[k] is the segment of target block k.
    std::vector<uint32_t> pk_seg;
    uint64_t ell_down_entries = 0, ell_up_entries = 0;
    // statistics
    uint64_t m_up = 0, m_down = 0, m_up_inf = 0, m_down_inf = 0;
    uint64_t m_down_intra = 0, m_up_intra = 0;
    uint32_t n_chains = 0, max_segment = 0;
};

// Builds the layout; returns false with a message on invalid input (ranks out
// of range, arcs that do not go up/down, more than 2^32-1 finite arcs).
bool build_layout(const CchInput &in, const LayoutParams &lp, Layout &out, std::string &err);

// Verifies the properties the kernels rely on: every finite arc joins an
// etree ancestor/descendant pair, and the schedule relaxes every arc after its
// tail is final (and, upward, before its head is read).
bool check_layout(const CchInput &in, const Layout &L, std::string &err);

}  // namespace bfg
