// layout.hpp — host-side derivation of the device-resident CCH layout.
//
// Input: the rank-aligned CCH the reference engine holds (CchTopo up/down CSR
// + CchWeights decoded to u32, route/src/formats/cch_topo.rs:250-286,
// route/src/formats/cch_weights.rs:171-187).  Output: everything the kernels
// read, derived once per context:
//   * INF-filtered flats (route/src/matrix/bucket_ch.rs:191-424):
//       up   — push CSR by tail rank (UpAdjFlat), for the upward walk;
//       pull — DOWN arcs keyed by head (DownReverseAdjFlat), re-indexed by
//              sweep position, for the downward pull sweep.
//   * the elimination tree of the symmetrised finite-arc pattern (Liu), so
//     every arc joins an ancestor/descendant pair: the up search space of s
//     is a subset of ancestors(s) and nodes that share no ancestry share no arc.
//   * a task schedule: chains of the top part of the tree and groups of
//     small bottom subtrees; tasks of one level are independent, and inside a
//     task the nodes of one step are independent.
//   * sweep positions: nodes ordered (task level, task, step, rank).
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
    uint32_t bottom_limit = 1024;   // subtree size that becomes a bottom task
    uint32_t group_limit = 4096;    // max nodes packed into one bottom task
};

struct Layout {
    uint32_t n = 0;
    uint32_t max_depth = 0;         // etree depth of the deepest node (root = 0)
    // per rank
    std::vector<uint32_t> parent, depth, pre, subtree;
    std::vector<uint32_t> pos_of_rank, rank_of_pos;
    // up flat by rank (push)
    std::vector<uint32_t> up_off, up_tgt, up_w;
    // pull flat by position (in-arcs of the node at position p)
    std::vector<uint32_t> in_off, in_src, in_w;
    // per position copies the sweep needs
    std::vector<uint32_t> pos_pre, pos_subtree, pos_depth;
    // schedule
    std::vector<uint32_t> step_pos;        // n_steps+1: step s covers positions [step_pos[s], step_pos[s+1])
    std::vector<uint32_t> task_step;       // n_tasks+1: task t covers steps [task_step[t], task_step[t+1])
    std::vector<uint32_t> level_task;      // n_levels+1: level l covers tasks [level_task[l], level_task[l+1])
    // statistics
    uint64_t m_up = 0, m_down = 0, m_up_inf = 0, m_down_inf = 0;
    uint32_t n_top = 0, n_chains = 0, n_bottom_tasks = 0, max_steps_per_task = 0;
};

// Builds the layout; returns false with a message on invalid input (ranks out
// of range, arcs that do not go up/down, more than 2^32-1 finite arcs).
bool build_layout(const CchInput &in, const LayoutParams &lp, Layout &out, std::string &err);

// Verifies the property the sweep relies on: for every finite arc the higher
// endpoint is an etree ancestor of the lower one.  O(m * depth) worst case;
// used by tests and on request.
bool check_ancestry(const CchInput &in, const Layout &L, std::string &err);

}  // namespace bfg
