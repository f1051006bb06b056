// meet.hpp — launchers of the many-to-many join kernels (meet.hip).
//
// The bucket many-to-many of the reference (table_bucket_parallel_bounded,
// route/src/matrix/bucket_ch.rs:3436-3520: a forward UP search per source
// filling buckets, a backward search per target over DOWN arcs reversed, and
// a join min(d_fw + d_bw) at every node both searches settle) has, on the
// elimination tree, a fixed meeting set: the common ancestors of the source
// and the target.  Both searches are the upward phase of a PHAST batch, the
// forward one on the CCH and the backward one on its transpose, so their
// labels sit in two label arenas with the same positions; the join is then a
// min-plus product of 64-lane tiles over the top-region segments both tiles
// reach, plus the few pairs whose lowest common ancestor lies in the bottom
// region.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace bfg_meet {

// One side of a join: the upward labels of one batch of lanes (sorted by the
// pre-order of their node) in an arena [lane group][position][kl] of
// `stride` positions per lane, the 64-lane wave groups' reach bits over the
// top-region segments ([wave group][words]), and the original index of each
// sorted lane (a matrix row for sources, a column for targets).
struct Side {
    const uint32_t *lab = nullptr;
    uint32_t stride = 0, kl = 0, n_lanes = 0;
    const unsigned long long *reach = nullptr;
    const uint32_t *index = nullptr;
    // [wave group][segment][8]: the least label of the segment's positions over
    // each 8 consecutive lanes of the wave group (INF where unreached), or null
    // (no pruning)
    const uint32_t *segmin = nullptr;
    // 16-bit copies of the final labels of the reached (segment, wave group)
    // pairs (0xFFFF for INF, else min(label, CAP16); 64 positions x 64 lanes
    // a pair at lab16[4096 slot ..]) and slot[wave group * nseg + segment]
    // (NONE where the pair was not reached), or null (no 16-bit join); 64
    // lanes a group
    const uint16_t *lab16 = nullptr;
    const uint32_t *slot = nullptr;
    // per pair slot and lane: the least 16-bit label of the segment,
    // lmin16[64 slot + lane], or null (blocks pruned by 8-lane minima alone)
    const uint16_t *lmin16 = nullptr;
};

// reach[g * words + s / 64] bit s % 64: wave group g (lanes 64g ..) has a
// seed in the subtree of segment s's head (the segment is on the root path of
// one of its lanes).
hipError_t launch_reach(hipStream_t st, const uint32_t *seed_off, const uint32_t *seed_pre, uint32_t n_lanes,
                        const uint32_t *pos_pre, const uint32_t *pos_sub, const uint32_t *seg_pos, uint32_t nseg,
                        uint32_t words, unsigned long long *reach);

// out[A.index[a] * ld + B.index[b]] = min over the top positions x of the
// segments both lanes' wave groups reach of A.lab(a, x) + B.lab(b, x)
// (saturating; INF where no segment is common), INF where above thr.  With
// both sides' segment minima (nseg segments a wave group), a tile visits the
// common segments deepest first and skips a segment where, for each 8 x 8
// block of its cells, the two lane groups' minima add up to no less than the
// block's largest cell so far: no cell can fall there.
// splits > 1 cuts each tile's common segments over that many wavefronts that
// merge with atomic min: the output cells must hold INF before (launch_fill).
// visits (when not null) gains the label positions the tiles visited.
// only (when not null): tile t = (source wave group) * (target wave groups)
// + (target wave group) is computed only where only[t] != 0.
hipError_t launch_join(hipStream_t st, const Side &A, const Side &B, const uint32_t *seg_pos, uint32_t words,
                       uint32_t nseg, uint32_t thr, uint32_t splits, uint32_t *out, uint64_t ld,
                       unsigned long long *visits, const uint32_t *only = nullptr);

// The join on both sides' 16-bit labels (lab16 / slot set, kl 64, one
// wavefront a tile): packed saturating u16 adds and mins, two cells an
// instruction.  Two finite labels sum to at most 0xFFFE and a sum through INF
// is 0xFFFF, so a cell of 0xFFFF is INF and a cell below CAP16 is exact (a
// sum through a capped label is at least CAP16); a tile with a kept cell in
// [CAP16, 0xFFFE] while thr admits such a value is flagged in redo[t] (zero
// beforehand) and counted in visits[1], and launch_join with only = redo
// recomputes it from the u32 labels.  visits[0] gains the positions visited,
// as launch_join.  With per-lane minima on both sides, a segment that passes
// the segment test is skipped still when no cell exceeds the sum of its two
// lanes' minima.
constexpr uint32_t CAP16 = 0x7FFF;
hipError_t launch_join16(hipStream_t st, const Side &A, const Side &B, const uint32_t *seg_pos, uint32_t words,
                         uint32_t nseg, uint32_t thr, uint32_t *out, uint64_t ld, unsigned long long *visits,
                         uint32_t *redo);

// p[0 .. count) = value (a kernel, so it may sit in a captured graph).
hipError_t launch_fill(hipStream_t st, uint32_t *p, uint64_t count, uint32_t value);

// The pairs (pairs[2k] = source lane, pairs[2k+1] = target lane) of nodes in
// one bottom tree: min over their common bottom ancestors y of the path slots
// (path[j * d1 + depth(seed) - depth(y)] of each lane's seed j) merged into
// the cell with atomic min when within thr.
struct PathSide {
    const uint32_t *seed_off = nullptr, *seed_pos = nullptr, *path = nullptr, *index = nullptr;
};
hipError_t launch_bottom_pairs(hipStream_t st, const uint32_t *pairs, uint32_t n_pairs, const PathSide &A,
                               const PathSide &B, uint32_t d1, const uint32_t *pos_parent,
                               const uint32_t *pos_depth, uint32_t n_top, uint32_t thr, uint32_t *out, uint64_t ld);

}  // namespace bfg_meet
