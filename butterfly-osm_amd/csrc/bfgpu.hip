// bfgpu.hip — MI355X (gfx950) K-lane batched PHAST: kernels + C ABI.
//
// One batch of source lanes (sources are CCH ranks; lanes are sorted by the
// pre-order of their source in the elimination tree so a group of KL lanes
// shares most of its root path).  Labels live in an arena
// [group][position][KL]: one node's KL labels are one coalesced KL*4-byte
// access, so every arc read is reused KL times.
//
// Upward phase — the reference's per-lane heap Dijkstra over UP arcs
// (batched_phast.rs:156-189, isochrone_handler.rs:388-411) as an ancestor
// walk, since the up search space lies on the source's root path:
//   fill_reach   labels := INF for the (segment, 64-lane wave group) pairs the
//                upward phase reaches (a pre-order test on the lanes' seeds;
//                no other pair is ever read as an upward label)
//   bottom_walk  one wavefront per lane walks the bottom part of its root path
//                (depth-indexed slots), relaxing UP arcs into bottom ancestors
//                (slots) and top ancestors (atomic min into the arena)
//   tri<UP>      per (segment, wave group), deepest segment level first: the
//                segment's labels in registers (one lane per thread), its
//                intra-segment UP arcs as 8x8 blocks, last octet to first
//   rows<PUSH>   per (target block, wave group): the UP arcs leaving a
//                segment as rows (tail, W weights) reduced in registers and
//                pushed with one atomic min per (target, lane)
// Downward phase — the rank-descending DOWN scan (batched_phast.rs:205-250,
// :538-590) scheduled by the elimination tree:
//   rows<DOWN>   per (unit of <= 128 rows of a segment, wave group): DOWN arcs
//                from outside the segment as rows (source, W weights), a
//                min-plus product of the sources' labels and the segment's
//                weight block; a unit holding all rows of its segment merges
//                with the upward labels and stores, split units use atomics
//   tri<DOWN>    per (segment, wave group): intra-segment DOWN blocks
//   bottom_path  matrices: the bottom paths above the targets, pulled top-down
//   bottom_chunk per (bottom chunk, group): seeds from the lanes' bottom path
//                slots, then DOWN arcs from top-region sources, grouped by
//                source, into the chunk's LDS labels
//   bottom_pull  per (node, group) of a bottom height level (descending): min
//                over DOWN arcs from bottom-region sources
// Output — gather: target extraction (batched_phast.rs:693-701) through an
// LDS transpose into row-major rows with the per-lane limit mask; it also
// writes full distance rows and isochrone reach bits.
// Arithmetic is u32 with saturating adds (v_add_u32 clamp) and min, so the
// labels equal the unique shortest up-down distances the reference computes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <shared_mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/bfgpu.h"
#include "errors.hpp"
#include "layout.hpp"
#include "meet.hpp"

using bfg::INF;
using bfg::NONE;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(e_ == hipErrorOutOfMemory ? BFG_OUT_OF_MEMORY : BFG_DEVICE_ERROR,   \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                  \
    } while (0)

constexpr int NT = 256;   // threads per workgroup of the gather kernel
constexpr int NTB = 512;  // threads per workgroup of the bottom chunk kernel

__device__ __forceinline__ uint32_t sat_add(uint32_t a, uint32_t b) {
    return __builtin_elementwise_add_sat(a, b);
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}

// ---------------------------------------------------------------------------
// bottom_walk: one wavefront per lane (lanes in sorted order).  Seeds of lane
// l are j in [seed_off[l], seed_off[l+1]) (<= 64), at positions seed_pos[j].
// Seeds in the top region go straight into the arena; every other seed walks
// up its bottom path, the walkers of one lane merging where paths meet (the
// union is visited in ascending rank).  Path slot pathbuf[j*D1 + i] holds the
// label of the node i levels above seed j.
// ---------------------------------------------------------------------------
template <int KL>
__global__ __launch_bounds__(64) void bottom_walk_kernel(
    const uint32_t *__restrict__ seed_off, const uint32_t *__restrict__ seed_rank,
    const uint32_t *__restrict__ seed_pos, const uint32_t *__restrict__ seed_cost,
    const uint32_t *__restrict__ seed_depth, const uint32_t *__restrict__ lane_limit,
    const uint32_t *__restrict__ parent, const uint32_t *__restrict__ pos_parent,
    const uint32_t *__restrict__ pos_depth, const uint32_t *__restrict__ up_off,
    const uint32_t *__restrict__ up_tgt, const uint32_t *__restrict__ up_w, uint32_t n_top,
    uint32_t *__restrict__ pathbuf, uint32_t D1, uint32_t *__restrict__ arena, uint32_t n) {
    const uint32_t l = blockIdx.x;
    const uint32_t t = threadIdx.x;
    const uint32_t g = l / KL, ln = l % KL;
    uint32_t *lab = arena + (size_t)g * n * KL + ln;
    const uint32_t j0 = seed_off[l];
    const uint32_t ns = seed_off[l + 1] - j0;
    const uint32_t lim = lane_limit[l];
    for (uint32_t j = 0; j < ns; ++j) {
        uint32_t *pb = pathbuf + (size_t)(j0 + j) * D1;
        for (uint32_t i = t; i < D1; i += 64) pb[i] = INF;
    }
    __syncthreads();
    uint32_t cur = NONE, cpos = NONE;
    if (t < ns) {
        const uint32_t p = seed_pos[j0 + t], c = seed_cost[j0 + t];
        if (p < n_top) {
            atomicMin(lab + (size_t)p * KL, c);
        } else {
            cur = seed_rank[j0 + t];
            cpos = p;
            pathbuf[(size_t)(j0 + t) * D1] = c;
        }
    }
    __syncthreads();
    for (;;) {
        const uint32_t x = wave_min(cur);
        if (x == NONE) break;
        const bool at = (cur == x);
        const uint64_t mask = __ballot(at);
        const uint32_t first = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
        const uint32_t px = (uint32_t)__shfl((int)cpos, (int)first, 64);
        const uint32_t dxn = pos_depth[px];
        uint32_t mine = INF;
        size_t my_slot = 0;
        if (at) {
            my_slot = (size_t)(j0 + t) * D1 + (seed_depth[j0 + t] - dxn);
            mine = pathbuf[my_slot];
        }
        const uint32_t dx = wave_min(mine);
        __syncthreads();
        if (at) pathbuf[my_slot] = dx;
        if (dx != INF && dx <= lim) {
            const uint32_t js = j0 + first;
            const uint32_t dsj = seed_depth[js];
            uint32_t *pb = pathbuf + (size_t)js * D1;
            const uint32_t a1 = up_off[px + 1];
            for (uint32_t a = up_off[px] + t; a < a1; a += 64) {
                const uint32_t y = up_tgt[a];
                const uint32_t nd = sat_add(dx, up_w[a]);
                if (y < n_top) {
                    atomicMin(lab + (size_t)y * KL, nd);
                } else {
                    uint32_t *slot = pb + (dsj - pos_depth[y]);
                    if (nd < *slot) *slot = nd;
                }
            }
        }
        __syncthreads();
        if (at) {
            const uint32_t pp = pos_parent[cpos];
            if (pp == NONE || pp < n_top) { cur = NONE; cpos = NONE; }
            else { cur = parent[x]; cpos = pp; }
        }
    }
}

// A slot is the KL threads that carry one source group's KL lanes.  With
// KL = 64 a slot is exactly one wavefront, so everything indexed by the slot
// (sources, arcs, weight rows) is wave-uniform: readfirstlane moves the slot
// index to a scalar register and the arc reads become scalar loads that every
// lane sees for free.
template <int KL>
__device__ __forceinline__ uint32_t slot_id() {
    const uint32_t s = threadIdx.x / KL;
    if constexpr (KL >= 64) return (uint32_t)__builtin_amdgcn_readfirstlane((int)s);
    return s;
}

// Relaxes the arcs [a0, a1) of one source (label d of this lane) into the
// LDS labels.  KL >= 64: a0, a1 are wave-uniform and the arc data comes in
// through scalar loads.  KL < 64 (two slots per wavefront): the slot's lanes
// load up to KL arcs with one coalesced access and broadcast them by shuffle.
template <int KL>
__device__ __forceinline__ void relax_arcs(uint32_t *L, uint32_t d, uint32_t a0, uint32_t a1,
                                           const uint16_t *__restrict__ av, const uint32_t *__restrict__ aw,
                                           uint32_t lane) {
    if constexpr (KL >= 64) {
        for (uint32_t base = a0; base < a1; base += 64) {
            const uint32_t cnt = min(64u, a1 - base);
            const uint32_t wl = __lane_id();
            uint32_t mv = 0, mw = INF;
            if (wl < cnt) { mv = av[base + wl]; mw = aw[base + wl]; }
#pragma unroll 4
            for (uint32_t j = 0; j < cnt; ++j) {
                const uint32_t v = (uint32_t)__builtin_amdgcn_readlane((int)mv, (int)j);
                const uint32_t w = (uint32_t)__builtin_amdgcn_readlane((int)mw, (int)j);
                if (d != INF) atomicMin(&L[v * KL + lane], sat_add(d, w));
            }
        }
    } else {
        for (uint32_t base = a0; base < a1; base += KL) {
            const uint32_t cnt = min((uint32_t)KL, a1 - base);
            uint32_t mv = 0, mw = INF;
            if (lane < cnt) { mv = av[base + lane]; mw = aw[base + lane]; }
            for (uint32_t j = 0; j < cnt; ++j) {
                const uint32_t v = (uint32_t)__shfl((int)mv, (int)j, KL);
                const uint32_t w = (uint32_t)__shfl((int)mw, (int)j, KL);
                if (d != INF) atomicMin(&L[v * KL + lane], sat_add(d, w));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Wave tiles of the top region.  A wavefront carries 64 source lanes (64/KL
// groups of KL; wave group w holds groups w*64/KL ..) and each thread holds,
// for its lane, the labels of one block of up to 64 positions in registers.
// A row (key, 64 weights) is wave-uniform: its key and weights come in
// through scalar loads, each thread reads the key's label of its lane (one
// coalesced load per row) and does 64 saturating adds and mins.
// ---------------------------------------------------------------------------
constexpr int TILE_THREADS = 256;  // four wavefronts: four wave groups on one tile
constexpr uint32_t MAX_GRID_Y = 65535;  // most workgroups along y in one launch
constexpr uint32_t XCDS = 8;            // MI355X: accelerator complex dies, each with its own L2

// Read-only tables the kernels index with wave-uniform offsets: read through
// the constant address space, whose uniform loads become scalar loads.
typedef __attribute__((address_space(4))) const uint32_t cu32;
__device__ __forceinline__ const cu32 *as_const(const uint32_t *p) { return (const cu32 *)p; }

// Index of this thread's wavefront in the workgroup, as a scalar (the
// compiler cannot prove threadIdx.x / 64 uniform by itself).
__device__ __forceinline__ uint32_t wave_id() {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
}

template <int KL>
__device__ __forceinline__ uint32_t *lane_labels(uint32_t *arena, uint32_t n, uint32_t wg, uint32_t *group) {
    const uint32_t l = threadIdx.x % 64;
    const uint32_t g = wg * (64 / KL) + l / KL;
    if (group) *group = g;
    return arena + (size_t)g * n * KL + l % KL;
}

// Lane j's value of v, as a scalar (j wave-uniform).
__device__ __forceinline__ uint32_t rl(uint32_t v, uint32_t j) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)j);
}

// Which wave groups reach which segments.  A lane's upward labels are finite
// exactly on the root paths of its seeds, so wave group wg (lanes wg*64 ..,
// sorted order) reaches the segment headed by position p iff a seed of one of
// its lanes lies in p's subtree: a pre-order interval test, no labels read.
struct Reach {
    const uint32_t *seed_off, *seed_pre, *pos_pre, *pos_sub;
    uint32_t n_lanes;
};

__device__ __forceinline__ bool wave_reaches(const Reach &R, uint32_t wg, uint32_t p) {
    const uint32_t gl = wg * 64 + threadIdx.x % 64;
    const uint32_t pr = R.pos_pre[p], ps = R.pos_sub[p];
    bool r = false;
    if (gl < R.n_lanes)
        for (uint32_t j = R.seed_off[gl]; j < R.seed_off[gl + 1]; ++j) r |= R.seed_pre[j] - pr < ps;
    return __ballot(r) != 0;
}

// fill_reach: wavefront w of workgroup (wave group blockIdx.x, blockIdx.y)
// takes segment 4*blockIdx.y + w and, if its wave group reaches it, sets the
// segment's labels of the wave group to INF before the upward phase pushes
// into them.  Unreached segments are never read as such (see tri_kernel and
// rows_kernel).
template <int KL>
__global__ __launch_bounds__(TILE_THREADS) void fill_reach_kernel(uint32_t nseg, const uint32_t *__restrict__ seg_pos,
                                                                  Reach R, uint32_t *__restrict__ arena,
                                                                  uint32_t n) {
    const uint32_t s = blockIdx.y * (TILE_THREADS / 64) + wave_id();
    if (s >= nseg) return;  // wave-uniform; this kernel has no barrier
    const uint32_t p0 = seg_pos[s], len = seg_pos[s + 1] - p0;
    if (!wave_reaches(R, blockIdx.x, p0)) return;
    uint32_t *lab = lane_labels<KL>(arena, n, blockIdx.x, nullptr);
    for (uint32_t j = 0; j < len; ++j) lab[(size_t)(p0 + j) * KL] = INF;
}

// acc[j] = min(acc[j], lab[key[r]] + wr[8k(r - r0) + j]) over rows [r0, r1)
// of k octets (8k weights, the first of the tile's N columns; k is
// wave-uniform, an octet past it is skipped by a scalar branch); two rows
// per step (min(acc, x0, x1) is one v_min3), the labels of the next two rows
// loaded one step ahead.  The keys of 64 rows come in with one vector load
// (lane j: row c0 + j) and reach the label addresses by v_readlane, so a
// step waits on no scalar load of its own besides the weights.  Reads past a
// chunk's last row re-read that row (label and weights), so an odd last row
// pairs with itself: min(acc, x, x) is exact.
template <int KL, int N>
__device__ __forceinline__ void rows_w(uint32_t (&acc)[N], uint32_t r0, uint32_t r1, const uint32_t *key,
                                       const cu32 *wr, uint32_t k, const uint32_t *lab) {
    const uint32_t wl = __lane_id(), W = 8 * k;
    for (uint32_t c0 = r0; c0 < r1; c0 += 64) {
        const uint32_t cl = min(64u, r1 - c0) - 1;  // last row of the chunk, relative
        const uint32_t kv = key[c0 + min(wl, cl)];
        uint32_t n0 = lab[(size_t)rl(kv, 0) * KL];
        uint32_t n1 = lab[(size_t)rl(kv, min(1u, cl)) * KL];
        const cu32 *W0 = wr + (size_t)(c0 - r0) * W;
        for (uint32_t j = 0; j <= cl; j += 2) {
            const uint32_t a0 = n0, a1 = n1;
            const cu32 *W1 = j < cl ? W0 + W : W0;
            n0 = lab[(size_t)rl(kv, min(j + 2, cl)) * KL];
            n1 = lab[(size_t)rl(kv, min(j + 3, cl)) * KL];
#pragma unroll
            for (int o = 0; o < N / 8; ++o) {
                if (o > 0 && (uint32_t)o >= k) break;
#pragma unroll
                for (int q = 8 * o; q < 8 * o + 8; ++q)
                    acc[q] = min(acc[q], min(sat_add(a0, W0[q]), sat_add(a1, W1[q])));
            }
            // the next labels are waited for here, after this step's adds:
            // the loads stay one step ahead of their use
            asm volatile("" : "+v"(n0), "+v"(n1));
            W0 += 2 * W;
        }
    }
}

// Intra-segment arcs of a segment of at most 8*NO nodes in registers
// (acc[j] = label of node j for this lane): 8x8 blocks T[8a + o] of tw, DOWN
// from the first octet to the last (the diagonal block of an octet first, i
// ascending), UP from the last to the first (i descending).  The block ids
// come in with one vector load (lane i holds T[i]) and every block's 64
// weights with one more each (lane j holds weight j), all issued before the
// first is used, so a wavefront waits on its loads once; a weight reaches
// the adds as a scalar through v_readlane.  Blocks without arcs load from
// tw[0 .. 63] (the table always holds one block) and are skipped.
template <bool DOWN, int NO>
__device__ __forceinline__ void tri_apply(uint32_t (&acc)[8 * NO], const uint32_t *T, const uint32_t *tw) {
    constexpr int NB = NO * (NO + 1) / 2;
    const uint32_t lane = __lane_id();
    const uint32_t kv = T[lane];
    uint32_t wv[NB];
    int b = 0;
#pragma unroll
    for (int aa = 0; aa < NO; ++aa) {
        const int a = DOWN ? aa : NO - 1 - aa;
#pragma unroll
        for (int oo = 0; oo < NO; ++oo) {
            const int o = DOWN ? a + oo : a - oo;
            if (o < 0 || o >= NO) continue;
            const uint32_t k = (uint32_t)__builtin_amdgcn_readlane((int)kv, a * 8 + o);
            wv[b++] = tw[(size_t)(k == NONE ? 0u : k) * 64 + lane];
        }
    }
    // every load above is complete here, and none is moved below
#pragma unroll
    for (int i = 0; i < NB; ++i) asm volatile("" ::"v"(wv[i]));
    b = 0;
#pragma unroll
    for (int aa = 0; aa < NO; ++aa) {
        const int a = DOWN ? aa : NO - 1 - aa;
#pragma unroll
        for (int oo = 0; oo < NO; ++oo) {
            const int o = DOWN ? a + oo : a - oo;
            if (o < 0 || o >= NO) continue;
            const uint32_t k = (uint32_t)__builtin_amdgcn_readlane((int)kv, a * 8 + o);
            const uint32_t wb = wv[b++];
            if (k == NONE) continue;
#pragma unroll
            for (int ii = 0; ii < 8; ++ii) {
                const int i = DOWN ? ii : 7 - ii;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    if (o == a && (DOWN ? j <= i : j >= i)) continue;
                    const uint32_t w = (uint32_t)__builtin_amdgcn_readlane((int)wb, i * 8 + j);
                    acc[o * 8 + j] = min(acc[o * 8 + j], sat_add(acc[a * 8 + i], w));
                }
            }
        }
    }
}

// One unit of rows_kernel at width W: rows [ra, rb) with weights from wr,
// merged into positions out[j] (PUSH) or p0 + j (j < len) of the arena with
// one return-less atomic min per finite (position, lane).  A DOWN unit that
// holds all rows of its segment (whole) and whose wave group does not reach
// the segment stores instead: those upward labels are INF and never written.
// tri_apply for levels of many segments, where enough wavefronts are in
// flight to hide load latency and VALU issue sets the time: the block ids
// come in with one vector load, each block's 64 weights through scalar
// loads (shared in the scalar cache by the workgroup's four wavefronts) and
// reach the adds as scalar operands, with no v_readlane.  Scheduling
// barriers keep half a block's weights in SGPRs at a time.
template <bool DOWN, int NO>
__device__ __forceinline__ void tri_apply_s(uint32_t (&acc)[8 * NO], const uint32_t *T, const cu32 *tw) {
    const uint32_t kv = T[__lane_id()];
#pragma unroll
    for (int aa = 0; aa < NO; ++aa) {
        const int a = DOWN ? aa : NO - 1 - aa;
#pragma unroll
        for (int oo = 0; oo < NO; ++oo) {
            const int o = DOWN ? a + oo : a - oo;
            if (o < 0 || o >= NO) continue;
            const uint32_t k = rl(kv, a * 8 + o);
            if (k == NONE) continue;
            const cu32 *W = tw + (size_t)k * 64;
#pragma unroll
            for (int ii = 0; ii < 8; ++ii) {
                const int i = DOWN ? ii : 7 - ii;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    if (o == a && (DOWN ? j <= i : j >= i)) continue;
                    acc[o * 8 + j] = min(acc[o * 8 + j], sat_add(acc[a * 8 + i], W[i * 8 + j]));
                }
                if (ii % 4 == 3) __builtin_amdgcn_sched_barrier(0);  // half a block's weights at a time
            }
        }
    }
}

template <int KL, int W>
__device__ __forceinline__ void push_unit(uint32_t ra, uint32_t rb, const uint32_t *K, const cu32 *wr, const cu32 *out,
                                          uint32_t *lab) {
    uint32_t acc[W];
#pragma unroll
    for (int j = 0; j < W; ++j) acc[j] = INF;
    rows_w<KL, W>(acc, ra, rb, K, wr, W / 8, lab);
#pragma unroll
    for (int j = 0; j < W; ++j)
        if (out[j] != NONE && acc[j] != INF) atomicMin(lab + (size_t)out[j] * KL, acc[j]);
}

// A DOWN unit of width W: rows [ra, rb) of one run of a segment (8W/8
// weights each, from wr), reduced into acc[W] and merged into the segment's
// labels p0 .. p0 + len: stored when the unit holds all rows of its segment
// (whole) and its wave group does not reach the segment (the upward labels
// are INF and were never written; columns past W get INF), else merged by
// one atomic min per finite (node, lane).
template <int KL, int W>
__device__ __forceinline__ void down_unit(uint32_t ra, uint32_t rb, const uint32_t *K, const cu32 *wr, uint32_t p0,
                                          uint32_t len, uint32_t *lab, bool whole, bool reach) {
    uint32_t acc[W];
#pragma unroll
    for (int j = 0; j < W; ++j) acc[j] = INF;
    rows_w<KL, W>(acc, ra, rb, K, wr, W / 8, lab);
    if (whole && !reach) {
#pragma unroll
        for (int j = 0; j < 64; ++j)
            if (j < len) lab[(size_t)(p0 + j) * KL] = j < W ? acc[j] : INF;
        return;
    }
#pragma unroll
    for (int j = 0; j < W; ++j)
        if (j < len && acc[j] != INF) atomicMin(lab + (size_t)(p0 + j) * KL, acc[j]);
}

// rows: wavefront w of workgroup (blockIdx.x, blockIdx.y) takes unit
// u0 + blockIdx.y for wave group 4*blockIdx.x + w < wgs: the four wavefronts
// of a workgroup read the same weights, so three of them find the rows in
// the scalar cache.
//   DOWN (PUSH = false): unit = rows [r0[u], r1[u]) of one run c of
//     segment seg[u] & UNIT_SEG_MASK (DOWN arcs from outside the segment;
//     c = seg[u] >> UNIT_RUN_SHIFT & 7, 8(8 - c) weights a row from woff[u]
//     on), reduced by the down_unit of that width and merged into the
//     segment's labels (which hold the upward labels): by plain stores if the
//     unit is flagged UNIT_WHOLE (it holds all rows of its segment), else by
//     one atomic min per finite (node, lane).  The intra arcs follow in
//     tri_kernel.
//   UP (PUSH = true): unit = target block u, rows [r0[u], r0[u+1]) keyed by
//     tails in segment seg[u] (final up labels), width wid[u], weights from
//     woff[u], merged into targets tgt[64u ..] with atomic min; a wavefront
//     whose wave group does not reach the segment stops at once.
constexpr uint32_t UNIT_WHOLE = 0x80000000u;
constexpr uint32_t UNIT_RUN_SHIFT = 28;  // DOWN: the unit's run c in bits 28..30
constexpr uint32_t UNIT_SEG_MASK = (1u << UNIT_RUN_SHIFT) - 1;

// The block gate of bounded calls (the K-lane form of the reference's
// active-block skipping, batched_phast.rs:1040-1064, isochrone_handler.rs:
// 414-450).  Outside the upward search space a label is a min over DOWN arcs
// from ancestors, so a DOWN row (source u, weights to a segment) can lower a
// label to d <= limit only if label(u) + (the row's least weight) <= limit.
// A unit none of whose rows passes that bound for any lane of its wave group
// is skipped (its labels stay the upward ones, or INF); a segment none of
// whose labels is within its lane's limit skips its intra-segment arcs.
// Skipped values are >= the true labels and > the limits, so every label
// within its limit stays exact and every other one is masked on output.
struct Gate {
    const uint32_t *rmin;        // least weight of each DOWN row; nullptr: no gate
    const uint32_t *lane_limit;  // per sorted lane
    uint32_t n_lanes;
    uint8_t *seg_live;           // [wave group][segment]: a label within a limit (statistics)
    uint32_t nseg;
};

// Whether some row of [r0, r1) passes the bound for some lane of this
// wavefront (wave-uniform).  Keys and least weights of 64 rows come in with
// one vector load each; eight labels are loaded before any is compared.
template <int KL>
__device__ __forceinline__ bool rows_live(const Gate &G, uint32_t wg, uint32_t r0, uint32_t r1,
                                          const uint32_t *__restrict__ key, const uint32_t *lab) {
    const uint32_t wl = __lane_id(), gl = wg * 64 + wl;
    const uint32_t lim = gl < G.n_lanes ? G.lane_limit[gl] : 0u;
    const bool valid = gl < G.n_lanes;
    for (uint32_t c0 = r0; c0 < r1; c0 += 64) {
        const uint32_t cl = min(64u, r1 - c0) - 1;
        const uint32_t kv = key[c0 + min(wl, cl)], mv = G.rmin[c0 + min(wl, cl)];
        bool live = false;
        for (uint32_t j = 0; j <= cl; j += 8) {
            uint32_t x[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) x[q] = lab[(size_t)rl(kv, min(j + q, cl)) * KL];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const uint32_t d = sat_add(x[q], rl(mv, min(j + q, cl)));
                live |= (x[q] != INF) & (d <= lim);
            }
        }
        if (__ballot(live && valid) != 0) return true;
    }
    return false;
}

template <int KL, bool PUSH, bool GATE>
__global__ __launch_bounds__(TILE_THREADS) void rows_kernel(
    uint32_t u0, uint32_t u1, const uint32_t *__restrict__ seg, const uint32_t *__restrict__ r0,
    const uint32_t *__restrict__ r1, const uint32_t *__restrict__ wid,
    const uint32_t *__restrict__ woff, const uint32_t *__restrict__ seg_pos, const uint32_t *__restrict__ tgt,
    const uint32_t *__restrict__ key, const uint32_t *__restrict__ wts, Reach R, Gate G,
    uint32_t *__restrict__ arena, uint32_t n, uint32_t wgs) {
    const uint32_t u = u0 + blockIdx.y, wg = blockIdx.x * (TILE_THREADS / 64) + wave_id();
    if (u >= u1 || wg >= wgs) return;  // wave-uniform; this kernel has no barrier
    uint32_t *lab = lane_labels<KL>(arena, n, wg, nullptr);
    const uint32_t sw = seg[u], s = sw & UNIT_SEG_MASK;
    const bool reach = wave_reaches(R, wg, seg_pos[s]);
    if (PUSH && !reach) return;  // the segment's labels are all INF: nothing to push
    if (GATE && !rows_live<KL>(G, wg, r0[u], r1[u], key, lab)) {
        // no row of the unit reaches a limit: a whole unit of an unreached
        // wave group stores INF (its upward labels were never written); the
        // upward labels (or the INF of split segments) stand otherwise
        if ((sw & UNIT_WHOLE) && !reach) {
            const uint32_t p0 = seg_pos[s], len = seg_pos[s + 1] - p0;
            for (uint32_t j = 0; j < len; ++j) lab[(size_t)(p0 + j) * KL] = INF;
        }
        return;
    }
    if constexpr (PUSH) {
        const uint32_t ra = r0[u], rb = r0[u + 1], W = wid[u];
        const cu32 *wr = as_const(wts) + woff[u], *out = as_const(tgt) + (size_t)u * 64;
        if (W == 16) push_unit<KL, 16>(ra, rb, key, wr, out, lab);
        else if (W == 32) push_unit<KL, 32>(ra, rb, key, wr, out, lab);
        else push_unit<KL, 64>(ra, rb, key, wr, out, lab);
    } else {
        const uint32_t ra = r0[u], rb = r1[u], p0 = seg_pos[s], len = seg_pos[s + 1] - p0;
        const cu32 *wr = as_const(wts) + woff[u];
        const bool whole = (sw & UNIT_WHOLE) != 0;
        switch ((sw >> UNIT_RUN_SHIFT) & 7u) {  // the unit's run: width 8(8 - c)
        case 0: down_unit<KL, 64>(ra, rb, key, wr, p0, len, lab, whole, reach); break;
        case 1: down_unit<KL, 56>(ra, rb, key, wr, p0, len, lab, whole, reach); break;
        case 2: down_unit<KL, 48>(ra, rb, key, wr, p0, len, lab, whole, reach); break;
        case 3: down_unit<KL, 40>(ra, rb, key, wr, p0, len, lab, whole, reach); break;
        case 4: down_unit<KL, 32>(ra, rb, key, wr, p0, len, lab, whole, reach); break;
        case 5: down_unit<KL, 24>(ra, rb, key, wr, p0, len, lab, whole, reach); break;
        case 6: down_unit<KL, 16>(ra, rb, key, wr, p0, len, lab, whole, reach); break;
        default: down_unit<KL, 8>(ra, rb, key, wr, p0, len, lab, whole, reach); break;
        }
    }
}

// ---------------------------------------------------------------------------
// Packed 16-bit DOWN rows.  One wavefront carries two wave groups, A and B
// (128 lanes): thread l holds lane l of A in the low and lane l of B in the
// high half of each accumulator, and v_pk_add_u16 (clamp) + v_pk_min_u16 relax
// both with one instruction each, a weight broadcast from a scalar half-word
// (weights packed two per dword, min(w, 0xFFFF)).  Arithmetic saturating at
// 0xFFFF equals the u32 arithmetic as long as every finite candidate stays
// below 0xFFFF: a finite source label x of a row r must satisfy x <= 0xFFFE -
// (largest finite weight of r), i.e. x + 1 <= cap[r] with cap[r] = 0xFFFF -
// that weight (0 when it is 0xFFFF or more).  Then 0xFFFF means INF and every
// other value is the exact u32 result.  A unit with a label past its row's
// cap (any lane of A or B) is redone with the u32 rows (down_unit), so the
// results are bit-identical to the u32 kernel either way.
// ---------------------------------------------------------------------------
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ u16x2 as_u16x2(uint32_t x) { return __builtin_bit_cast(u16x2, x); }

// min(x, 0xFFFF) of two u32 labels into one packed word (A low, B high)
__device__ __forceinline__ uint32_t pack16(uint32_t a, uint32_t b) {
    return min(a, 0xFFFFu) | (min(b, 0xFFFFu) << 16);
}

// acc[j] = min(acc[j], {labA, labB}[key[r]] + w16[..]) over rows [r0, r1) of
// k octets, two rows per step, labels one step ahead (as rows_w).  Returns
// (per lane) whether a finite label passed its row's cap.
// A pointer that is the same in every lane, as a scalar (the compiler cannot
// prove it uniform by itself).
typedef __attribute__((address_space(1))) const uint32_t gu32;  // global memory

__device__ __forceinline__ const gu32 *uniform_ptr(const uint32_t *p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return (const gu32 *)((uint64_t)hi << 32 | lo);
}

// Label loads of one row for wave groups A and B.  KL = 64: a row's labels
// of a wave group are one 256-byte line at a wave-uniform address, so the
// address is a scalar base (wave group base + key * 256) and the lane's
// constant offset, with no per-row vector address arithmetic.
template <int KL>
struct RowLab {
    const gu32 *a, *b;  // KL = 64: the wave groups' label bases (wave-uniform); else the lane's pointers
    uint32_t lane;
    __device__ __forceinline__ uint32_t A(uint32_t key) const {
        if constexpr (KL == 64) return (a + (size_t)key * 64)[lane];
        return a[(size_t)key * KL];
    }
    __device__ __forceinline__ uint32_t B(uint32_t key) const {
        if constexpr (KL == 64) return (b + (size_t)key * 64)[lane];
        return b[(size_t)key * KL];
    }
};

template <int KL, int N>
__device__ __forceinline__ bool rows_w16(u16x2 (&acc)[N], uint32_t r0, uint32_t r1, const uint32_t *key,
                                         const uint32_t *cap, const cu32 *wr, uint32_t k, const RowLab<KL> &LB) {
    const uint32_t wl = __lane_id(), W2 = 4 * k;  // dwords per row
    bool bad = false;
    for (uint32_t c0 = r0; c0 < r1; c0 += 64) {
        const uint32_t cl = min(64u, r1 - c0) - 1;
        const uint32_t kv = key[c0 + min(wl, cl)], cv = cap[c0 + min(wl, cl)];
        uint32_t nA0 = LB.A(rl(kv, 0)), nB0 = LB.B(rl(kv, 0));
        uint32_t nA1 = LB.A(rl(kv, min(1u, cl))), nB1 = LB.B(rl(kv, min(1u, cl)));
        const cu32 *W0 = wr + (size_t)(c0 - r0) * W2;
        for (uint32_t j = 0; j <= cl; j += 2) {
            const uint32_t c0v = rl(cv, j), c1v = rl(cv, min(j + 1, cl));
            bad |= (nA0 + 1u > c0v) | (nB0 + 1u > c0v) | (nA1 + 1u > c1v) | (nB1 + 1u > c1v);
            const u16x2 a0 = as_u16x2(pack16(nA0, nB0)), a1 = as_u16x2(pack16(nA1, nB1));
            const cu32 *W1 = j < cl ? W0 + W2 : W0;
            const uint32_t k2 = rl(kv, min(j + 2, cl)), k3 = rl(kv, min(j + 3, cl));
            nA0 = LB.A(k2);
            nB0 = LB.B(k2);
            nA1 = LB.A(k3);
            nB1 = LB.B(k3);
#pragma unroll
            for (int o = 0; o < N / 8; ++o) {
                if (o > 0 && (uint32_t)o >= k) break;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const u16x2 w = as_u16x2(W0[o * 4 + q]), v = as_u16x2(W1[o * 4 + q]);
                    u16x2 &x0 = acc[o * 8 + 2 * q], &x1 = acc[o * 8 + 2 * q + 1];
                    // four independent adds, then the mins (a packed op reading the
                    // result of the one just before it waits a cycle)
                    const u16x2 t0 = __builtin_elementwise_add_sat(a0, w.xx);
                    const u16x2 t1 = __builtin_elementwise_add_sat(a1, v.xx);
                    const u16x2 t2 = __builtin_elementwise_add_sat(a0, w.yy);
                    const u16x2 t3 = __builtin_elementwise_add_sat(a1, v.yy);
                    x0 = __builtin_elementwise_min(x0, __builtin_elementwise_min(t0, t1));
                    x1 = __builtin_elementwise_min(x1, __builtin_elementwise_min(t2, t3));
                }
            }
            asm volatile("" : "+v"(nA0), "+v"(nB0), "+v"(nA1), "+v"(nB1));
            W0 += 2 * W2;
        }
    }
    return bad;
}

// The u32 fallback of a 16-bit unit, compact rather than fast (it runs only
// where a label does not fit 16 bits): column j of the unit is the min over
// its rows of label + weight j, the rows' labels re-read per column.  Merges
// as down_unit.
template <int KL>
__device__ __noinline__ void down_unit_u32(uint32_t ra, uint32_t rb, const uint32_t *K, const cu32 *wr, uint32_t W,
                                           uint32_t p0, uint32_t len, uint32_t *lab, bool whole, bool reach) {
    for (uint32_t j = 0; j < len; ++j) {
        uint32_t m = INF;
        if (j < W)
            for (uint32_t r = ra; r < rb; ++r) m = min(m, sat_add(lab[(size_t)K[r] * KL], wr[(size_t)(r - ra) * W + j]));
        if (whole && !reach) lab[(size_t)(p0 + j) * KL] = m;
        else if (m != INF) atomicMin(lab + (size_t)(p0 + j) * KL, m);
    }
}

// A DOWN unit of width W for wave groups A and B (B absent when !hasB: its
// pointer then repeats A's and nothing is written to it).  Merges as
// down_unit, per wave group: stored when the unit is whole and the wave
// group does not reach the segment, else one atomic min per finite value.
template <int KL, int W>
__device__ __forceinline__ void down_unit16(uint32_t ra, uint32_t rb, const uint32_t *K, const uint32_t *cap,
                                            const cu32 *wr16, const cu32 *wr32, uint32_t p0, uint32_t len,
                                            uint32_t *labA, uint32_t *labB, bool hasB, bool whole, bool reachA,
                                            bool reachB) {
    u16x2 acc[W];
#pragma unroll
    for (int j = 0; j < W; ++j) acc[j] = (u16x2){0xFFFF, 0xFFFF};
    RowLab<KL> LB{(const gu32 *)labA, (const gu32 *)labB, 0};
    if constexpr (KL == 64) {  // wave-uniform bases (lane_labels added the lane)
        LB.lane = __lane_id();
        LB.a = uniform_ptr(labA - LB.lane);
        LB.b = uniform_ptr(labB - LB.lane);
    }
    const bool bad = rows_w16<KL, W>(acc, ra, rb, K, cap, wr16, W / 8, LB);
    if (__ballot(bad) != 0) {  // a label too large for 16 bits: u32 arithmetic, per wave group
        down_unit_u32<KL>(ra, rb, K, wr32, W, p0, len, labA, whole, reachA);
        if (hasB) down_unit_u32<KL>(ra, rb, K, wr32, W, p0, len, labB, whole, reachB);
        return;
    }
    const bool stA = whole && !reachA, stB = hasB && whole && !reachB;
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const uint32_t vA = acc[j].x == 0xFFFF ? INF : (uint32_t)acc[j].x;
        const uint32_t vB = acc[j].y == 0xFFFF ? INF : (uint32_t)acc[j].y;
        if (j < len) {
            if (stA) labA[(size_t)(p0 + j) * KL] = vA;
            else if (vA != INF) atomicMin(labA + (size_t)(p0 + j) * KL, vA);
            if (stB) labB[(size_t)(p0 + j) * KL] = vB;
            else if (hasB && vB != INF) atomicMin(labB + (size_t)(p0 + j) * KL, vB);
        }
    }
    // columns past W of a stored unit: INF
    for (uint32_t j = W; j < len; ++j) {
        if (stA) labA[(size_t)(p0 + j) * KL] = INF;
        if (stB) labB[(size_t)(p0 + j) * KL] = INF;
    }
}

// rows16: wavefront w of workgroup (blockIdx.x, blockIdx.y) takes DOWN unit
// u0 + blockIdx.y for the wave groups 2p and 2p + 1, p = 4 blockIdx.x + w <
// ceil(wgs / 2).  As rows_kernel<KL, false, GATE> otherwise; the weights come
// from the packed u16 copy, the u32 ones serve the fallback.
template <int KL, bool GATE>
__global__ __launch_bounds__(TILE_THREADS) __attribute__((amdgpu_waves_per_eu(5))) void rows16_kernel(
    uint32_t u0, uint32_t u1, const uint32_t *__restrict__ seg, const uint32_t *__restrict__ r0,
    const uint32_t *__restrict__ r1, const uint32_t *__restrict__ woff, const uint32_t *__restrict__ seg_pos,
    const uint32_t *__restrict__ key, const uint32_t *__restrict__ cap, const uint32_t *__restrict__ wts16,
    const uint32_t *__restrict__ wts, Reach R, Gate G, uint32_t *__restrict__ arena, uint32_t n, uint32_t wgs) {
    const uint32_t u = u0 + blockIdx.y, pr = blockIdx.x * (TILE_THREADS / 64) + wave_id();
    const uint32_t wgA = 2 * pr, wgB = 2 * pr + 1;
    if (u >= u1 || wgA >= wgs) return;  // wave-uniform; this kernel has no barrier
    const bool hasB = wgB < wgs;
    uint32_t *labA = lane_labels<KL>(arena, n, wgA, nullptr);
    uint32_t *labB = hasB ? lane_labels<KL>(arena, n, wgB, nullptr) : labA;
    const uint32_t sw = seg[u], s = sw & UNIT_SEG_MASK;
    const uint32_t p0 = seg_pos[s], len = seg_pos[s + 1] - p0;
    const bool reachA = wave_reaches(R, wgA, p0), reachB = hasB && wave_reaches(R, wgB, p0);
    const bool whole = (sw & UNIT_WHOLE) != 0;
    if constexpr (GATE) {
        const bool liveA = rows_live<KL>(G, wgA, r0[u], r1[u], key, labA);
        const bool liveB = hasB && rows_live<KL>(G, wgB, r0[u], r1[u], key, labB);
        if (!liveA && !liveB) {  // as rows_kernel: INF where a whole unit's wave group is unreached
            if (whole)
                for (uint32_t j = 0; j < len; ++j) {
                    if (!reachA) labA[(size_t)(p0 + j) * KL] = INF;
                    if (hasB && !reachB) labB[(size_t)(p0 + j) * KL] = INF;
                }
            return;
        }
    }
    const uint32_t ra = r0[u], rb = r1[u];
    const cu32 *wr32 = as_const(wts) + woff[u], *wr16 = as_const(wts16) + woff[u] / 2;
    switch ((sw >> UNIT_RUN_SHIFT) & 7u) {  // the unit's run: width 8(8 - c)
    case 0: down_unit16<KL, 64>(ra, rb, key, cap, wr16, wr32, p0, len, labA, labB, hasB, whole, reachA, reachB); break;
    case 1: down_unit16<KL, 56>(ra, rb, key, cap, wr16, wr32, p0, len, labA, labB, hasB, whole, reachA, reachB); break;
    case 2: down_unit16<KL, 48>(ra, rb, key, cap, wr16, wr32, p0, len, labA, labB, hasB, whole, reachA, reachB); break;
    case 3: down_unit16<KL, 40>(ra, rb, key, cap, wr16, wr32, p0, len, labA, labB, hasB, whole, reachA, reachB); break;
    case 4: down_unit16<KL, 32>(ra, rb, key, cap, wr16, wr32, p0, len, labA, labB, hasB, whole, reachA, reachB); break;
    case 5: down_unit16<KL, 24>(ra, rb, key, cap, wr16, wr32, p0, len, labA, labB, hasB, whole, reachA, reachB); break;
    case 6: down_unit16<KL, 16>(ra, rb, key, cap, wr16, wr32, p0, len, labA, labB, hasB, whole, reachA, reachB); break;
    default: down_unit16<KL, 8>(ra, rb, key, cap, wr16, wr32, p0, len, labA, labB, hasB, whole, reachA, reachB); break;
    }
}

// tri: wavefront w of workgroup (blockIdx.x, blockIdx.y) takes segment
// first + blockIdx.y < first + count of one level for wave group
// 4*blockIdx.x + w < wgs (the workgroup shares the segment's blocks), loads its
// labels, relaxes its intra-segment arcs (tri_apply) and stores them.  UP: a
// wave group that does not reach the segment skips it, only writing INF
// where the DOWN plan merges into the labels with atomics (split[s]).  SC:
// scalar-loaded blocks (tri_apply_s) for levels of many segments.
template <int KL, bool DOWN, bool SC, bool GATE>
__global__ __launch_bounds__(TILE_THREADS) void tri_kernel(uint32_t first, uint32_t count,
                                                           const uint32_t *__restrict__ seg_pos,
                                                           const uint32_t *__restrict__ tab,
                                                           const uint32_t *__restrict__ tw,
                                                           const uint8_t *__restrict__ split, Reach R, Gate G,
                                                           uint32_t *__restrict__ arena, uint32_t n,
                                                           uint32_t wgs) {
    const uint32_t i = blockIdx.y, wg = blockIdx.x * (TILE_THREADS / 64) + wave_id();
    if (i >= count || wg >= wgs) return;  // wave-uniform; this kernel has no barrier
    const uint32_t s = first + i;
    uint32_t *lab = lane_labels<KL>(arena, n, wg, nullptr);
    const uint32_t p0 = seg_pos[s], len = seg_pos[s + 1] - p0;
    if constexpr (!DOWN) {
        if (!wave_reaches(R, wg, p0)) {
            if (split[s])
                for (uint32_t j = 0; j < len; ++j) lab[(size_t)(p0 + j) * KL] = INF;
            return;
        }
    }
    uint32_t acc[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) acc[j] = j < len ? lab[(size_t)(p0 + j) * KL] : INF;
    if constexpr (DOWN && GATE) {
        {  // intra arcs only raise labels from above the limits: skip them then
            const uint32_t gl = wg * 64 + __lane_id();
            const uint32_t lim = gl < G.n_lanes ? G.lane_limit[gl] : 0u;
            bool live = false;
#pragma unroll
            for (int j = 0; j < 64; ++j) live |= (j < len) & (acc[j] != INF) & (acc[j] <= lim);
            const bool alive = __ballot(live && gl < G.n_lanes) != 0;
            if (__lane_id() == 0) G.seg_live[(size_t)wg * G.nseg + s] = alive ? 1 : 0;
            if (!alive) return;
        }
    }
    if constexpr (SC) tri_apply_s<DOWN, 8>(acc, tab + (size_t)s * 64, as_const(tw));
    else tri_apply<DOWN, 8>(acc, tab + (size_t)s * 64, tw);
#pragma unroll
    for (int j = 0; j < 64; ++j)
        if (j < len) lab[(size_t)(p0 + j) * KL] = acc[j];
}

// Upward phase over lists of reached units (the bucket's searches, where
// nothing reads an unreached (segment, wave group) pair): item i = 4 *
// blockIdx.x + w of a list (unit u[i], wave group g[i]) goes to wavefront w;
// lists are sorted by unit, so a workgroup's wavefronts mostly share theirs.
//   fill_list: labels of reached segment u := INF
//   tri_list:  the intra-segment UP blocks of segment u (tri_kernel's UP work)
//   push_list: the UP push rows of target block u (rows_kernel's PUSH work)
template <int KL>
__global__ __launch_bounds__(TILE_THREADS) void fill_list_kernel(uint32_t count, const uint32_t *__restrict__ lu,
                                                                 const uint32_t *__restrict__ lg,
                                                                 const uint32_t *__restrict__ seg_pos,
                                                                 uint32_t *__restrict__ arena, uint32_t n) {
    const uint32_t i = blockIdx.x * (TILE_THREADS / 64) + wave_id();
    if (i >= count) return;  // wave-uniform; this kernel has no barrier
    const uint32_t s = lu[i];
    uint32_t *lab = lane_labels<KL>(arena, n, lg[i], nullptr);
    const uint32_t p0 = seg_pos[s], len = seg_pos[s + 1] - p0;
    for (uint32_t j = 0; j < len; ++j) lab[(size_t)(p0 + j) * KL] = INF;
}

template <int KL>
__global__ __launch_bounds__(TILE_THREADS) void tri_list_kernel(uint32_t count, const uint32_t *__restrict__ lu,
                                                                const uint32_t *__restrict__ lg,
                                                                const uint32_t *__restrict__ seg_pos,
                                                                const uint32_t *__restrict__ tab,
                                                                const uint32_t *__restrict__ tw,
                                                                uint32_t *__restrict__ arena, uint32_t n,
                                                                uint32_t *__restrict__ segmin, uint32_t nseg,
                                                                uint16_t *__restrict__ lab16,
                                                                uint32_t *__restrict__ slot,
                                                                uint16_t *__restrict__ lmin16, uint32_t item0) {
    // a chain's latency is the level's critical path while the other side's
    // pushes fill the SIMDs: its wavefronts issue first
    __builtin_amdgcn_s_setprio(3);
    const uint32_t i = blockIdx.x * (TILE_THREADS / 64) + wave_id();
    if (i >= count) return;
    const uint32_t s = lu[i];
    uint32_t *lab = lane_labels<KL>(arena, n, lg[i], nullptr);
    const uint32_t p0 = seg_pos[s], len = seg_pos[s + 1] - p0;
    uint32_t acc[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) acc[j] = j < len ? lab[(size_t)(p0 + j) * KL] : INF;
    tri_apply<false, 8>(acc, tab + (size_t)s * 64, tw);
#pragma unroll
    for (int j = 0; j < 64; ++j)
        if (j < len) lab[(size_t)(p0 + j) * KL] = acc[j];
    // the segment's labels are final now: their least over each 8 lanes of the
    // wave group bounds every meeting of those lanes through the segment (the
    // join's pruning)
    uint32_t m = INF;
#pragma unroll
    for (int j = 0; j < 64; ++j)
        if (j < len) m = min(m, acc[j]);
    if constexpr (KL == 64) {  // the lane's least label as the join's 16-bit bound
        if (lab16) lmin16[(size_t)(item0 + i) * 64 + __lane_id()] = (uint16_t)(m != INF ? min(m, bfg_meet::CAP16) : 0xFFFFu);
    }
#pragma unroll
    for (int o = 4; o > 0; o >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, o, 64));
    const uint32_t lane = __lane_id();
    if (lane % 8 == 0) segmin[((size_t)lg[i] * nseg + s) * 8 + lane / 8] = m;
    if constexpr (KL == 64) {  // the 16-bit copy the join reads (positions past len: INF)
        if (lab16) {
            uint16_t *q = lab16 + (size_t)(item0 + i) * 4096 + lane;
#pragma unroll
            for (int j = 0; j < 64; ++j)
                q[j * 64] = (uint16_t)(j < len && acc[j] != INF ? min(acc[j], bfg_meet::CAP16) : 0xFFFFu);
            if (lane == 0) slot[(size_t)lg[i] * nseg + s] = item0 + i;
        }
    }
}

// Upward pushes on column pairs.  A label enters as 16 bits, 0xFFFF for INF
// and min(label, PUSH_CAP) else, and a weight the same way (the packed copy
// holds columns 2q, 2q + 1 of a row in dword q).  Two finite values sum to at
// most 0xFFFE and a sum through INF saturates to 0xFFFF, so a column of
// 0xFFFF is INF and a column below PUSH_CAP is the u32 result (a sum through
// a capped value is at least PUSH_CAP).
constexpr uint32_t PUSH_CAP = 0x7FFF;

__device__ __forceinline__ uint32_t lab16_of(uint32_t x) {
    return x >= PUSH_CAP ? (x == INF ? 0xFFFFu : PUSH_CAP) : x;
}

// rows_w with acc[q] = columns 2q (low) and 2q + 1 (high): per two rows and
// column pair, two v_pk_add_u16 (clamp; the row's label broadcast by
// op_sel) and two v_pk_min_u16, against three u32 instructions per two
// updates.  Rows, keys and label prefetch as rows_w; wr16 holds 4k dwords a
// row.
template <int KL, int N>
__device__ __forceinline__ void rows_w16c(u16x2 (&acc)[N / 2], uint32_t r0, uint32_t r1, const uint32_t *key,
                                          const cu32 *wr16, uint32_t k, const uint32_t *lab) {
    const uint32_t wl = __lane_id(), W2 = 4 * k;
    for (uint32_t c0 = r0; c0 < r1; c0 += 64) {
        const uint32_t cl = min(64u, r1 - c0) - 1;
        const uint32_t kv = key[c0 + min(wl, cl)];
        uint32_t n0 = lab[(size_t)rl(kv, 0) * KL];
        uint32_t n1 = lab[(size_t)rl(kv, min(1u, cl)) * KL];
        const cu32 *W0 = wr16 + (size_t)(c0 - r0) * W2;
        for (uint32_t j = 0; j <= cl; j += 2) {
            const u16x2 a0 = as_u16x2(lab16_of(n0)), a1 = as_u16x2(lab16_of(n1));
            const cu32 *W1 = j < cl ? W0 + W2 : W0;
            n0 = lab[(size_t)rl(kv, min(j + 2, cl)) * KL];
            n1 = lab[(size_t)rl(kv, min(j + 3, cl)) * KL];
#pragma unroll
            for (int o = 0; o < N / 8; ++o) {
                if (o > 0 && (uint32_t)o >= k) break;
#pragma unroll
                for (int q = 4 * o; q < 4 * o + 4; ++q) {
                    const u16x2 s0 = __builtin_elementwise_add_sat(a0.xx, as_u16x2(W0[q]));
                    const u16x2 s1 = __builtin_elementwise_add_sat(a1.xx, as_u16x2(W1[q]));
                    acc[q] = __builtin_elementwise_min(acc[q], __builtin_elementwise_min(s0, s1));
                }
            }
            asm volatile("" : "+v"(n0), "+v"(n1));
            W0 += 2 * W2;
        }
    }
}

// The u32 push of a unit whose 16-bit columns passed the cap, compact rather
// than fast (labels re-read per column): merges as push_unit.
template <int KL>
__device__ __noinline__ void push_unit_u32(uint32_t ra, uint32_t rb, const uint32_t *K, const cu32 *wr, uint32_t W,
                                           const cu32 *out, uint32_t *lab) {
    for (uint32_t j = 0; j < W; ++j) {
        const uint32_t t = out[j];
        if (t == NONE) continue;
        uint32_t m = INF;
        for (uint32_t r = ra; r < rb; ++r) m = min(m, sat_add(lab[(size_t)K[r] * KL], wr[(size_t)(r - ra) * W + j]));
        if (m != INF) atomicMin(lab + (size_t)t * KL, m);
    }
}

// push_unit on column pairs; a unit with a column in [PUSH_CAP, 0xFFFE] on
// any lane is pushed again in u32 (the exact columns already merged are
// true sums, so merging them first changes nothing).
template <int KL, int W>
__device__ __forceinline__ void push_unit16(uint32_t ra, uint32_t rb, const uint32_t *K, const cu32 *wr16,
                                            const cu32 *wr, const cu32 *out, uint32_t *lab) {
    u16x2 acc[W / 2];
#pragma unroll
    for (int q = 0; q < W / 2; ++q) acc[q] = (u16x2){0xFFFF, 0xFFFF};
    rows_w16c<KL, W>(acc, ra, rb, K, wr16, W / 8, lab);
    bool wide = false;
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const uint32_t v = (j & 1) ? acc[j / 2].y : acc[j / 2].x;
        wide |= v >= PUSH_CAP && v != 0xFFFFu;
        if (out[j] != NONE && v < PUSH_CAP) atomicMin(lab + (size_t)out[j] * KL, v);
    }
    if (__ballot(wide) != 0) push_unit_u32<KL>(ra, rb, K, wr, W, out, lab);
}

// wts16 (when not null): the packed 16-bit copy of wts, a row's columns in
// pairs from woff[u] / 2 on
template <int KL>
__global__ __launch_bounds__(TILE_THREADS) void push_list_kernel(
    uint32_t count, const uint32_t *__restrict__ lu, const uint32_t *__restrict__ lg, const uint32_t *__restrict__ r0,
    const uint32_t *__restrict__ wid, const uint32_t *__restrict__ woff, const uint32_t *__restrict__ tgt,
    const uint32_t *__restrict__ key, const uint32_t *__restrict__ wts, const uint32_t *__restrict__ wts16,
    uint32_t *__restrict__ arena, uint32_t n) {
    const uint32_t i = blockIdx.x * (TILE_THREADS / 64) + wave_id();
    if (i >= count) return;
    const uint32_t u = lu[i];
    uint32_t *lab = lane_labels<KL>(arena, n, lg[i], nullptr);
    const uint32_t ra = r0[u], rb = r0[u + 1], W = wid[u];
    const cu32 *wr = as_const(wts) + woff[u], *out = as_const(tgt) + (size_t)u * 64;
    if (wts16) {
        const cu32 *wr16 = as_const(wts16) + woff[u] / 2;
        if (W == 16) push_unit16<KL, 16>(ra, rb, key, wr16, wr, out, lab);
        else if (W == 32) push_unit16<KL, 32>(ra, rb, key, wr16, wr, out, lab);
        else push_unit16<KL, 64>(ra, rb, key, wr16, wr, out, lab);
        return;
    }
    if (W == 16) push_unit<KL, 16>(ra, rb, key, wr, out, lab);
    else if (W == 32) push_unit<KL, 32>(ra, rb, key, wr, out, lab);
    else push_unit<KL, 64>(ra, rb, key, wr, out, lab);
}

// The UP pushes of the PHAST upward phase (rows_kernel<KL, true, false>) on
// 16-bit column pairs: workgroup (blockIdx.x, blockIdx.y) takes push block
// u0 + blockIdx.y for the wave groups 4 blockIdx.x + w that reach its
// segment; wts16 as for push_list_kernel.
template <int KL>
__global__ __launch_bounds__(TILE_THREADS) void rows_push16_kernel(
    uint32_t u0, uint32_t u1, const uint32_t *__restrict__ seg, const uint32_t *__restrict__ r0,
    const uint32_t *__restrict__ wid, const uint32_t *__restrict__ woff, const uint32_t *__restrict__ seg_pos,
    const uint32_t *__restrict__ tgt, const uint32_t *__restrict__ key, const uint32_t *__restrict__ wts,
    const uint32_t *__restrict__ wts16, Reach R, uint32_t *__restrict__ arena, uint32_t n, uint32_t wgs) {
    const uint32_t u = u0 + blockIdx.y, wg = blockIdx.x * (TILE_THREADS / 64) + wave_id();
    if (u >= u1 || wg >= wgs) return;  // wave-uniform; this kernel has no barrier
    const uint32_t s = seg[u] & UNIT_SEG_MASK;
    if (!wave_reaches(R, wg, seg_pos[s])) return;  // the segment's labels are all INF: nothing to push
    uint32_t *lab = lane_labels<KL>(arena, n, wg, nullptr);
    const uint32_t ra = r0[u], rb = r0[u + 1], W = wid[u];
    const cu32 *wr = as_const(wts) + woff[u], *wr16 = as_const(wts16) + woff[u] / 2;
    const cu32 *out = as_const(tgt) + (size_t)u * 64;
    if (W == 16) push_unit16<KL, 16>(ra, rb, key, wr16, wr, out, lab);
    else if (W == 32) push_unit16<KL, 32>(ra, rb, key, wr16, wr, out, lab);
    else push_unit16<KL, 64>(ra, rb, key, wr16, wr, out, lab);
}

// bottom_gate: per (bottom tree, lane group), whether the tree can hold a
// label within a lane's limit: some bottom chunk holding part of the tree
// was left open by the chunk gate (chunk_ok).  Thread (blockIdx.x * NT +
// threadIdx.x) takes tree t of group blockIdx.y.
__global__ __launch_bounds__(NT) void bottom_gate_kernel(uint32_t n_bt, const uint32_t *__restrict__ bt_ch,
                                                         const uint8_t *__restrict__ chunk_ok, uint32_t nch,
                                                         uint8_t *__restrict__ open_b) {
    const uint32_t t = blockIdx.x * NT + threadIdx.x, g = blockIdx.y;
    if (t >= n_bt) return;
    uint8_t o = 0;
    for (uint32_t c = bt_ch[2 * t]; c <= bt_ch[2 * t + 1]; ++c) o |= chunk_ok[(size_t)g * nch + c];
    open_b[(size_t)g * n_bt + t] = o;
}

// The bottom gate as the bottom kernels read it: open_b = nullptr means no
// gate; open_b[g * n_bt + t] for lane group g and bottom tree t.
struct BGate {
    const uint8_t *open_b;
    const uint32_t *pos_bt;  // tree of bottom position p at pos_bt[p - n_top]
    uint32_t n_bt, n_top;
};

template <int KL>
__device__ __forceinline__ bool bottom_open(const BGate &B, uint32_t g, uint32_t p) {
    return B.open_b[(size_t)g * B.n_bt + B.pos_bt[p - B.n_top]] != 0;
}

// The chunk gate: per (bottom chunk, lane group), whether a lane can settle a
// node of the chunk within its limit: a seed's path slot within the limit,
// or a top source u of the chunk with label(u) + (least weight of u's arcs
// into the chunk) within it.  A closed chunk's labels are INF.
struct CGate {
    const uint32_t *cs_min;      // least weight of each chunk source's arcs; nullptr: no gate
    const uint32_t *lane_limit;  // per sorted lane
    uint8_t *chunk_ok;           // [group][chunk]
    uint32_t nch;
};

// Relaxes the grouped arcs of sources [k0, k1) into the LDS labels: sources
// go in batches of UNR per slot — ids, arc ranges and labels of a batch are
// loaded before any is relaxed, so UNR label reads are in flight per slot.
template <int KL, int NS>
__device__ __forceinline__ void relax_sources(uint32_t *L, const uint32_t *__restrict__ lab, uint32_t k0, uint32_t k1,
                                              const uint32_t *__restrict__ src, const uint32_t *__restrict__ aoff,
                                              const uint16_t *__restrict__ av, const uint32_t *__restrict__ aw,
                                              uint32_t slot, uint32_t lane) {
    constexpr int UNR = 8;
    if constexpr (KL >= 64) {
        // wavefront `slot` takes sources k0 + slot + NS*m (round robin); the
        // ids and arc ranges of 64 of them arrive in one vector load and are
        // read lane by lane
        const uint32_t wl = __lane_id();
        const uint32_t mine = k1 > k0 + slot ? (k1 - k0 - slot + NS - 1) / NS : 0u;
        for (uint32_t mb = 0; mb < mine; mb += 64) {
            const uint32_t cntk = min(64u, mine - mb);
            uint32_t ms = 0, ma0 = 0, ma1 = 0;
            if (wl < cntk) {
                const uint32_t k = k0 + slot + (mb + wl) * NS;
                ms = src[k];
                ma0 = aoff[k];
                ma1 = aoff[k + 1];
            }
            for (uint32_t j0 = 0; j0 < cntk; j0 += UNR) {
                uint32_t d[UNR];
#pragma unroll
                for (int r = 0; r < UNR; ++r) {
                    const uint32_t u = (uint32_t)__builtin_amdgcn_readlane((int)ms, (int)(j0 + r));
                    d[r] = (j0 + r < cntk) ? lab[(size_t)u * KL + lane] : INF;
                }
#pragma unroll
                for (int r = 0; r < UNR; ++r) {
                    if (j0 + r >= cntk) break;
                    const uint32_t a0 = (uint32_t)__builtin_amdgcn_readlane((int)ma0, (int)(j0 + r));
                    const uint32_t a1 = (uint32_t)__builtin_amdgcn_readlane((int)ma1, (int)(j0 + r));
                    relax_arcs<KL>(L, d[r], a0, a1, av, aw, lane);
                }
            }
        }
    } else {
        // sources in batches of UNR per slot: ids, arc ranges and labels of a
        // batch are loaded before any is relaxed
        for (uint32_t kb = k0 + slot; kb < k1; kb += UNR * NS) {
            uint32_t u[UNR], a0[UNR], a1[UNR], d[UNR];
#pragma unroll
            for (int r = 0; r < UNR; ++r) {
                const uint32_t k = kb + r * NS;
                const bool ok = k < k1;
                u[r] = ok ? src[k] : 0;
                a0[r] = ok ? aoff[k] : 0;
                a1[r] = ok ? aoff[k + 1] : 0;
            }
#pragma unroll
            for (int r = 0; r < UNR; ++r) d[r] = a1[r] > a0[r] ? lab[(size_t)u[r] * KL + lane] : INF;
#pragma unroll
            for (int r = 0; r < UNR; ++r) relax_arcs<KL>(L, d[r], a0[r], a1[r], av, aw, lane);
        }
    }
}

// ---------------------------------------------------------------------------
// bottom_chunk: all bottom chunks in one launch, group blockIdx.y.  A chunk
// holds bottom nodes of neighbouring subtrees (pre-order), which share their
// top ancestors: labels in LDS seeded from the lane's bottom path slots, then
// the DOWN arcs from top sources grouped by source.  Results go to the arena;
// arcs between bottom nodes follow in bottom_pull.
// ---------------------------------------------------------------------------
template <int KL, int BB, int NTH>
__global__ __launch_bounds__(NTH) void bottom_chunk_kernel(
    const uint32_t *__restrict__ ch_off, const uint32_t *__restrict__ ch_pos, const uint32_t *__restrict__ cs_off,
    const uint32_t *__restrict__ cs_src, const uint32_t *__restrict__ cs_aoff, const uint16_t *__restrict__ ca_v,
    const uint32_t *__restrict__ ca_w, const uint32_t *__restrict__ pos_pre, const uint32_t *__restrict__ pos_sub,
    const uint32_t *__restrict__ pos_depth, const uint32_t *__restrict__ seed_off,
    const uint32_t *__restrict__ seed_pre, const uint32_t *__restrict__ seed_depth,
    const uint32_t *__restrict__ pathbuf, uint32_t D1, uint32_t n_lanes, CGate CG, uint32_t *__restrict__ arena,
    uint32_t n) {
    constexpr int NS = NTH / KL;
    __shared__ uint32_t L[BB * KL];
    const uint32_t c = blockIdx.x, g = blockIdx.y;
    const uint32_t lane = threadIdx.x % KL, slot = slot_id<KL>();
    const uint32_t i0 = ch_off[c], len = ch_off[c + 1] - i0;
    uint32_t *lab = arena + (size_t)g * n * KL;
    uint32_t sj0 = 0, sj1 = 0;
    const uint32_t gl = g * KL + lane;
    if (gl < n_lanes) { sj0 = seed_off[gl]; sj1 = seed_off[gl + 1]; }
    for (uint32_t i = slot; i < len; i += NS) {
        uint32_t v = INF;
        if (sj0 < sj1) {
            const uint32_t p = ch_pos[i0 + i];
            const uint32_t pr = pos_pre[p], ps = pos_sub[p], pd = pos_depth[p];
            for (uint32_t j = sj0; j < sj1; ++j) {
                const uint32_t sp = seed_pre[j];
                if (sp >= pr && sp - pr < ps) v = min(v, pathbuf[(size_t)j * D1 + (seed_depth[j] - pd)]);
            }
        }
        L[i * KL + lane] = v;
    }
    if (CG.cs_min) {  // the chunk gate (workgroup-uniform outcome)
        const uint32_t lim = gl < n_lanes ? CG.lane_limit[gl] : 0u;
        bool live = false;
        if (gl < n_lanes) {
            for (uint32_t i = slot; i < len; i += NS) {
                const uint32_t v = L[i * KL + lane];
                live |= (v != INF) & (v <= lim);
            }
            for (uint32_t k = cs_off[c] + slot; k < cs_off[c + 1]; k += NS) {
                const uint32_t x = lab[(size_t)cs_src[k] * KL + lane];
                live |= (x != INF) & (sat_add(x, CG.cs_min[k]) <= lim);
            }
        }
        const bool open = __syncthreads_or(live) != 0;
        if (threadIdx.x == 0) CG.chunk_ok[(size_t)g * CG.nch + c] = open ? 1 : 0;
        if (!open) {
            for (uint32_t i = slot; i < len; i += NS) lab[(size_t)ch_pos[i0 + i] * KL + lane] = INF;
            return;
        }
    } else {
        __syncthreads();
    }
    relax_sources<KL, NS>(L, lab, cs_off[c], cs_off[c + 1], cs_src, cs_aoff, ca_v, ca_w, slot, lane);
    __syncthreads();
    for (uint32_t i = slot; i < len; i += NS) lab[(size_t)ch_pos[i0 + i] * KL + lane] = L[i * KL + lane];
}

// bottom_path for KL = 64, where a slot is one wavefront and everything
// indexed by the path is wave-uniform.  A wavefront walks its path's
// elements in chunks of 64: lane j gathers element j's position, arc range
// and pre-order interval once, and each element reads them back by
// v_readlane.  An element's arcs come in as vectors (lane j: arc a0 + j, its
// label index and weight), those of the next element while the current
// element's labels are in flight, so the walk waits once per element: on
// its labels, the first 32 arcs' in one batch (the rest, rare, in further
// batches).  Lanes past an element's arcs read label 0 with weight INF.
__device__ __forceinline__ void path_arcs(uint32_t wl, uint32_t li, uint32_t a, uint32_t as, uint32_t a1, uint32_t ib,
                                          const uint16_t *__restrict__ in_dd, const uint32_t *__restrict__ in_src,
                                          const uint32_t *__restrict__ in_w, uint32_t &mi, uint32_t &mw) {
    mi = 0;
    mw = INF;
    a += wl;
    if (a < a1) {
        mi = a < as ? li - in_dd[a - ib] : in_src[a];
        mw = in_w[a];
    }
}

__device__ void bottom_path_wave(uint32_t e0, uint32_t e1, uint32_t n_top, const uint32_t *__restrict__ path_pos,
                                 const uint32_t *__restrict__ in_off, const uint32_t *__restrict__ in_split,
                                 const uint16_t *__restrict__ in_dd, const uint32_t *__restrict__ in_src,
                                 const uint32_t *__restrict__ in_w, const uint32_t *__restrict__ pos_pre,
                                 const uint32_t *__restrict__ pos_sub, const uint32_t *__restrict__ pos_depth,
                                 uint32_t sj0, uint32_t sj1, const uint32_t *__restrict__ seed_pre,
                                 const uint32_t *__restrict__ seed_depth, const uint32_t *__restrict__ pathbuf,
                                 uint32_t D1, uint32_t *__restrict__ lab) {
    const uint32_t wl = __lane_id();
    const uint32_t ib = in_off[n_top];  // first arc of a bottom head: in_dd's origin
    // the lane's first seed, kept in registers; further seeds (rare) are read per element
    const uint32_t sp0 = sj0 < sj1 ? seed_pre[sj0] : 0u, sd0 = sj0 < sj1 ? seed_depth[sj0] : 0u;
    for (uint32_t c0 = e0; c0 < e1; c0 += 64) {
        const uint32_t cn = min(64u, e1 - c0);
        uint32_t vp = 0, va0 = 0, vas = 0, va1 = 0, vpr = 0, vps = 0, vpd = 0;
        if (wl < cn) {
            vp = path_pos[c0 + wl];
            va0 = in_off[vp];
            vas = in_split[vp];
            va1 = in_off[vp + 1];
            vpr = pos_pre[vp];
            vps = pos_sub[vp];
            vpd = pos_depth[vp];
        }
        uint32_t mi, mw;
        path_arcs(wl, n_top + c0, rl(va0, 0), rl(vas, 0), rl(va1, 0), ib, in_dd, in_src, in_w, mi, mw);
        for (uint32_t j = 0; j < cn; ++j) {
            const uint32_t li = n_top + c0 + j;  // this element's label index
            const uint32_t a0 = rl(va0, j), a1 = rl(va1, j);
            uint32_t x[32];
#pragma unroll
            for (int r = 0; r < 32; ++r) x[r] = lab[(size_t)rl(mi, r) * 64];
            const uint32_t cw = mw;
            // the next element's first arcs
            if (j + 1 < cn)
                path_arcs(wl, li + 1, rl(va0, j + 1), rl(vas, j + 1), rl(va1, j + 1), ib, in_dd, in_src, in_w, mi,
                          mw);
            uint32_t m0 = INF, m1 = INF, m2 = INF, m3 = INF;
            if (sj0 < sj1) {
                const uint32_t pr = rl(vpr, j), ps = rl(vps, j), pd = rl(vpd, j);
                if (sp0 - pr < ps) m0 = pathbuf[(size_t)sj0 * D1 + (sd0 - pd)];
                for (uint32_t q = sj0 + 1; q < sj1; ++q) {
                    const uint32_t sp = seed_pre[q];
                    if (sp - pr < ps) m0 = min(m0, pathbuf[(size_t)q * D1 + (seed_depth[q] - pd)]);
                }
            }
#pragma unroll
            for (int r = 0; r < 32; r += 4) {
                m0 = min(m0, sat_add(x[r], rl(cw, r)));
                m1 = min(m1, sat_add(x[r + 1], rl(cw, r + 1)));
                m2 = min(m2, sat_add(x[r + 2], rl(cw, r + 2)));
                m3 = min(m3, sat_add(x[r + 3], rl(cw, r + 3)));
            }
            // arcs past the first 32 of this element
            for (uint32_t base = a0 + 32; base < a1; base += 32) {
                const uint32_t as = rl(vas, j);
                uint32_t bi, bw;
                path_arcs(wl & 31u, li, base, as, a1, ib, in_dd, in_src, in_w, bi, bw);
#pragma unroll
                for (int r = 0; r < 32; ++r) x[r] = lab[(size_t)rl(bi, r) * 64];
#pragma unroll
                for (int r = 0; r < 32; r += 4) {
                    m0 = min(m0, sat_add(x[r], rl(bw, r)));
                    m1 = min(m1, sat_add(x[r + 1], rl(bw, r + 1)));
                    m2 = min(m2, sat_add(x[r + 2], rl(bw, r + 2)));
                    m3 = min(m3, sat_add(x[r + 3], rl(bw, r + 3)));
                }
            }
            lab[(size_t)li * 64] = min(min(m0, m1), min(m2, m3));
        }
    }
}

// bottom_path: the bottom region restricted to what a matrix reads.  A
// target's label depends only on its ancestors, so the bottom sweep of a
// matrix needs the bottom paths from the top region down to its targets
// alone.  Slot (KL threads, one per lane) `blockIdx.x * NS + slot` of group
// blockIdx.y walks path k, path_pos[path_off[k] .. path_off[k+1]) top-down:
// each node starts from the lane's upward label (its bottom path slots) and
// pulls over all its DOWN in-arcs.  Labels are compact: top positions as
// they are, path element i at n_top + i.  Top sources are final; a bottom
// source lies above the node on the same path (at the element its depth
// gives) and was written by this same thread.
template <int KL>
__global__ __launch_bounds__(NT) void bottom_path_kernel(
    uint32_t n_paths, const uint32_t *__restrict__ path_off, const uint32_t *__restrict__ path_pos, uint32_t n_top,
    const uint32_t *__restrict__ in_off, const uint32_t *__restrict__ in_split, const uint16_t *__restrict__ in_dd,
    const uint32_t *__restrict__ in_src, const uint32_t *__restrict__ in_w,
    const uint32_t *__restrict__ pos_pre, const uint32_t *__restrict__ pos_sub, const uint32_t *__restrict__ pos_depth,
    const uint32_t *__restrict__ seed_off, const uint32_t *__restrict__ seed_pre,
    const uint32_t *__restrict__ seed_depth, const uint32_t *__restrict__ pathbuf, uint32_t D1, uint32_t n_lanes,
    uint32_t *__restrict__ arena, uint32_t n) {
    constexpr int NS = NT / KL;
    const uint32_t lane = threadIdx.x % KL, slot = slot_id<KL>();
    // Workgroups go to the XCDs round robin (XCD b % XCDS).  XCD x walks the
    // groups g = x (mod XCDS), GPX of them side by side on each path chunk,
    // chunk after chunk: one XCD's L2 then holds the top labels of its own
    // groups only, for the few neighbouring paths in flight.  Blocks past the
    // last group (the padding of GPX) return at once.
    const uint32_t groups = (n_lanes + KL - 1) / KL, gpx = (groups + XCDS - 1) / XCDS;
    const uint32_t x = blockIdx.x % XCDS, j = blockIdx.x / XCDS;
    const uint32_t g = x + XCDS * (j % gpx);
    const uint32_t k = j / gpx * NS + slot;
    if (g >= groups || k >= n_paths) return;  // this kernel has no barrier
    uint32_t *lab = arena + (size_t)g * n * KL + lane;
    uint32_t sj0 = 0, sj1 = 0;
    const uint32_t gl = g * KL + lane;
    if (gl < n_lanes) {
        sj0 = seed_off[gl];
        sj1 = seed_off[gl + 1];
    }
    if constexpr (KL == 64) {
        bottom_path_wave(path_off[k], path_off[k + 1], n_top, path_pos, in_off, in_split, in_dd, in_src, in_w, pos_pre,
                         pos_sub, pos_depth, sj0, sj1, seed_pre, seed_depth, pathbuf, D1, lab);
        return;
    }
    const uint32_t ib = in_off[n_top];  // first arc of a bottom head: in_dd's origin
    for (uint32_t i = path_off[k]; i < path_off[k + 1]; ++i) {
        const uint32_t p = path_pos[i], li = n_top + i;  // this node's label index
        uint32_t m0 = INF, m1 = INF, m2 = INF, m3 = INF;
        if (sj0 < sj1) {
            const uint32_t pr = pos_pre[p], ps = pos_sub[p], pd = pos_depth[p];
            for (uint32_t j = sj0; j < sj1; ++j) {
                const uint32_t sp = seed_pre[j];
                if (sp >= pr && sp - pr < ps) m0 = min(m0, pathbuf[(size_t)j * D1 + (seed_depth[j] - pd)]);
            }
        }
        // bottom sources: the ancestors in_dd levels up, path elements li - in_dd
        uint32_t a = in_off[p];
        const uint32_t as = in_split[p], a1 = in_off[p + 1];
        for (; a + 2 <= as; a += 2) {
            const uint32_t d0 = in_dd[a - ib], d1 = in_dd[a + 1 - ib];
            m0 = min(m0, sat_add(lab[(size_t)(li - d0) * KL], in_w[a]));
            m1 = min(m1, sat_add(lab[(size_t)(li - d1) * KL], in_w[a + 1]));
        }
        if (a < as) m0 = min(m0, sat_add(lab[(size_t)(li - in_dd[a - ib]) * KL], in_w[a]));
        // top sources: final labels at their positions
        for (a = as; a + 4 <= a1; a += 4) {
            const uint32_t u0 = in_src[a], u1 = in_src[a + 1], u2 = in_src[a + 2], u3 = in_src[a + 3];
            const uint32_t w0 = in_w[a], w1 = in_w[a + 1], w2 = in_w[a + 2], w3 = in_w[a + 3];
            m0 = min(m0, sat_add(lab[(size_t)u0 * KL], w0));
            m1 = min(m1, sat_add(lab[(size_t)u1 * KL], w1));
            m2 = min(m2, sat_add(lab[(size_t)u2 * KL], w2));
            m3 = min(m3, sat_add(lab[(size_t)u3 * KL], w3));
        }
        for (; a < a1; ++a) m0 = min(m0, sat_add(lab[(size_t)in_src[a] * KL], in_w[a]));
        lab[(size_t)li * KL] = min(min(m0, m1), min(m2, m3));
    }
}

// bottom_pull: bottom nodes [p0, p1) of one height level, group blockIdx.y:
// the chunk result min the DOWN arcs from bottom sources (higher levels).
template <int KL>
__global__ __launch_bounds__(NT) void bottom_pull_kernel(uint32_t p0, uint32_t p1, const uint32_t *__restrict__ in_off,
                                                         const uint32_t *__restrict__ in_split,
                                                         const uint32_t *__restrict__ in_src,
                                                         const uint32_t *__restrict__ in_w, BGate B,
                                                         uint32_t *__restrict__ arena, uint32_t n) {
    constexpr int NS = NT / KL;
    const uint32_t lane = threadIdx.x % KL, slot = slot_id<KL>();
    const uint32_t p = p0 + blockIdx.x * NS + slot;
    if (p >= p1) return;
    if (B.open_b && !bottom_open<KL>(B, blockIdx.y, p)) return;  // a dead tree: never read
    uint32_t *lab = arena + (size_t)blockIdx.y * n * KL + lane;
    uint32_t a = in_off[p];
    const uint32_t a1 = in_split[p];
    if (a == a1) return;
    if constexpr (KL >= 64) {
        uint32_t m = lab[(size_t)p * KL];
        const uint32_t wl = __lane_id();
        for (uint32_t base = a; base < a1; base += 64) {
            const uint32_t cnt = min(64u, a1 - base);
            uint32_t ms = 0, mw = INF;
            if (wl < cnt) { ms = in_src[base + wl]; mw = in_w[base + wl]; }
#pragma unroll 8
            for (uint32_t j = 0; j < cnt; ++j) {
                const uint32_t u = (uint32_t)__builtin_amdgcn_readlane((int)ms, (int)j);
                const uint32_t w = (uint32_t)__builtin_amdgcn_readlane((int)mw, (int)j);
                m = min(m, sat_add(lab[(size_t)u * KL], w));
            }
        }
        lab[(size_t)p * KL] = m;
        return;
    }
    uint32_t m0 = lab[(size_t)p * KL], m1 = INF, m2 = INF, m3 = INF;
    for (; a + 4 <= a1; a += 4) {
        const uint32_t u0 = in_src[a], u1 = in_src[a + 1], u2 = in_src[a + 2], u3 = in_src[a + 3];
        const uint32_t w0 = in_w[a], w1 = in_w[a + 1], w2 = in_w[a + 2], w3 = in_w[a + 3];
        m0 = min(m0, sat_add(lab[(size_t)u0 * KL], w0));
        m1 = min(m1, sat_add(lab[(size_t)u1 * KL], w1));
        m2 = min(m2, sat_add(lab[(size_t)u2 * KL], w2));
        m3 = min(m3, sat_add(lab[(size_t)u3 * KL], w3));
    }
    for (; a < a1; ++a) m0 = min(m0, sat_add(lab[(size_t)in_src[a] * KL], in_w[a]));
    lab[(size_t)p * KL] = min(min(m0, m1), min(m2, m3));
}

// ---------------------------------------------------------------------------
// gather: target columns of one batch into rows, limit-masked.  Tile = 64
// targets x KL lanes through LDS.  Row of lane gl: row_map[gl] if given, else
// gl.  Optional reach bits per row (t0 is a multiple of 64).
// ---------------------------------------------------------------------------
template <int KL>
__global__ __launch_bounds__(NT) void gather_kernel(
    const uint32_t *__restrict__ arena, uint32_t n, const uint32_t *__restrict__ tgt_pos, uint32_t T,
    const uint32_t *__restrict__ lane_limit, uint32_t n_lanes, const uint32_t *__restrict__ row_map, BGate B,
    uint32_t *__restrict__ out, uint64_t ld, unsigned long long *__restrict__ bits, uint64_t bits_ld, uint32_t g0,
    uint32_t lane_lo, uint32_t row_base) {
    __shared__ uint32_t tile[64][KL + 1];
    const uint32_t t0 = blockIdx.x * 64;
    const uint32_t g = g0 + blockIdx.y;
    const uint32_t *lab = arena + (size_t)g * n * KL;
    for (uint32_t idx = threadIdx.x; idx < 64 * KL; idx += NT) {
        const uint32_t ti = idx / KL, lane = idx % KL;
        const uint32_t t = t0 + ti;
        uint32_t v = INF;
        const uint32_t gl = g * KL + lane;
        if (t < T && gl >= lane_lo && gl < n_lanes) {
            const uint32_t p = tgt_pos[t];
            // positions of dead bottom trees were never written (with a gate,
            // tgt_pos holds positions)
            if (!B.open_b || p < B.n_top || bottom_open<KL>(B, g, p)) {
                v = lab[(size_t)p * KL + lane];
                if (v > lane_limit[gl]) v = INF;
            }
        }
        tile[ti][lane] = v;
    }
    __syncthreads();
    if (out) {
        for (uint32_t idx = threadIdx.x; idx < 64 * KL; idx += NT) {
            const uint32_t lane = idx / 64, ti = idx % 64;
            const uint32_t gl = g * KL + lane, t = t0 + ti;
            if (gl >= lane_lo && gl < n_lanes && t < T) {
                const uint64_t row = (row_map ? row_map[gl] : gl) - row_base;
                out[row * ld + t] = tile[ti][lane];
            }
        }
    }
    if (bits) {
        const uint32_t w = threadIdx.x / 64, ti = threadIdx.x % 64;
        for (uint32_t lane = w; lane < KL; lane += NT / 64) {
            const uint32_t gl = g * KL + lane;
            const bool r = (t0 + ti < T) && tile[ti][lane] != INF;
            const unsigned long long word = __ballot(r);
            if (ti == 0 && gl >= lane_lo && gl < n_lanes) {
                const uint64_t row = row_map ? row_map[gl] : gl;
                bits[row * bits_ld + (t0 / 64)] = word;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Isochrone lists: per lane, the settled nodes (rank, time <= limit) in
// ascending rank, the output of run_phast_bounded_fast_seeded
// (isochrone_handler.rs:453-471), for every lane of a batch at once.  Ranks
// are cut into superblocks of SB_RANKS (64 reach-bit words).
//   iso_count  per row: the reach bits' popcount per superblock, turned into
//              the row's exclusive offsets by superblock, and the row total
//   iso_scan   the batch's row totals into row starts (one workgroup)
//   iso_emit   per (superblock, lane group): 64-rank tiles through LDS as in
//              gather; per lane a ballot over the tile's ranks and the
//              position of each settled rank among them (mbcnt) give its slot
// ---------------------------------------------------------------------------
constexpr uint32_t SB_RANKS = 4096;
constexpr int ISO_THREADS = 256;

// Row rows_sorted[blockIdx.x] of the reach bits (bits_ld words per row):
// sb_off[i * nsb + b] = settled ranks of the row before superblock b,
// row_tot[i] = all of them.  Thread t takes a run of consecutive superblocks.
__global__ __launch_bounds__(ISO_THREADS) void iso_count_kernel(const unsigned long long *__restrict__ bits,
                                                                uint64_t words, uint64_t bits_ld,
                                                                const uint32_t *__restrict__ rows_sorted, uint32_t nsb,
                                                                uint32_t *__restrict__ sb_off,
                                                                uint32_t *__restrict__ row_tot) {
    __shared__ uint32_t part[ISO_THREADS];
    const uint32_t i = blockIdx.x, t = threadIdx.x;
    const unsigned long long *row = bits + (uint64_t)rows_sorted[i] * bits_ld;
    const uint32_t per = (nsb + ISO_THREADS - 1) / ISO_THREADS;
    const uint32_t b0 = min(nsb, t * per), b1 = min(nsb, b0 + per);
    uint32_t sum = 0;
    for (uint32_t b = b0; b < b1; ++b) {
        const uint64_t w1 = min<uint64_t>(words, (uint64_t)(b + 1) * (SB_RANKS / 64));
        for (uint64_t w = (uint64_t)b * (SB_RANKS / 64); w < w1; ++w) sum += __popcll(row[w]);
    }
    part[t] = sum;
    __syncthreads();
    // exclusive scan of the per-thread sums (Hillis-Steele in LDS)
    for (uint32_t o = 1; o < ISO_THREADS; o <<= 1) {
        const uint32_t v = t >= o ? part[t - o] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;  // ranks before this thread's run
    for (uint32_t b = b0; b < b1; ++b) {
        sb_off[(uint64_t)i * nsb + b] = run;
        const uint64_t w1 = min<uint64_t>(words, (uint64_t)(b + 1) * (SB_RANKS / 64));
        for (uint64_t w = (uint64_t)b * (SB_RANKS / 64); w < w1; ++w) run += __popcll(row[w]);
    }
    if (t == ISO_THREADS - 1) row_tot[i] = part[t];
}

// row_start[i] = base + row_tot[0] + .. + row_tot[i - 1]; total[0] = base + all.
__global__ __launch_bounds__(ISO_THREADS) void iso_scan_kernel(const uint32_t *__restrict__ row_tot, uint32_t nl,
                                                               uint64_t base, uint64_t *__restrict__ row_start,
                                                               uint64_t *__restrict__ total) {
    __shared__ uint64_t part[ISO_THREADS];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (nl + ISO_THREADS - 1) / ISO_THREADS;
    const uint32_t i0 = min(nl, t * per), i1 = min(nl, i0 + per);
    uint64_t sum = 0;
    for (uint32_t i = i0; i < i1; ++i) sum += row_tot[i];
    part[t] = sum;
    __syncthreads();
    for (uint32_t o = 1; o < ISO_THREADS; o <<= 1) {
        const uint64_t v = t >= o ? part[t - o] : 0ull;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint64_t run = base + part[t] - sum;
    for (uint32_t i = i0; i < i1; ++i) {
        row_start[i] = run;
        run += row_tot[i];
    }
    if (t == ISO_THREADS - 1) total[0] = base + part[t];
}

// Superblock blockIdx.x of lane group blockIdx.y: 64 tiles of 64 ranks, each
// read as in gather (limit mask, bottom gate); lane gl's pairs go from
// row_start[slot[gl]] + sb_off[slot[gl] * nsb + b] on, in rank order.
template <int KL>
__global__ __launch_bounds__(NT) void iso_emit_kernel(const uint32_t *__restrict__ arena, uint32_t n_lab,
                                                      const uint32_t *__restrict__ pos_of_rank, uint32_t n,
                                                      const unsigned long long *__restrict__ bits, uint64_t bits_ld,
                                                      const uint32_t *__restrict__ row_map, uint32_t n_lanes,
                                                      const uint32_t *__restrict__ slot,
                                                      const uint64_t *__restrict__ row_start,
                                                      const uint32_t *__restrict__ sb_off, uint32_t nsb,
                                                      uint2 *__restrict__ pairs) {
    __shared__ uint32_t tile[64][KL + 1];
    __shared__ uint64_t base[KL];
    __shared__ unsigned long long word[KL];
    const uint32_t b = blockIdx.x, g = blockIdx.y;
    const uint32_t *lab = arena + (size_t)g * n_lab * KL;
    for (uint32_t lane = threadIdx.x; lane < KL; lane += NT) {
        const uint32_t gl = g * KL + lane;
        base[lane] = gl < n_lanes ? row_start[slot[gl]] + sb_off[(uint64_t)slot[gl] * nsb + b] : 0ull;
    }
    for (uint32_t tt = 0; tt < SB_RANKS / 64; ++tt) {
        const uint32_t t0 = b * SB_RANKS + tt * 64;
        if (t0 >= n) break;  // workgroup-uniform
        __syncthreads();     // the previous tile is consumed
        for (uint32_t lane = threadIdx.x; lane < KL; lane += NT) {  // this tile's reach bits per lane
            const uint32_t gl = g * KL + lane;
            word[lane] = gl < n_lanes ? bits[(uint64_t)row_map[gl] * bits_ld + t0 / 64] : 0ull;
        }
        __syncthreads();
        for (uint32_t idx = threadIdx.x; idx < 64 * KL; idx += NT) {
            const uint32_t ti = idx / KL, lane = idx % KL;
            // a set bit marks a settled rank (label within the limit): only those are read
            tile[ti][lane] = (word[lane] >> ti) & 1ull ? lab[(size_t)pos_of_rank[t0 + ti] * KL + lane] : INF;
        }
        __syncthreads();
        // wavefront w takes lanes w, w + NT/64, ..; thread ti one rank
        const uint32_t w = threadIdx.x / 64, ti = threadIdx.x % 64;
        for (uint32_t lane = w; lane < KL; lane += NT / 64) {
            const uint32_t v = tile[ti][lane];
            const bool r = (t0 + ti < n) && v != INF && g * KL + lane < n_lanes;
            const uint64_t m = __ballot(r);
            const uint64_t at = base[lane];
            if (r) {
                const uint32_t before = __popcll(m & ((1ull << ti) - 1ull));
                pairs[at + before] = make_uint2(t0 + ti, v);
            }
            if (ti == 0) base[lane] = at + __popcll(m);  // this wavefront alone owns the lane
        }
    }
}

// table_gather: seeded table cells of one batch (table_phast_lopsided,
// bucket_ch.rs:6785-6811).  Tile = 64 targets x KL lanes through LDS, as in
// gather.  Cell (lane gl, target t): the minimum over t's seeds q of
// F - tpart[q] (saturating), F the lane's label at tlab[q] if F <= the
// lane's field bound; INF above thr.  A target seed rank equal to one of the
// lane's seed ranks makes the cell a conflict: INF, and conf = 1.
template <int KL>
__global__ __launch_bounds__(NT) void table_gather_kernel(
    const uint32_t *__restrict__ arena, uint32_t n, uint32_t T, const uint32_t *__restrict__ toff,
    const uint32_t *__restrict__ tlab, const uint32_t *__restrict__ trank, const uint32_t *__restrict__ tpart,
    const uint32_t *__restrict__ seed_off, const uint32_t *__restrict__ seed_rank,
    const uint32_t *__restrict__ lane_limit, uint32_t n_lanes, const uint32_t *__restrict__ row_map, uint32_t thr,
    uint32_t *__restrict__ out, uint8_t *__restrict__ conf) {
    __shared__ uint32_t tile[64][KL + 1];
    const uint32_t t0 = blockIdx.x * 64;
    const uint32_t g = blockIdx.y;
    const uint32_t *lab = arena + (size_t)g * n * KL;
    for (uint32_t idx = threadIdx.x; idx < 64 * KL; idx += NT) {
        const uint32_t ti = idx / KL, lane = idx % KL;
        const uint32_t t = t0 + ti, gl = g * KL + lane;
        uint32_t best = INF, clash = 0;
        if (t < T && gl < n_lanes) {
            const uint32_t lim = lane_limit[gl], s0 = seed_off[gl], s1 = seed_off[gl + 1];
            for (uint32_t q = toff[t]; q < toff[t + 1]; ++q) {
                const uint32_t f = lab[(size_t)tlab[q] * KL + lane], part = tpart[q], r = trank[q];
                if (f != INF && f <= lim) best = min(best, f > part ? f - part : 0u);
                for (uint32_t k = s0; k < s1; ++k) clash |= seed_rank[k] == r;
            }
            if (clash || (thr != INF && best > thr)) best = INF;
        }
        tile[ti][lane] = best;
        if (conf && t < T && gl < n_lanes) conf[(uint64_t)(row_map ? row_map[gl] : gl) * T + t] = (uint8_t)clash;
    }
    __syncthreads();
    for (uint32_t idx = threadIdx.x; idx < 64 * KL; idx += NT) {
        const uint32_t lane = idx / 64, ti = idx % 64;
        const uint32_t gl = g * KL + lane, t = t0 + ti;
        if (gl < n_lanes && t < T) {
            const uint64_t row = row_map ? row_map[gl] : gl;
            out[row * T + t] = tile[ti][lane];
        }
    }
}

// ---------------------------------------------------------------------------
// host context
// ---------------------------------------------------------------------------
// Device allocations the library has made (every one goes through DevBuf).
std::atomic<uint64_t> g_device_allocs{0};

// A device buffer: owned (hipMalloc'd, grown on demand) or borrowed (a fixed
// range of memory the library holds elsewhere, which is never reallocated: a
// request beyond its size fails with hipErrorOutOfMemory).
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    bool borrowed = false;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p && !borrowed) (void)hipFree(p);
        if (!borrowed) {
            p = nullptr;
            n = 0;
        }
    }
    hipError_t alloc(size_t count) {
        if (borrowed) return count <= n ? hipSuccess : hipErrorOutOfMemory;
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e == hipSuccess) {
            n = count;
            g_device_allocs.fetch_add(1, std::memory_order_relaxed);
        }
        else p = nullptr;
        return e;
    }
    // at least `count` elements; contents are not kept when it grows
    hipError_t ensure(size_t count) { return count <= n && p ? hipSuccess : alloc(count); }
    void borrow(T *ptr, size_t count) {
        release();
        p = ptr;
        n = count;
        borrowed = true;
    }
    hipError_t upload(const std::vector<T> &v, hipStream_t s) {
        hipError_t e = alloc(v.size());
        if (e != hipSuccess) return e;
        if (v.empty()) return hipSuccess;
        return hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
    }
};

// Waits for the stream when it goes out of scope: declared after host vectors
// that asynchronous uploads read, it keeps them alive until the copies are
// done, also when an error returns early.
struct StreamDrain {
    hipStream_t s;
    ~StreamDrain() { (void)hipStreamSynchronize(s); }
};

constexpr int SEG = 64;     // nodes of a segment (one register tile)
constexpr int CHUNK = 256;  // nodes of a bottom chunk (LDS = CHUNK * KL * 4 bytes)
// Subtree size at or below which a node is in the bottom region.  Small: the
// dense separator chains of the nested dissection become register tiles and a
// matrix's target paths through the bottom stay short.
constexpr uint32_t BOTTOM_LIMIT = 256;

bfg::LayoutParams layout_params() {
    bfg::LayoutParams lp;
    lp.segment_limit = SEG;
    lp.block_limit = CHUNK;
    lp.bottom_limit = BOTTOM_LIMIT;
    if (const char *e = std::getenv("BFG_BOTTOM_LIMIT")) {  // tuning override
        const long v = std::strtol(e, nullptr, 10);
        if (v > 0 && v < (1L << 30)) lp.bottom_limit = (uint32_t)v;
    }
    return lp;
}

// Two-channel sweeps (time, length along the time-optimal path): arcs as
// packed (time << 32 | length) weights in pull and push lists of their own,
// built by bfg_ctx_attach_length.
struct TwoChannel {
    // top region: per segment s, DOWN arcs from outside [x_off[s], x_off[s+1])
    // grouped by source, x_sj = source position << 6 | head offset in the
    // segment; per top head position p, arcs from its own segment
    // [i_off[p], i_off[p+1]), i_k = tail offset in the segment
    DevBuf<uint32_t> x_off, x_sj, i_off, i_k;
    DevBuf<uint64_t> x_w, i_w;
    // bottom heads p >= n_top: DOWN in-arcs [b_off[p - n_top], ..), sources b_src
    DevBuf<uint32_t> b_off, b_src;
    DevBuf<uint64_t> b_w;
    // UP arcs by tail position (targets are positions)
    DevBuf<uint32_t> u_off, u_tgt;
    DevBuf<uint64_t> u_w;
    DevBuf<uint32_t> pos_of_rank;
    // per call: labels [group][position][64 lanes], seeds, limits, outputs
    DevBuf<uint64_t> lab, s_lab;
    DevBuf<uint64_t> top;  // the upward phase's top-region labels, [lane][top position]
    DevBuf<uint32_t> s_off, s_pos, lim, o_t, o_l;
    uint64_t bytes = 0;
    // per segment level, the most outside arcs one of its segments has (host)
    std::vector<uint32_t> lvl_xmax;
};

}  // namespace

// DOWN work plan for batches of one wave-group count: units (segment |
// run | UNIT_WHOLE, rows [r0, r1), weights from w) by level.  Plans are shared
// by the calls in flight and never change once built.
struct Plan {
    uint32_t wgs = 0;
    std::vector<uint32_t> lvl;  // units of level l: [lvl[l], lvl[l+1])
    DevBuf<uint32_t> seg, r0, r1, w;
    DevBuf<uint8_t> split;      // 1 where the plan merges a segment's rows with atomics
};

// A call's lanes: per lane seeds and limit (host), original lane order.
struct Req {
    std::vector<uint32_t> soff, srank, scost, limit;
    std::vector<uint32_t> major;  // optional: lanes are sorted by it first (the request of a lane)
};

// The device state of one call in flight: its stream, label arena, seeds,
// target paths and gate flags.  A context runs up to `max_works` calls at
// once, each on its own Work (and stream); a Work whose buffers are borrowed
// from the library's static scratch array (`fixed`) never allocates.
struct Work {
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<hipEvent_t> dom_ev;  // event pairs around each DOWN row launch of a batch
    bool fixed = false;              // buffers borrowed from the scratch array
    bool time_dom = true;            // record the dominant kernel's launch times
    uint32_t lane_cap = 0;           // lanes per batch of the current call (multiple of 64)
    DevBuf<uint32_t> arena, pathbuf;
    DevBuf<uint32_t> seed_off, seed_rank, seed_pos, seed_cost, seed_depth, seed_pre, lane_limit, row_map;
    DevBuf<uint32_t> tgt_pos, path_off, path_pos;
    // gate flags of a bounded call: per (wave group, segment) whether a label
    // is within its lane's limit (statistics); per (lane group, bottom chunk)
    // and (lane group, bottom tree) whether the chunk gate left it open
    DevBuf<uint8_t> seg_live, chunk_ok, open_b;
    bool gate = false;
    // bottom paths above the targets of the current matrix call; while
    // paths_on, the downward sweep covers the top region and these paths only
    bool paths_on = false;  // while on, labels are compact (see label_stride)
    uint32_t n_paths = 0;
    uint64_t path_nodes = 0, path_arcs = 0;
    // host copies of the staged path tables, kept for their asynchronous
    // uploads; the last element of each path (its target); whether path_pos
    // of this call is on the device yet
    std::vector<uint32_t> path_off_h, path_pos_h, tgt_pos_h, path_leaf;
    bool path_pos_staged = false;
    // upward phase only (the many-to-many join): labels of the top region
    // alone (see label_stride), path slots for the bottom; the join's reach
    // bits and bottom pairs
    bool up_only = false;
    uint32_t cap_hint = 0;  // lanes per batch chosen by the caller (0: run_all chooses)
    DevBuf<unsigned long long> reach, visits;
    DevBuf<uint32_t> pairs;
    // upward-only batches: the reached (segment, wave group) pairs by segment
    // level (tri/fill) and the reached (push block, wave group) pairs
    // (push), host lists and their device copies; item ranges per level
    std::vector<uint32_t> h_tu, h_tg, h_pu, h_pg, h_toff, h_poff, h_mark;
    DevBuf<uint32_t> tu, tg, pu, pg;
    DevBuf<uint32_t> segmin;  // upward-only batches: [wave group][segment][8 lanes] least final label
    // upward-only batches of 64-lane groups: the 16-bit label of each reached
    // (segment, wave group) pair (0xFFFF for INF, else min(label, CAP16)),
    // [list item][64 positions][64 lanes], and the item of each pair,
    // [wave group][segment] (NONE: not reached)
    DevBuf<uint16_t> lab16, lmin16;  // lmin16: [list item][64 lanes] the least 16-bit label of the segment
    DevBuf<uint32_t> slot, redo;
    bool lab16_ok = false;  // the last upward-only batch wrote lab16 and slot
    std::vector<uint32_t> h_pairs;
    // host staging of a call (kept, so a fixed Work reuses its capacity)
    std::vector<uint32_t> h_bt, h_pending, h_order, h_soff, h_srank, h_scost, h_lim, h_rows, h_sdepth, h_spre,
        h_spos, h_src, h_tgt;
    std::vector<uint64_t> h_key;
    std::vector<std::pair<uint32_t, uint32_t>> h_elem;
    Req req;  // the request of a call that must not allocate (the scratch Work)
    bfg_timings last{};

    Work() = default;
    Work(const Work &) = delete;
    Work &operator=(const Work &) = delete;
    ~Work() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto &e : dom_ev) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct bfg_ctx {
    int device = 0;
    bfg::Layout L;
    // configuration: calls hold it shared, configure / tune / attach_length
    // exclusively
    std::shared_mutex cfg;
    uint32_t lanes = 64;        // KL: sources per lane group (16, 32, 64 or 128)
    uint32_t lane_cap = 0;      // lanes per batch when cap_fixed (multiple of 64)
    bool cap_fixed = false;
    // tuning set by bfg_ctx_tune (0: the defaults or their environment overrides)
    uint32_t plan_rows = 0, tri_segs = 0;
    // resident graph
    DevBuf<uint32_t> parent, pos_parent, pos_pre, pos_sub, pos_depth, seg_pos;
    DevBuf<uint32_t> in_off, in_split, in_src, in_w;
    DevBuf<uint32_t> ch_off, ch_pos, cs_off, cs_src, cs_aoff, ca_w;
    DevBuf<uint16_t> ca_v, in_dd;
    DevBuf<uint32_t> up_off, up_tgt, up_w;
    DevBuf<uint32_t> xr_off, xr_run, xw_off, xr_src, xr_w, tb_dn, tb_up, tb_w, pt_blk, pt_tgt, pr_off, pr_wid,
        pw_off, pr_x, pr_w;
    // pr_w as 16 bits in column pairs (0xFFFF for INF, else min(w, 0x7FFF)):
    // the upward-list pushes
    DevBuf<uint32_t> pr_w16;
    DevBuf<uint32_t> pk_seg;
    // gate tables: the least weight of each DOWN row (xr order) and of each
    // bottom chunk source's arcs (cs order); the bottom trees (maximal
    // subtrees of the bottom region) in pre-order: the tree of each bottom
    // position and the chunks each tree touches, [bt_ch[2t], bt_ch[2t+1]]
    DevBuf<uint32_t> xr_min, cs_min, pos_bt, bt_ch;
    uint32_t n_bt = 0;
    // the packed 16-bit DOWN rows: weights min(w, 0xFFFF) two per dword (xr_w
    // order) and per row the cap 0xFFFF - (its largest finite weight), 0 when
    // that is 0xFFFF or more; label_bits 16 uses them, 32 keeps the u32 rows
    DevBuf<uint32_t> xr_w16, xr_cap;
    uint32_t label_bits = 16;
    DevBuf<uint32_t> pos_of_rank_d;  // position of each rank (the dense gathers' target table)
    std::vector<uint64_t> lvl_dom_bytes;  // algorithmic bytes of a level's DOWN rows per wave group
    uint64_t top_in_arcs = 0;
    // bdep[p]: levels between bottom position p and the top region (0 for a
    // bottom root)
    std::vector<uint32_t> bdep;
    // DOWN plans by wave-group count (a few recent ones; calls in flight hold
    // theirs), and the one-wave-group plan the scratch Work uses, built at
    // creation
    std::mutex plan_mu;
    std::map<uint32_t, std::shared_ptr<const Plan>> plans;
    std::vector<uint32_t> plan_lru;
    std::shared_ptr<const Plan> plan1;
    // Works: idle ones wait in the pool; at most max_works exist
    std::mutex pool_mu;
    std::condition_variable pool_cv;
    std::vector<std::unique_ptr<Work>> works;
    std::vector<Work *> idle;
    uint32_t max_works = 4;
    // the Work over the static scratch array (bfg_matrix_device_noalloc)
    std::unique_ptr<Work> scratch_work;
    // timings of the last finished call
    std::mutex last_mu;
    bfg_timings last{};
    uint64_t graph_bytes = 0;
    std::mutex t2_mu;                // two-channel calls share the label buffers below
    std::unique_ptr<TwoChannel> t2;  // set by bfg_ctx_attach_length
    // the many-to-many join: the context over the transposed graph (same
    // positions, set by bfg_ctx_pair_reverse; not owned) and the context this
    // one is the reverse of
    bfg_ctx *rev = nullptr, *rev_of = nullptr;
    // host tables of the upward lists: the segment of each top position, the
    // segment holding each segment head's parent (NONE at roots), each
    // segment's level (built by bfg_ctx_pair_reverse for both contexts)
    std::vector<uint32_t> seg_of_pos, seg_par, seg_lvl;
    std::vector<uint32_t> top_anc;  // per bottom position p (index p - n_top): its first top ancestor, or NONE
    std::vector<uint32_t> bt_root;  // per bottom position p (index p - n_top): the root of its bottom tree

    bfg_ctx() = default;
    bfg_ctx(const bfg_ctx &) = delete;
    bfg_ctx &operator=(const bfg_ctx &) = delete;
    ~bfg_ctx() {
        if (rev) rev->rev_of = nullptr;  // a pair ends with either of its contexts
        if (rev_of) rev_of->rev = nullptr;
        (void)hipSetDevice(device);
        scratch_work.reset();
        works.clear();  // each Work drains and releases its stream
    }
};

namespace {

uint32_t d1_of(const bfg_ctx *c) { return c->L.max_height + 1; }

// ---------------------------------------------------------------------------
// Works and their pool
// ---------------------------------------------------------------------------
int work_init(Work *w) {
    HIP_TRY(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
    for (auto &e : w->ev) HIP_TRY(hipEventCreate(&e));
    return BFG_SUCCESS;
}

// A Work of the pool held for one call; it goes back to the pool (its stream
// drained) when the lease ends.
struct WorkLease {
    bfg_ctx *c = nullptr;
    Work *w = nullptr;
    WorkLease() = default;
    WorkLease(const WorkLease &) = delete;
    WorkLease &operator=(const WorkLease &) = delete;
    ~WorkLease() {
        if (!c || !w) return;
        (void)hipStreamSynchronize(w->stream);
        {
            std::lock_guard<std::mutex> g(c->pool_mu);
            c->idle.push_back(w);
        }
        c->pool_cv.notify_one();
    }
};

int acquire_work(bfg_ctx *c, WorkLease &lease) {
    std::unique_lock<std::mutex> g(c->pool_mu);
    for (;;) {
        if (!c->idle.empty()) {
            lease.w = c->idle.back();  // the most recent one: its buffers are likely the right size
            c->idle.pop_back();
            lease.c = c;
            return BFG_SUCCESS;
        }
        if (c->works.size() < c->max_works) {
            std::unique_ptr<Work> w(new Work());
            const int rc = work_init(w.get());
            if (rc) return rc;
            lease.w = w.get();
            lease.c = c;
            c->works.push_back(std::move(w));
            return BFG_SUCCESS;
        }
        c->pool_cv.wait(g);
    }
}

// Frees the label arenas of the idle Works (after an allocation failure).
void trim_idle(bfg_ctx *c) {
    std::lock_guard<std::mutex> g(c->pool_mu);
    for (Work *w : c->idle) w->arena.release();
}

// A call in progress: the configuration held shared, the device current and
// a Work leased (released in reverse order).
struct CallScope {
    std::shared_lock<std::shared_mutex> cfg;
    WorkLease lease;
    explicit CallScope(bfg_ctx *c) : cfg(c->cfg) {}
};

int begin_call(bfg_ctx *c, CallScope &cs) {
    HIP_TRY(hipSetDevice(c->device));
    return acquire_work(c, cs.lease);
}

// Labels per lane in the arena: all n positions, or for a matrix the top
// region [0, n_top) followed by the nodes of its target paths (path_pos
// element i at n_top + i).
uint32_t label_stride(const bfg_ctx *c, const Work *w) {
    if (w->up_only) return c->L.n_top;
    return w->paths_on ? c->L.n_top + (uint32_t)w->path_nodes : c->L.n;
}

// A failed allocation of a Work over the scratch array is a request beyond
// that array's limits.
int alloc_fail(const Work *w, hipError_t e, const char *what) {
    if (w->fixed)
        return fail(BFG_NOT_SUPPORTED, std::string(what) + " exceeds the scratch array of the allocation-free call");
    return fail(e == hipErrorOutOfMemory ? BFG_OUT_OF_MEMORY : BFG_DEVICE_ERROR,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define ENSURE(w, buf, count, what)                                      \
    do {                                                                 \
        hipError_t e_ = (w)->buf.ensure(count);                          \
        if (e_ != hipSuccess) return alloc_fail((w), e_, what);          \
    } while (0)

int ensure_lane_buffers(bfg_ctx *c, Work *w, size_t n_lanes, size_t n_seeds) {
    ENSURE(w, seed_off, n_lanes + 1, "seed offsets");
    ENSURE(w, lane_limit, n_lanes, "lane limits");
    ENSURE(w, row_map, n_lanes, "row map");
    ENSURE(w, seed_rank, n_seeds, "seeds");
    ENSURE(w, seed_pos, n_seeds, "seeds");
    ENSURE(w, seed_cost, n_seeds, "seeds");
    ENSURE(w, seed_depth, n_seeds, "seeds");
    ENSURE(w, seed_pre, n_seeds, "seeds");
    ENSURE(w, pathbuf, n_seeds * (size_t)d1_of(c), "path slots");
    return BFG_SUCCESS;
}

int ensure_arena(bfg_ctx *c, Work *w, size_t need) {
    hipError_t e = w->arena.ensure(need);
    if (e == hipErrorOutOfMemory && !w->fixed) {  // idle Works may hold the memory
        trim_idle(c);
        e = w->arena.ensure(need);
    }
    if (e != hipSuccess) return alloc_fail(w, e, "label arena");
    return BFG_SUCCESS;
}

// Lanes per batch: 60 % of free HBM (plus this Work's own arena) over the
// bytes one lane needs (labels, path slots, and `extra` per lane of outputs
// the caller keeps on the device per batch), capped by the request; a
// multiple of 64 (the top-region kernels carry 64 lanes per wavefront).
uint32_t auto_lane_cap(bfg_ctx *c, Work *w, uint32_t want, size_t extra) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)8 << 30;
    const size_t have = free_b + w->arena.n * sizeof(uint32_t);
    const size_t per_lane = (size_t)label_stride(c, w) * 4 + (size_t)d1_of(c) * 4 + 64 + extra;
    const size_t budget = (size_t)(have * 0.6);
    size_t cap = std::max<size_t>(budget / std::max<size_t>(per_lane, 1), 64);
    cap = std::min<size_t>(cap, 16384);
    cap = std::min<size_t>(cap, ((size_t)want + 63) / 64 * 64);
    cap = std::max<size_t>(cap / 64 * 64, 64);
    return (uint32_t)cap;
}

// Wavefronts a level's DOWN rows aim to keep in flight, and the fewest and the
// most rows one unit takes (the longest unit bounds its level's time).
constexpr uint64_t PLAN_WAVES = 8192;
constexpr uint32_t PLAN_MIN_ROWS = 64;
constexpr uint32_t PLAN_MAX_ROWS = 256;
constexpr size_t PLAN_CACHE = 8;  // DOWN plans kept per context

uint32_t plan_max_rows() {
    if (const char *e = std::getenv("BFG_PLAN_MAX_ROWS")) {  // tuning override
        const long v = std::strtol(e, nullptr, 10);
        if (v >= PLAN_MIN_ROWS && v < (1L << 24)) return (uint32_t)v;
    }
    return PLAN_MAX_ROWS;
}

// DOWN levels of more than this many segments relax their intra-segment
// blocks with scalar-loaded weights (tri_apply_s); fewer segments leave too
// few wavefronts to hide the loads, and the vector-loaded form waits only
// once.  The upward phase keeps the vector form: few wave groups reach a
// wide level's segments, so its levels are short of wavefronts too.
constexpr uint32_t TRI_SCALAR_SEGS = 64;

uint32_t tri_scalar_segs() {
    if (const char *e = std::getenv("BFG_TRI_SCALAR_SEGS")) {  // tuning override
        const long v = std::strtol(e, nullptr, 10);
        if (v >= 0 && v < (1L << 30)) return (uint32_t)v;
    }
    return TRI_SCALAR_SEGS;
}

// DOWN work plan for batches of `wgs` wave groups.  Per level, a row budget
// R = rows of the level * wgs / PLAN_WAVES, clamped to [PLAN_MIN_ROWS,
// PLAN_MAX_ROWS]: a segment with at most R rows is one whole unit; a longer
// one is cut into units of R rows merged with atomics.  Uploads on stream s.
int build_plan(bfg_ctx *c, hipStream_t s, uint32_t wgs, Plan &P) {
    const bfg::Layout &L = c->L;
    const uint32_t nsl = (uint32_t)L.seglvl_seg.size() - 1;
    const uint64_t rmax = c->plan_rows ? c->plan_rows : plan_max_rows();
    if (L.seg_pos.size() - 1 > UNIT_SEG_MASK) return fail(BFG_NOT_SUPPORTED, "too many segments for the DOWN plan");
    std::vector<uint32_t> us, u0, u1, uw;
    std::vector<uint8_t> split(L.seg_pos.size() - 1, 0);
    P.wgs = wgs;
    P.lvl.assign(1, 0);
    uint64_t covered = 0;
    for (uint32_t l = 0; l < nsl; ++l) {
        const uint32_t s0 = L.seglvl_seg[l], s1 = L.seglvl_seg[l + 1];
        const uint64_t rows = L.xr_off[s1] - L.xr_off[s0];
        const uint64_t R = std::min<uint64_t>(
            rmax, std::max<uint64_t>(PLAN_MIN_ROWS, (rows * wgs + PLAN_WAVES - 1) / PLAN_WAVES));
        for (uint32_t sg = s0; sg < s1; ++sg) {
            const uint32_t a = L.xr_off[sg], b = L.xr_off[sg + 1];
            const uint32_t *run = &L.xr_run[8 * (size_t)sg];
            uint32_t nrun = 0;
            for (uint32_t k = 0, st = 0; k < 8; st = run[k++]) nrun += run[k] > st;
            uint64_t w = L.xw_off[sg];
            if (nrun == 0) {  // no rows: a whole unit still gives unreached wave groups their INF labels
                us.push_back(sg | 7u << UNIT_RUN_SHIFT | UNIT_WHOLE);
                u0.push_back(a);
                u1.push_back(a);
                uw.push_back((uint32_t)w);
            }
            for (uint32_t k = 0, st = 0; k < 8; st = run[k++]) {
                const uint32_t W = 8 * (8 - k), cs = sg | k << UNIT_RUN_SHIFT;
                if (run[k] == st) continue;  // an empty run (the increment still advances st)
                if (nrun == 1 && run[k] - st <= R) {  // the whole segment in one unit
                    us.push_back(cs | UNIT_WHOLE);
                    u0.push_back(a + st);
                    u1.push_back(a + run[k]);
                    uw.push_back((uint32_t)w);
                } else {
                    for (uint64_t r = st; r < run[k]; r += R) {
                        us.push_back(cs);
                        u0.push_back((uint32_t)(a + r));
                        u1.push_back((uint32_t)(a + std::min<uint64_t>(r + R, run[k])));
                        uw.push_back((uint32_t)(w + (r - st) * W));
                    }
                    split[sg] = 1;
                }
                w += (uint64_t)(run[k] - st) * W;
            }
            if (w != L.xw_off[sg + 1] || a + run[7] != b) return fail(BFG_INTERNAL_ERROR, "row runs disagree");
        }
        P.lvl.push_back((uint32_t)us.size());
    }
    for (size_t k = 0; k < us.size(); ++k) covered += u1[k] - u0[k];
    if (covered != L.xr_off.back()) return fail(BFG_INTERNAL_ERROR, "DOWN plan does not cover every row");
    const StreamDrain drain{s};
    HIP_TRY(P.seg.upload(us, s));
    HIP_TRY(P.r0.upload(u0, s));
    HIP_TRY(P.r1.upload(u1, s));
    HIP_TRY(P.w.upload(uw, s));
    HIP_TRY(P.split.upload(split, s));
    HIP_TRY(hipStreamSynchronize(s));  // the host vectors die with this frame
    return BFG_SUCCESS;
}

// The DOWN plan for `wgs` wave groups: cached per context (a few recent
// ones), built on first use.  A Work over the scratch array takes the
// one-wave-group plan built at creation and builds nothing.
int get_plan(bfg_ctx *c, Work *w, uint32_t wgs, std::shared_ptr<const Plan> &out) {
    if (w->fixed) {
        if (wgs != 1 || !c->plan1) return fail(BFG_INTERNAL_ERROR, "no plan for the allocation-free call");
        out = c->plan1;
        return BFG_SUCCESS;
    }
    std::lock_guard<std::mutex> g(c->plan_mu);
    auto it = c->plans.find(wgs);
    if (it == c->plans.end()) {
        std::shared_ptr<Plan> P(new Plan());
        const int rc = build_plan(c, w->stream, wgs, *P);
        if (rc) return rc;
        it = c->plans.emplace(wgs, std::move(P)).first;
    }
    out = it->second;
    auto &lru = c->plan_lru;
    lru.erase(std::remove(lru.begin(), lru.end(), wgs), lru.end());
    lru.push_back(wgs);
    while (lru.size() > PLAN_CACHE) {  // calls in flight keep theirs alive
        c->plans.erase(lru.front());
        lru.erase(lru.begin());
    }
    return BFG_SUCCESS;
}

int stage_path_elements(bfg_ctx *c, Work *w);

struct PhaseTimes {
    float up_ms = 0, sweep_ms = 0, dom_ms = 0;
    uint32_t launches = 0, dom_launches = 0;
    uint64_t dom_bytes = 0;
};

Gate gate_of(const bfg_ctx *c, const Work *w, uint32_t nl) {
    return Gate{w->gate ? c->xr_min.p : nullptr, w->lane_limit.p, nl, w->seg_live.p,
                (uint32_t)c->L.seg_pos.size() - 1};
}

BGate bgate_of(const bfg_ctx *c, const Work *w) {
    return BGate{w->gate ? w->open_b.p : nullptr, c->pos_bt.p, c->n_bt, c->L.n_top};
}

CGate cgate_of(const bfg_ctx *c, const Work *w) {
    return CGate{w->gate ? c->cs_min.p : nullptr, w->lane_limit.p, w->chunk_ok.p, (uint32_t)c->L.ch_off.size() - 1};
}

// The host tables of the upward lists (see bfg_ctx): built once per context.
void upward_list_tables(bfg_ctx *c) {
    if (!c->seg_of_pos.empty()) return;
    const bfg::Layout &L = c->L;
    const uint32_t nseg = (uint32_t)L.seg_pos.size() - 1, nsl = (uint32_t)L.seglvl_seg.size() - 1;
    c->seg_of_pos.assign(L.n_top, NONE);
    c->seg_par.assign(nseg, NONE);
    c->seg_lvl.assign(nseg, 0);
    for (uint32_t s = 0; s < nseg; ++s)
        for (uint32_t p = L.seg_pos[s]; p < L.seg_pos[s + 1]; ++p) c->seg_of_pos[p] = s;
    for (uint32_t l = 0; l < nsl; ++l)
        for (uint32_t s = L.seglvl_seg[l]; s < L.seglvl_seg[l + 1]; ++s) c->seg_lvl[s] = l;
    for (uint32_t s = 0; s < nseg; ++s) {
        const uint32_t q = L.pos_parent[L.seg_pos[s]];
        c->seg_par[s] = q != NONE && q < L.n_top ? c->seg_of_pos[q] : NONE;
    }
    // bottom positions come by descending height, so a parent before its children
    c->top_anc.assign(L.n - L.n_top, NONE);
    c->bt_root.assign(L.n - L.n_top, NONE);
    for (uint32_t p = L.n_top; p < L.n; ++p) {
        const uint32_t q = L.pos_parent[p];
        const bool root = q == NONE || q < L.n_top;
        c->top_anc[p - L.n_top] = root ? q : c->top_anc[q - L.n_top];
        c->bt_root[p - L.n_top] = root ? p : c->bt_root[q - L.n_top];
    }
}

// The reached units of an upward-only batch of nl lanes (seeds staged in
// h_soff / h_spos): each lane's root path in segments, from the segment of
// its first top ancestor up, marked once per wave group (a path already
// marked for the group is marked up to the root).  The pairs are sorted by
// segment (levels ascend with segment ids), then wave group, and each
// reached segment's push blocks listed block by block with its wave groups
// side by side, so the four wavefronts of a workgroup mostly share a unit
// (and its weights in the scalar cache).  Uploaded on the Work's stream.
int build_up_lists(bfg_ctx *c, Work *w, uint32_t nl) {
    const bfg::Layout &L = c->L;
    const uint32_t nseg = (uint32_t)L.seg_pos.size() - 1, nsl = (uint32_t)L.seglvl_seg.size() - 1;
    if (c->seg_of_pos.size() != L.n_top) return fail(BFG_INTERNAL_ERROR, "upward lists without their tables");
    std::vector<uint32_t> &mark = w->h_mark, &toff = w->h_toff, &poff = w->h_poff;
    mark.assign(nseg, 0);
    toff.assign(nsl + 1, 0);
    std::vector<uint32_t> us, gs, soff(nseg + 1, 0);  // in discovery order; per segment counts
    for (uint32_t i = 0; i < nl; ++i) {
        const uint32_t g = i / 64;
        for (uint32_t j = w->h_soff[i]; j < w->h_soff[i + 1]; ++j) {
            uint32_t p = w->h_spos[j];
            if (p >= L.n_top) p = c->top_anc[p - L.n_top];
            if (p == NONE) continue;  // a bottom tree without a top ancestor
            for (uint32_t s = c->seg_of_pos[p]; s != NONE && mark[s] != g + 1; s = c->seg_par[s]) {
                mark[s] = g + 1;
                us.push_back(s);
                gs.push_back(g);
                soff[s + 1]++;
            }
        }
    }
    // counting sort by segment (a stable one: wave groups ascend within a segment)
    for (uint32_t s = 0; s < nseg; ++s) soff[s + 1] += soff[s];
    std::vector<uint32_t> &tu = w->h_tu, &tg = w->h_tg, &pu = w->h_pu, &pg = w->h_pg;
    tu.resize(us.size());
    tg.resize(gs.size());
    {
        std::vector<uint32_t> at(soff.begin(), soff.end() - 1);
        for (size_t k = 0; k < us.size(); ++k) {
            const uint32_t d = at[us[k]]++;
            tu[d] = us[k];
            tg[d] = gs[k];
        }
    }
    for (uint32_t l = 0; l <= nsl; ++l) toff[l] = soff[L.seglvl_seg[l]];
    // the push list: per level, every target block of a reached segment once
    // per reached wave group; sized first, then filled by threads over ranges
    // of levels of about equal size
    poff.assign(nsl + 1, 0);
    uint64_t total_p = 0;
    for (uint32_t l = 0; l < nsl; ++l) {
        for (uint32_t s = L.seglvl_seg[l]; s < L.seglvl_seg[l + 1]; ++s)
            total_p += (uint64_t)(L.pt_blk[s + 1] - L.pt_blk[s]) * (soff[s + 1] - soff[s]);
        if (total_p > 0xFFFFFFFFull) return fail(BFG_NOT_SUPPORTED, "more than 2^32 - 1 upward push units");
        poff[l + 1] = (uint32_t)total_p;
    }
    pu.resize(total_p);
    pg.resize(total_p);
    auto fill_levels = [&](uint32_t l0, uint32_t l1) {
        for (uint32_t l = l0; l < l1; ++l) {
            uint32_t o = poff[l];
            for (uint32_t s = L.seglvl_seg[l]; s < L.seglvl_seg[l + 1]; ++s)
                for (uint32_t b = L.pt_blk[s]; b < L.pt_blk[s + 1]; ++b)
                    for (uint32_t k = soff[s]; k < soff[s + 1]; ++k, ++o) {
                        pu[o] = b;
                        pg[o] = tg[k];
                    }
        }
    };
    const uint32_t nth = total_p < (1u << 20) ? 1u : std::min(8u, std::max(1u, std::thread::hardware_concurrency()));
    if (nth == 1) {
        fill_levels(0, nsl);
    } else {
        std::vector<std::thread> th;
        uint32_t l0 = 0;
        for (uint32_t t = 1; t <= nth && l0 < nsl; ++t) {
            const uint64_t goal = total_p * t / nth;
            uint32_t l1 = l0;
            while (l1 < nsl && (t == nth || poff[l1 + 1] <= goal)) ++l1;
            if (l1 == l0) l1 = l0 + 1;
            th.emplace_back(fill_levels, l0, l1);
            l0 = l1;
        }
        for (auto &h : th) h.join();
    }
    ENSURE(w, tu, std::max<size_t>(tu.size(), 1), "upward lists");
    ENSURE(w, tg, std::max<size_t>(tg.size(), 1), "upward lists");
    ENSURE(w, pu, std::max<size_t>(pu.size(), 1), "upward lists");
    ENSURE(w, pg, std::max<size_t>(pg.size(), 1), "upward lists");
    // the host vectors live in the Work until the batch has synchronised
    if (!tu.empty()) {
        HIP_TRY(hipMemcpyAsync(w->tu.p, tu.data(), tu.size() * 4, hipMemcpyHostToDevice, w->stream));
        HIP_TRY(hipMemcpyAsync(w->tg.p, tg.data(), tg.size() * 4, hipMemcpyHostToDevice, w->stream));
    }
    if (!pu.empty()) {
        HIP_TRY(hipMemcpyAsync(w->pu.p, pu.data(), pu.size() * 4, hipMemcpyHostToDevice, w->stream));
        HIP_TRY(hipMemcpyAsync(w->pg.p, pg.data(), pg.size() * 4, hipMemcpyHostToDevice, w->stream));
    }
    return BFG_SUCCESS;
}

// The upward phase of an upward-only batch over the lists of its reached
// units: INF fill, the bottom walk, then per segment level (deepest first)
// the intra-segment UP blocks and the UP pushes of the reached pairs only.
template <int KL>
int launch_up_lists(bfg_ctx *c, Work *w, uint32_t nl, PhaseTimes *pt) {
    const bfg::Layout &L = c->L;
    hipStream_t s = w->stream;
    const uint32_t n = label_stride(c, w), D1 = d1_of(c);
    const uint32_t nsl = (uint32_t)L.seglvl_seg.size() - 1;
    int rc = build_up_lists(c, w, nl);
    if (rc) return rc;
    constexpr uint32_t WPB = TILE_THREADS / 64;
    uint32_t nlaunch = 0;
    // per (wave group, segment, 8 lanes): the least label of the segment once
    // final, INF where the wave group does not reach it
    const uint32_t nseg = (uint32_t)L.seg_pos.size() - 1;
    const uint64_t nmin = (uint64_t)(nl + 63) / 64 * nseg * 8;
    ENSURE(w, segmin, std::max<size_t>(nmin, 1), "segment minima");
    HIP_TRY(bfg_meet::launch_fill(s, w->segmin.p, nmin, INF));
    const uint32_t total = (uint32_t)w->h_tu.size();
    uint16_t *l16 = nullptr;
    uint32_t *slot = nullptr;
    w->lab16_ok = false;
    if constexpr (KL == 64) {
        // the 16-bit copy serves the join alone: without memory for it the
        // join runs on the u32 labels
        if (w->lab16.ensure(std::max<size_t>((size_t)total * 4096, 1)) == hipSuccess &&
            w->slot.ensure(std::max<size_t>(nmin / 8, 1)) == hipSuccess &&
            w->lmin16.ensure(std::max<size_t>((size_t)total * 64, 1)) == hipSuccess) {
            HIP_TRY(bfg_meet::launch_fill(s, w->slot.p, nmin / 8, NONE));
            l16 = w->lab16.p;
            slot = w->slot.p;
            w->lab16_ok = true;
        } else {
            (void)hipGetLastError();
        }
    }
    HIP_TRY(hipEventRecord(w->ev[0], s));
    if (total) {
        hipLaunchKernelGGL(fill_list_kernel<KL>, dim3((total + WPB - 1) / WPB), dim3(TILE_THREADS), 0, s, total,
                           w->tu.p, w->tg.p, c->seg_pos.p, w->arena.p, n);
        ++nlaunch;
    }
    hipLaunchKernelGGL(bottom_walk_kernel<KL>, dim3(nl), dim3(64), 0, s, w->seed_off.p, w->seed_rank.p,
                       w->seed_pos.p, w->seed_cost.p, w->seed_depth.p, w->lane_limit.p, c->parent.p,
                       c->pos_parent.p, c->pos_depth.p, c->up_off.p, c->up_tgt.p, c->up_w.p, L.n_top, w->pathbuf.p,
                       D1, w->arena.p, n);
    ++nlaunch;
    const char *p32 = std::getenv("BFG_PUSH_U32");  // override: the u32 pushes alone
    const bool push16 = c->pr_w16.n > 0 && !(p32 && p32[0] == '1');
    for (int l = (int)nsl - 1; l >= 0; --l) {
        const uint32_t t0 = w->h_toff[l], tc = w->h_toff[l + 1] - t0;
        if (tc) {
            hipLaunchKernelGGL(tri_list_kernel<KL>, dim3((tc + WPB - 1) / WPB), dim3(TILE_THREADS), 0, s, tc,
                               w->tu.p + t0, w->tg.p + t0, c->seg_pos.p, c->tb_up.p, c->tb_w.p, w->arena.p, n,
                               w->segmin.p, nseg, l16, slot, w->lmin16.p, t0);
            ++nlaunch;
        }
        const uint32_t p0 = w->h_poff[l], pc = w->h_poff[l + 1] - p0;
        if (pc) {
            hipLaunchKernelGGL(push_list_kernel<KL>, dim3((pc + WPB - 1) / WPB), dim3(TILE_THREADS), 0, s, pc,
                               w->pu.p + p0, w->pg.p + p0, c->pr_off.p, c->pr_wid.p, c->pw_off.p, c->pt_tgt.p,
                               c->pr_x.p, c->pr_w.p, push16 ? c->pr_w16.p : nullptr, w->arena.p, n);
            ++nlaunch;
        }
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(w->ev[1], s));
    HIP_TRY(hipEventSynchronize(w->ev[1]));
    float a = 0;
    HIP_TRY(hipEventElapsedTime(&a, w->ev[0], w->ev[1]));
    pt->up_ms += a;
    pt->launches += nlaunch;
    return BFG_SUCCESS;
}

template <int KL>
int launch_phases(bfg_ctx *c, Work *w, const Plan &P, uint32_t nl, PhaseTimes *pt) {
    const bfg::Layout &L = c->L;
    hipStream_t s = w->stream;
    const uint32_t groups = (nl + KL - 1) / KL;
    const uint32_t wgs = (nl + 63) / 64;            // wave groups of the top-region kernels
    const uint32_t n = label_stride(c, w), D1 = d1_of(c);  // n: labels per lane in the arena
    const uint32_t nsl = (uint32_t)L.seglvl_seg.size() - 1;
    const uint32_t nseg = (uint32_t)L.seg_pos.size() - 1;
    uint32_t nlaunch = 0;
    HIP_TRY(hipEventRecord(w->ev[0], s));
    const Reach R{w->seed_off.p, w->seed_pre.p, c->pos_pre.p, c->pos_sub.p, nl};
    const Gate G = gate_of(c, w, nl), G0{nullptr, nullptr, 0, nullptr, 0};
    const CGate CG = cgate_of(c, w);
    const BGate B = bgate_of(c, w);
    constexpr uint32_t WPB = TILE_THREADS / 64;  // wave groups per workgroup
    const uint32_t wgx = (wgs + WPB - 1) / WPB;  // workgroups across the wave groups
    const uint32_t tss = c->tri_segs ? c->tri_segs - 1 : tri_scalar_segs();
    if (w->up_only) return launch_up_lists<KL>(c, w, nl, pt);
    if (nseg) {
        hipLaunchKernelGGL(fill_reach_kernel<KL>, dim3(wgs, (nseg + WPB - 1) / WPB), dim3(TILE_THREADS), 0, s, nseg,
                           c->seg_pos.p, R, w->arena.p, n);
        ++nlaunch;
    }
    hipLaunchKernelGGL(bottom_walk_kernel<KL>, dim3(nl), dim3(64), 0, s, w->seed_off.p, w->seed_rank.p,
                       w->seed_pos.p, w->seed_cost.p, w->seed_depth.p, w->lane_limit.p, c->parent.p,
                       c->pos_parent.p, c->pos_depth.p, c->up_off.p, c->up_tgt.p, c->up_w.p, L.n_top, w->pathbuf.p,
                       D1, w->arena.p, n);
    ++nlaunch;
    uint32_t ndom = 0;  // DOWN row launches of this batch
    const char *p32 = std::getenv("BFG_PUSH_U32");  // override: the u32 pushes alone
    const bool push16 = c->pr_w16.n > 0 && !(p32 && p32[0] == '1');
    for (int l = (int)nsl - 1; l >= 0; --l) {
        const uint32_t s0 = L.seglvl_seg[l], s1 = L.seglvl_seg[l + 1];
        for (uint32_t b = s0; b < s1; b += MAX_GRID_Y, ++nlaunch)
            hipLaunchKernelGGL((tri_kernel<KL, false, false, false>), dim3(wgx, std::min(s1 - b, MAX_GRID_Y)),
                               dim3(TILE_THREADS), 0, s, b, std::min(s1 - b, MAX_GRID_Y), c->seg_pos.p, c->tb_up.p,
                               c->tb_w.p, P.split.p, R, G0, w->arena.p, n, wgs);
        const uint32_t k0 = L.pt_blk[s0], k1 = L.pt_blk[s1];
        for (uint32_t b = k0; b < k1; b += MAX_GRID_Y, ++nlaunch)
            if (push16)
                hipLaunchKernelGGL(rows_push16_kernel<KL>, dim3(wgx, std::min(k1 - b, MAX_GRID_Y)),
                                   dim3(TILE_THREADS), 0, s, b, std::min(k1, b + MAX_GRID_Y), c->pk_seg.p,
                                   c->pr_off.p, c->pr_wid.p, c->pw_off.p, c->seg_pos.p, c->pt_tgt.p, c->pr_x.p,
                                   c->pr_w.p, c->pr_w16.p, R, w->arena.p, n, wgs);
            else
                hipLaunchKernelGGL((rows_kernel<KL, true, false>), dim3(wgx, std::min(k1 - b, MAX_GRID_Y)),
                                   dim3(TILE_THREADS), 0, s, b, std::min(k1, b + MAX_GRID_Y), c->pk_seg.p,
                                   c->pr_off.p, c->pr_off.p, c->pr_wid.p, c->pw_off.p, c->seg_pos.p, c->pt_tgt.p,
                                   c->pr_x.p, c->pr_w.p, R, G0, w->arena.p, n, wgs);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(w->ev[1], s));
    for (uint32_t l = 0; l < nsl; ++l) {
        const uint32_t u0 = P.lvl[l], u1 = P.lvl[l + 1];
        if (u1 > u0) {
            const bool timed = w->time_dom;
            if (timed && w->dom_ev.size() < 2 * (size_t)(ndom + 1)) {
                for (int k = 0; k < 2; ++k) {
                    hipEvent_t e;
                    HIP_TRY(hipEventCreate(&e));
                    w->dom_ev.push_back(e);
                }
            }
            if (timed) HIP_TRY(hipEventRecord(w->dom_ev[2 * ndom], s));
            const uint32_t pairs = (wgs + 1) / 2, px = (pairs + WPB - 1) / WPB;  // the 16-bit kernel's wave-group pairs
            for (uint32_t b = u0; b < u1; b += MAX_GRID_Y, ++nlaunch) {
                const dim3 grid(wgx, std::min(u1 - b, MAX_GRID_Y)), grid16(px, std::min(u1 - b, MAX_GRID_Y));
                if (c->label_bits == 16 && w->gate)
                    hipLaunchKernelGGL((rows16_kernel<KL, true>), grid16, dim3(TILE_THREADS), 0, s, b,
                                       std::min(u1, b + MAX_GRID_Y), P.seg.p, P.r0.p, P.r1.p, P.w.p, c->seg_pos.p,
                                       c->xr_src.p, c->xr_cap.p, c->xr_w16.p, c->xr_w.p, R, G, w->arena.p, n, wgs);
                else if (c->label_bits == 16)
                    hipLaunchKernelGGL((rows16_kernel<KL, false>), grid16, dim3(TILE_THREADS), 0, s, b,
                                       std::min(u1, b + MAX_GRID_Y), P.seg.p, P.r0.p, P.r1.p, P.w.p, c->seg_pos.p,
                                       c->xr_src.p, c->xr_cap.p, c->xr_w16.p, c->xr_w.p, R, G, w->arena.p, n, wgs);
                else if (w->gate)
                    hipLaunchKernelGGL((rows_kernel<KL, false, true>), grid, dim3(TILE_THREADS), 0, s, b,
                                       std::min(u1, b + MAX_GRID_Y), P.seg.p, P.r0.p, P.r1.p, nullptr, P.w.p,
                                       c->seg_pos.p, c->pt_tgt.p, c->xr_src.p, c->xr_w.p, R, G, w->arena.p, n, wgs);
                else
                    hipLaunchKernelGGL((rows_kernel<KL, false, false>), grid, dim3(TILE_THREADS), 0, s, b,
                                       std::min(u1, b + MAX_GRID_Y), P.seg.p, P.r0.p, P.r1.p, nullptr, P.w.p,
                                       c->seg_pos.p, c->pt_tgt.p, c->xr_src.p, c->xr_w.p, R, G, w->arena.p, n, wgs);
            }
            if (timed) {
                HIP_TRY(hipEventRecord(w->dom_ev[2 * ndom + 1], s));
                ++ndom;
            }
            pt->dom_bytes += (uint64_t)wgs * c->lvl_dom_bytes[l];
        }
        const uint32_t s0 = L.seglvl_seg[l], s1 = L.seglvl_seg[l + 1];
        for (uint32_t b = s0; b < s1; b += MAX_GRID_Y, ++nlaunch) {
            const dim3 grid(wgx, std::min(s1 - b, MAX_GRID_Y));
            const uint32_t cnt = std::min(s1 - b, MAX_GRID_Y);
            const bool sc = s1 - s0 > tss;
            if (w->gate && sc)
                hipLaunchKernelGGL((tri_kernel<KL, true, true, true>), grid, dim3(TILE_THREADS), 0, s, b, cnt,
                                   c->seg_pos.p, c->tb_dn.p, c->tb_w.p, P.split.p, R, G, w->arena.p, n, wgs);
            else if (w->gate)
                hipLaunchKernelGGL((tri_kernel<KL, true, false, true>), grid, dim3(TILE_THREADS), 0, s, b, cnt,
                                   c->seg_pos.p, c->tb_dn.p, c->tb_w.p, P.split.p, R, G, w->arena.p, n, wgs);
            else if (sc)
                hipLaunchKernelGGL((tri_kernel<KL, true, true, false>), grid, dim3(TILE_THREADS), 0, s, b, cnt,
                                   c->seg_pos.p, c->tb_dn.p, c->tb_w.p, P.split.p, R, G, w->arena.p, n, wgs);
            else
                hipLaunchKernelGGL((tri_kernel<KL, true, false, false>), grid, dim3(TILE_THREADS), 0, s, b, cnt,
                                   c->seg_pos.p, c->tb_dn.p, c->tb_w.p, P.split.p, R, G, w->arena.p, n, wgs);
        }
    }
    if (w->paths_on) {
        constexpr int NSP = NT / KL;
        if (w->n_paths) {
            const int rc = stage_path_elements(c, w);
            if (rc) return rc;
            const uint32_t gpx = (groups + XCDS - 1) / XCDS;
            hipLaunchKernelGGL(bottom_path_kernel<KL>, dim3((w->n_paths + NSP - 1) / NSP * gpx * XCDS), dim3(NT), 0,
                               s, w->n_paths, w->path_off.p, w->path_pos.p, L.n_top, c->in_off.p, c->in_split.p,
                               c->in_dd.p, c->in_src.p, c->in_w.p, c->pos_pre.p, c->pos_sub.p, c->pos_depth.p,
                               w->seed_off.p, w->seed_pre.p, w->seed_depth.p, w->pathbuf.p, D1, nl, w->arena.p, n);
            ++nlaunch;
        }
    } else {
        const uint32_t nch = (uint32_t)L.ch_off.size() - 1;
        if (nch) {
            hipLaunchKernelGGL((bottom_chunk_kernel<KL, CHUNK, NTB>), dim3(nch, groups), dim3(NTB), 0, s,
                               c->ch_off.p, c->ch_pos.p, c->cs_off.p, c->cs_src.p, c->cs_aoff.p, c->ca_v.p, c->ca_w.p,
                               c->pos_pre.p, c->pos_sub.p, c->pos_depth.p, w->seed_off.p, w->seed_pre.p,
                               w->seed_depth.p, w->pathbuf.p, D1, nl, CG, w->arena.p, n);
            ++nlaunch;
            if (w->gate && c->n_bt) {  // tree flags from the chunk flags
                hipLaunchKernelGGL(bottom_gate_kernel, dim3((c->n_bt + NT - 1) / NT, groups), dim3(NT), 0, s, c->n_bt,
                                   c->bt_ch.p, w->chunk_ok.p, nch, w->open_b.p);
                ++nlaunch;
            }
        }
        constexpr int NSP = NT / KL;
        for (size_t b = 0; b + 1 < L.botlvl_pos.size(); ++b) {
            const uint32_t p0 = L.botlvl_pos[b], p1 = L.botlvl_pos[b + 1];
            hipLaunchKernelGGL(bottom_pull_kernel<KL>, dim3((p1 - p0 + NSP - 1) / NSP, groups), dim3(NT), 0, s, p0,
                               p1, c->in_off.p, c->in_split.p, c->in_src.p, c->in_w.p, B, w->arena.p, n);
            ++nlaunch;
        }
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(w->ev[2], s));
    HIP_TRY(hipEventSynchronize(w->ev[2]));
    float a = 0, b = 0;
    HIP_TRY(hipEventElapsedTime(&a, w->ev[0], w->ev[1]));
    HIP_TRY(hipEventElapsedTime(&b, w->ev[1], w->ev[2]));
    pt->up_ms += a;
    pt->sweep_ms += b;
    pt->launches += nlaunch;
    for (uint32_t k = 0; k < ndom; ++k) {
        float d = 0;
        HIP_TRY(hipEventElapsedTime(&d, w->dom_ev[2 * k], w->dom_ev[2 * k + 1]));
        pt->dom_ms += d;
    }
    pt->dom_launches += ndom;
    return BFG_SUCCESS;
}

// One batch: lanes [0, nl) in sorted order, staged in the Work's host seed
// lists (h_soff, h_srank, h_scost, h_lim, h_rows); runs the upward and
// downward phases; the output step is done by the caller.
int run_batch(bfg_ctx *c, Work *w, const Plan &P, uint32_t nl, PhaseTimes *pt) {
    const size_t ns = w->h_srank.size();
    int rc = ensure_lane_buffers(c, w, nl, ns);
    if (rc) return rc;
    w->h_sdepth.resize(ns);
    w->h_spre.resize(ns);
    w->h_spos.resize(ns);
    for (size_t j = 0; j < ns; ++j) {
        const uint32_t r = w->h_srank[j];
        w->h_sdepth[j] = c->L.depth[r];
        w->h_spre[j] = c->L.pre[r];
        w->h_spos[j] = c->L.pos_of_rank[r];
    }
    hipStream_t s = w->stream;
    const StreamDrain drain{s};  // a no-op once launch_phases has synchronised
    HIP_TRY(hipMemcpyAsync(w->seed_off.p, w->h_soff.data(), (nl + 1) * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(w->lane_limit.p, w->h_lim.data(), nl * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(w->row_map.p, w->h_rows.data(), nl * 4, hipMemcpyHostToDevice, s));
    if (ns) {
        HIP_TRY(hipMemcpyAsync(w->seed_rank.p, w->h_srank.data(), ns * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(w->seed_pos.p, w->h_spos.data(), ns * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(w->seed_cost.p, w->h_scost.data(), ns * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(w->seed_depth.p, w->h_sdepth.data(), ns * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(w->seed_pre.p, w->h_spre.data(), ns * 4, hipMemcpyHostToDevice, s));
    }
    switch (c->lanes) {
        case 16: rc = launch_phases<16>(c, w, P, nl, pt); break;
        case 64: rc = launch_phases<64>(c, w, P, nl, pt); break;
        default: rc = launch_phases<32>(c, w, P, nl, pt); break;
    }
    return rc;  // launch_phases synchronised the stream: the staged vectors may change
}

// Gathers targets d_tpos[0 .. T) of a batch into rows (d_out, ld) and/or reach
// bits (d_bits, bits_ld), rows through d_rows (nullptr: sorted lane order).
// With a gate and dense labels, d_tpos holds positions and the bottom gate
// applies.
int gather_out(bfg_ctx *c, Work *w, uint32_t nl, const uint32_t *d_tpos, uint32_t T, const uint32_t *d_rows,
               uint32_t *d_out, uint64_t ld, unsigned long long *d_bits, uint64_t bits_ld, float *out_ms,
               uint32_t lane_lo = 0, uint32_t row_base = 0) {
    hipStream_t s = w->stream;
    // the lane groups holding lanes [lane_lo, nl)
    const uint32_t g0 = lane_lo / c->lanes, groups = (nl + c->lanes - 1) / c->lanes - g0;
    dim3 grid((T + 63) / 64, groups);
    BGate B = bgate_of(c, w);
    if (w->paths_on) B.open_b = nullptr;  // compact labels: dead paths hold INF
    HIP_TRY(hipEventRecord(w->ev[0], s));
    const uint32_t n = label_stride(c, w);
    switch (c->lanes) {
        case 16:
            hipLaunchKernelGGL(gather_kernel<16>, grid, dim3(NT), 0, s, w->arena.p, n, d_tpos, T, w->lane_limit.p,
                               nl, d_rows, B, d_out, ld, d_bits, bits_ld, g0, lane_lo, row_base);
            break;
        case 64:
            hipLaunchKernelGGL(gather_kernel<64>, grid, dim3(NT), 0, s, w->arena.p, n, d_tpos, T, w->lane_limit.p,
                               nl, d_rows, B, d_out, ld, d_bits, bits_ld, g0, lane_lo, row_base);
            break;
        default:
            hipLaunchKernelGGL(gather_kernel<32>, grid, dim3(NT), 0, s, w->arena.p, n, d_tpos, T, w->lane_limit.p,
                               nl, d_rows, B, d_out, ld, d_bits, bits_ld, g0, lane_lo, row_base);
            break;
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(w->ev[1], s));
    HIP_TRY(hipEventSynchronize(w->ev[1]));
    float a = 0;
    HIP_TRY(hipEventElapsedTime(&a, w->ev[0], w->ev[1]));
    *out_ms += a;
    return BFG_SUCCESS;
}

int check_ranks(const bfg_ctx *c, const uint32_t *r, size_t k, const char *what) {
    for (size_t i = 0; i < k; ++i)
        if (r[i] >= c->L.n) return fail(BFG_INVALID_PARAMETER, std::string(what) + " rank out of range");
    return BFG_SUCCESS;
}

// Restricts the bottom sweep of one call to the staged target paths (compact
// labels); always switched off again.
struct PathsOn {
    Work *w;
    explicit PathsOn(Work *ww) : w(ww) { w->paths_on = true; }
    ~PathsOn() { w->paths_on = false; }
};

// Limit classes of about 19 % width (a quarter of a binary order of
// magnitude): lanes of one class share wave groups when the gate is on.
uint32_t limit_class(uint32_t lim) {
    if (lim == INF) return 0xFFFFu;
    const uint32_t lg = 31 - (uint32_t)__builtin_clz(lim | 1u);  // floor(log2)
    const uint32_t frac = lg >= 2 ? (lim >> (lg - 2)) & 3u : (lim << (2 - lg)) & 3u;  // next two bits
    return lg * 4 + frac;
}

// Runs every lane of `req` on Work w: lanes are ordered by the pre-order of
// their first seed and cut into batches of the lane capacity; after each
// batch, out(batch_first, batch_lanes, rows, &ms) writes the outputs, where
// rows[i] is the original index of sorted lane batch_first + i.  `extra` is
// the device bytes per lane the caller's outputs take per batch.  A lane
// limit below INF turns the block gate on.
template <class F>
int run_all(bfg_ctx *c, Work *w, const Req &req, size_t extra, F &&out) {
    const uint32_t L = (uint32_t)req.limit.size();
    w->gate = false;
    for (uint32_t i = 0; i < L; ++i) w->gate |= req.limit[i] != INF;
    if (w->fixed) w->lane_cap = 64;
    else if (c->cap_fixed) w->lane_cap = c->lane_cap;
    else w->lane_cap = w->cap_hint ? w->cap_hint : auto_lane_cap(c, w, L, extra);
    int rc = ensure_arena(c, w, (size_t)w->lane_cap * label_stride(c, w));
    if (rc) return rc;
    if (w->gate) {
        const size_t wgs = w->lane_cap / 64, groups = (w->lane_cap + c->lanes - 1) / c->lanes;
        ENSURE(w, seg_live, wgs * std::max<size_t>(c->L.seg_pos.size() - 1, 1), "gate flags");
        ENSURE(w, chunk_ok, groups * std::max<size_t>(c->L.ch_off.size() - 1, 1), "gate flags");
        ENSURE(w, open_b, groups * std::max<uint32_t>(c->n_bt, 1), "gate flags");
    }
    // sort key: pre-order of the first seed, then the lane index
    auto &key = w->h_key;
    key.resize(L);
    for (uint32_t i = 0; i < L; ++i) {
        const uint64_t k = req.soff[i + 1] > req.soff[i] ? c->L.pre[req.srank[req.soff[i]]] : 0u;
        key[i] = k << 32 | i;
    }
    if (req.major.empty() && w->gate) {
        // bounded lanes by limit class first: the block gate skips a unit only
        // when every lane of its wave group is out of reach, and lanes of
        // similar limits from neighbouring origins reach similar units
        std::sort(key.begin(), key.end(), [&](uint64_t a, uint64_t b) {
            const uint32_t ca = limit_class(req.limit[(uint32_t)a]), cb = limit_class(req.limit[(uint32_t)b]);
            return ca != cb ? ca < cb : a < b;
        });
    } else if (req.major.empty()) {
        std::sort(key.begin(), key.end());
    } else {  // each request's lanes stay together, in the order of their requests
        const std::vector<uint32_t> &mj = req.major;
        std::sort(key.begin(), key.end(), [&](uint64_t a, uint64_t b) {
            const uint32_t ma = mj[(uint32_t)a], mb = mj[(uint32_t)b];
            return ma != mb ? ma < mb : a < b;
        });
    }
    auto t_start = std::chrono::steady_clock::now();
    bfg_timings tm{};
    PhaseTimes pt;
    for (uint32_t b0 = 0; b0 < L; b0 += w->lane_cap) {
        const uint32_t nl = std::min(w->lane_cap, L - b0);
        w->h_soff.resize(nl + 1);
        w->h_lim.resize(nl);
        w->h_rows.resize(nl);
        w->h_srank.clear();
        w->h_scost.clear();
        for (uint32_t i = 0; i < nl; ++i) {
            const uint32_t o = (uint32_t)key[b0 + i];
            w->h_rows[i] = o;
            w->h_soff[i] = (uint32_t)w->h_srank.size();
            for (uint32_t k = req.soff[o]; k < req.soff[o + 1]; ++k) {
                w->h_srank.push_back(req.srank[k]);
                w->h_scost.push_back(req.scost[k]);
            }
            w->h_lim[i] = req.limit[o];
        }
        w->h_soff[nl] = (uint32_t)w->h_srank.size();
        static const Plan no_plan;  // an upward-only batch needs no DOWN plan
        std::shared_ptr<const Plan> P;
        if (!w->up_only) {
            rc = get_plan(c, w, (nl + 63) / 64, P);
            if (rc) return rc;
        }
        rc = run_batch(c, w, P ? *P : no_plan, nl, &pt);
        if (rc) return rc;
        if (w->gate && !w->fixed) {  // how much the gate left open (flags of this batch)
            const size_t wgs = (nl + 63) / 64, groups = (nl + c->lanes - 1) / c->lanes;
            const size_t nseg = c->L.seg_pos.size() - 1;
            std::vector<uint8_t> f(std::max(wgs * nseg, groups * (size_t)c->n_bt));
            HIP_TRY(hipMemcpy(f.data(), w->seg_live.p, wgs * nseg, hipMemcpyDeviceToHost));
            tm.gate_seg_units += wgs * nseg;
            for (size_t k = 0; k < wgs * nseg; ++k) tm.gate_seg_open += f[k];
            if (c->n_bt && !w->paths_on) {
                HIP_TRY(hipMemcpy(f.data(), w->open_b.p, groups * c->n_bt, hipMemcpyDeviceToHost));
                tm.gate_bt_units += groups * c->n_bt;
                for (size_t k = 0; k < groups * c->n_bt; ++k) tm.gate_bt_open += f[k];
            }
        }
        rc = out(b0, nl, w->h_rows, &tm.output_ms);
        if (rc) return rc;
        tm.batches++;
    }
    tm.up_ms = pt.up_ms;
    tm.sweep_ms = pt.sweep_ms;
    tm.sweep_launches = pt.launches;
    tm.dom_ms = pt.dom_ms;
    tm.dom_launches = pt.dom_launches;
    tm.dom_bytes = pt.dom_bytes;
    tm.lanes_run = L;
    tm.nodes_swept = w->paths_on ? c->L.n_top + w->path_nodes : c->L.n;
    tm.arcs_swept = w->paths_on ? c->top_in_arcs + w->path_arcs : c->L.m_down;
    tm.total_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    w->last = tm;
    std::lock_guard<std::mutex> g(c->last_mu);
    c->last = tm;
    return BFG_SUCCESS;
}

void single_source_req(Req &r, const uint32_t *src, uint32_t k, uint32_t threshold) {
    r.soff.resize(k + 1);
    r.srank.assign(src, src + k);
    r.scost.assign(k, 0);
    r.limit.assign(k, threshold);
    for (uint32_t i = 0; i <= k; ++i) r.soff[i] = i;
}

// Bottom paths of a matrix's targets: for every bottom target that is not an
// ancestor of another target, its bottom ancestors top-down (the top-region
// boundary first, the target last).  Stages the path offsets (a path's
// length is its target's bdep + 1) and the targets' label indices (tgt_pos):
// top targets by position, bottom targets by their element on a path
// (compact labels, see label_stride).  The elements themselves are written
// by stage_path_elements once the top-region kernels are queued.
int stage_target_paths(bfg_ctx *c, Work *w, const uint32_t *targets, uint32_t T) {
    const bfg::Layout &L = c->L;
    std::vector<uint32_t> &bt = w->h_bt, &pending = w->h_pending;
    bt.clear();
    pending.clear();
    for (uint32_t t = 0; t < T; ++t) {
        const uint32_t p = L.pos_of_rank[targets[t]];
        if (p >= L.n_top) bt.push_back(p);
    }
    std::sort(bt.begin(), bt.end(), [&](uint32_t a, uint32_t b) { return L.pos_pre[a] < L.pos_pre[b]; });
    bt.erase(std::unique(bt.begin(), bt.end()), bt.end());
    std::vector<uint32_t> &off = w->path_off_h, &leaf = w->path_leaf;
    off.assign(1, 0);
    leaf.clear();
    auto &elem = w->h_elem;  // (bottom target position, path element)
    elem.clear();
    for (size_t i = 0; i < bt.size(); ++i) {
        const uint32_t t = bt[i];
        pending.push_back(t);
        // the next target in pre-order lies in t's subtree iff t is an
        // ancestor of some later target: t is then on that target's path
        // (and so are the pending targets before it, its ancestors)
        if (i + 1 < bt.size() && L.pos_pre[bt[i + 1]] - L.pos_pre[t] < L.pos_subtree[t]) continue;
        const uint32_t e0 = off.back(), d0 = L.pos_depth[t] - c->bdep[t];  // depth of the path's first element
        off.push_back(e0 + c->bdep[t] + 1);
        leaf.push_back(t);
        for (uint32_t q : pending) elem.emplace_back(q, e0 + L.pos_depth[q] - d0);
        pending.clear();
    }
    std::sort(elem.begin(), elem.end());
    std::vector<uint32_t> &tp = w->tgt_pos_h;
    tp.resize(T);
    for (uint32_t t = 0; t < T; ++t) {
        const uint32_t p = L.pos_of_rank[targets[t]];
        if (p < L.n_top) {
            tp[t] = p;
        } else {
            auto it = std::lower_bound(elem.begin(), elem.end(), std::make_pair(p, 0u));
            if (it == elem.end() || it->first != p) return fail(BFG_INTERNAL_ERROR, "target missing from its path");
            tp[t] = L.n_top + it->second;
        }
    }
    ENSURE(w, path_off, off.size(), "target paths");
    ENSURE(w, path_pos, std::max<size_t>(off.back(), 1), "target paths");
    ENSURE(w, tgt_pos, T, "targets");
    // the host tables live in the Work until its next call, which starts
    // after this one has synchronised
    HIP_TRY(hipMemcpyAsync(w->path_off.p, off.data(), off.size() * 4, hipMemcpyHostToDevice, w->stream));
    HIP_TRY(hipMemcpyAsync(w->tgt_pos.p, tp.data(), (size_t)T * 4, hipMemcpyHostToDevice, w->stream));
    w->n_paths = (uint32_t)leaf.size();
    w->path_nodes = off.back();
    w->path_arcs = 0;
    w->path_pos_staged = false;
    return BFG_SUCCESS;
}

// The elements of the staged paths (a walk up from each path's target) and
// their DOWN in-arc count, uploaded on the stream.  Called with the
// top-region kernels already queued, so the walk overlaps them.
int stage_path_elements(bfg_ctx *c, Work *w) {
    if (w->path_pos_staged) return BFG_SUCCESS;
    const bfg::Layout &L = c->L;
    std::vector<uint32_t> &pos = w->path_pos_h;
    pos.resize(w->path_nodes);
    uint64_t arcs = 0;
    for (uint32_t k = 0; k < w->n_paths; ++k) {
        uint32_t x = w->path_leaf[k];
        for (uint32_t i = w->path_off_h[k + 1]; i-- > w->path_off_h[k];) {
            if (x == NONE || x < L.n_top) return fail(BFG_INTERNAL_ERROR, "bottom path shorter than its depth");
            pos[i] = x;
            arcs += L.in_off[x + 1] - L.in_off[x];
            x = L.pos_parent[x];
        }
        if (x != NONE && x >= L.n_top) return fail(BFG_INTERNAL_ERROR, "bottom path longer than its depth");
    }
    if (!pos.empty())
        HIP_TRY(hipMemcpyAsync(w->path_pos.p, pos.data(), pos.size() * 4, hipMemcpyHostToDevice, w->stream));
    w->path_arcs = arcs;
    w->path_pos_staged = true;
    return BFG_SUCCESS;
}

// Full rows for all ranks (dist and/or reach bits) into host buffers indexed
// by original lane.  Distances go through a device buffer of one batch's rows
// (counted in the batch's lane capacity) straight into dist_host.
int dense_rows(bfg_ctx *c, Work *w, const Req &req, uint32_t *dist_host, uint64_t *bits_host) {
    const uint32_t n = c->L.n;
    const uint32_t L = (uint32_t)req.limit.size();
    const uint64_t words = (n + 63) / 64;
    DevBuf<uint32_t> d_rows;
    DevBuf<unsigned long long> d_bits;
    const StreamDrain drain{w->stream};
    // reach bits of every lane land in request order on the device (rows
    // through row_map) and come back in one copy
    if (bits_host) HIP_TRY(d_bits.alloc((size_t)L * words));
    const size_t extra = dist_host ? (size_t)n * 4 : 0;
    int rc = run_all(c, w, req, extra, [&](uint32_t, uint32_t nl, const std::vector<uint32_t> &rows, float *oms) -> int {
        if (dist_host) {  // this batch's rows in sorted lane order, each copied to its request row
            HIP_TRY(d_rows.ensure((size_t)nl * n));
            int rc2 = gather_out(c, w, nl, c->pos_of_rank_d.p, n, nullptr, d_rows.p, n, nullptr, 0, oms);
            if (rc2) return rc2;
            for (uint32_t i = 0; i < nl; ++i)
                HIP_TRY(hipMemcpyAsync(dist_host + (size_t)rows[i] * n, d_rows.p + (size_t)i * n, (size_t)n * 4,
                                       hipMemcpyDeviceToHost, w->stream));
            HIP_TRY(hipStreamSynchronize(w->stream));
        }
        if (bits_host) {
            int rc2 = gather_out(c, w, nl, c->pos_of_rank_d.p, n, w->row_map.p, nullptr, n, d_bits.p, words, oms);
            if (rc2) return rc2;
        }
        return BFG_SUCCESS;
    });
    if (rc) return rc;
    if (bits_host) HIP_TRY(hipMemcpy(bits_host, d_bits.p, (size_t)L * words * 8, hipMemcpyDeviceToHost));
    return BFG_SUCCESS;
}

// Isochrone lists of every lane of `req`: reach bits (into d_bits_ext when
// given, else a buffer of the call; copied to bits_host when given), then per
// lane its settled (rank, time) pairs in ascending rank.  Lane i's list is
// pairs[start[i] .. start[i] + len[i]); lists are laid out batch by batch,
// by original lane order within a batch.  Pairs go to d_pairs_ext (device,
// `capacity` pairs) or to pairs_host (`capacity` pairs).  If the lists need
// more than capacity, len and start are still filled (every batch is
// counted) and BFG_BUFFER_TOO_SMALL is returned with no pair written past
// what fit.
int iso_lists(bfg_ctx *c, Work *w, const Req &req, uint64_t *bits_host, uint64_t *d_bits_ext, uint64_t *start,
              uint32_t *len, uint32_t *pairs_host, uint32_t *d_pairs_ext, uint64_t capacity, uint64_t *total_out) {
    const uint32_t n = c->L.n;
    const uint32_t L = (uint32_t)req.limit.size();
    const uint64_t words = (n + 63) / 64;
    const uint32_t nsb = (n + SB_RANKS - 1) / SB_RANKS;
    DevBuf<unsigned long long> d_bits_own;
    DevBuf<uint32_t> d_sorted, d_slot, d_sboff, d_rowtot;
    DevBuf<uint64_t> d_rowstart, d_total;
    DevBuf<uint2> d_pairs_own;
    hipStream_t s = w->stream;
    const StreamDrain drain{s};
    unsigned long long *d_bits = (unsigned long long *)d_bits_ext;
    if (!d_bits) {
        HIP_TRY(d_bits_own.alloc((size_t)L * words));
        d_bits = d_bits_own.p;
    }
    HIP_TRY(d_total.alloc(1));
    uint64_t running = 0;
    bool too_small = false;
    std::vector<uint32_t> sorted, slot, rowtot;
    const size_t extra = (size_t)nsb * 4 + 32;
    int rc = run_all(c, w, req, extra, [&](uint32_t, uint32_t nl, const std::vector<uint32_t> &rows, float *oms) -> int {
        int rc2 = gather_out(c, w, nl, c->pos_of_rank_d.p, n, w->row_map.p, nullptr, n, d_bits, words, oms);
        if (rc2) return rc2;
        // the batch's rows in original order, and each sorted lane's slot among them
        sorted.assign(rows.begin(), rows.end());
        std::sort(sorted.begin(), sorted.end());
        slot.resize(nl);
        for (uint32_t i = 0; i < nl; ++i)
            slot[i] = (uint32_t)(std::lower_bound(sorted.begin(), sorted.end(), rows[i]) - sorted.begin());
        HIP_TRY(d_sorted.ensure(nl));
        HIP_TRY(d_slot.ensure(nl));
        HIP_TRY(d_sboff.ensure((size_t)nl * nsb));
        HIP_TRY(d_rowtot.ensure(nl));
        HIP_TRY(d_rowstart.ensure(nl));
        HIP_TRY(hipMemcpyAsync(d_sorted.p, sorted.data(), nl * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(d_slot.p, slot.data(), nl * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipEventRecord(w->ev[0], s));
        iso_count_kernel<<<nl, ISO_THREADS, 0, s>>>(d_bits, words, words, d_sorted.p, nsb, d_sboff.p, d_rowtot.p);
        // host pairs go through a buffer of this batch's pairs (base 0)
        const uint64_t base = d_pairs_ext ? running : 0;
        iso_scan_kernel<<<1, ISO_THREADS, 0, s>>>(d_rowtot.p, nl, base, d_rowstart.p, d_total.p);
        HIP_TRY(hipGetLastError());
        rowtot.resize(nl);
        HIP_TRY(hipMemcpyAsync(rowtot.data(), d_rowtot.p, nl * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        uint64_t batch = 0;
        for (uint32_t i = 0; i < nl; ++i) {
            start[sorted[i]] = running + batch;
            len[sorted[i]] = rowtot[i];
            batch += rowtot[i];
        }
        if (too_small || running + batch > capacity) {
            too_small = true;
            running += batch;
            return BFG_SUCCESS;
        }
        uint2 *dst = (uint2 *)d_pairs_ext;
        if (!dst) {
            HIP_TRY(d_pairs_own.ensure(std::max<uint64_t>(batch, 1)));
            dst = d_pairs_own.p;
        }
        const uint32_t groups = (nl + c->lanes - 1) / c->lanes;
        const dim3 grid(nsb, groups);
        switch (c->lanes) {
            case 16:
                iso_emit_kernel<16><<<grid, NT, 0, s>>>(w->arena.p, n, c->pos_of_rank_d.p, n, d_bits, words,
                                                        w->row_map.p, nl, d_slot.p, d_rowstart.p, d_sboff.p, nsb,
                                                        dst);
                break;
            case 64:
                iso_emit_kernel<64><<<grid, NT, 0, s>>>(w->arena.p, n, c->pos_of_rank_d.p, n, d_bits, words,
                                                        w->row_map.p, nl, d_slot.p, d_rowstart.p, d_sboff.p, nsb,
                                                        dst);
                break;
            default:
                iso_emit_kernel<32><<<grid, NT, 0, s>>>(w->arena.p, n, c->pos_of_rank_d.p, n, d_bits, words,
                                                        w->row_map.p, nl, d_slot.p, d_rowstart.p, d_sboff.p, nsb,
                                                        dst);
                break;
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(w->ev[1], s));
        if (pairs_host && batch)
            HIP_TRY(hipMemcpyAsync(pairs_host + 2 * running, dst, batch * 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        float a = 0;
        HIP_TRY(hipEventElapsedTime(&a, w->ev[0], w->ev[1]));
        *oms += a;
        running += batch;
        return BFG_SUCCESS;
    });
    if (rc) return rc;
    if (total_out) *total_out = running;
    if (bits_host) HIP_TRY(hipMemcpy(bits_host, d_bits, (size_t)L * words * 8, hipMemcpyDeviceToHost));
    if (too_small)
        return fail(BFG_BUFFER_TOO_SMALL, "isochrone lists need " + std::to_string(running) + " pairs, capacity " +
                                              std::to_string(capacity));
    return BFG_SUCCESS;
}

// ---------------------------------------------------------------------------
// Two-channel seeded PHAST (time, length along the time-optimal path), the
// semantics of run_phast_bounded_fast_seeded_2ch (isochrone_handler.rs:500-625).
// A label is (time << 32 | length): lexicographic order is integer order, so
// the reference's "strictly faster, or equally fast and strictly shorter"
// relaxation (:543-567, :584-602) is a min, and a node's final label is the
// least (time, length) over the up-down paths from the lane's seeds.  Lanes
// are grouped by 64; labels live in [group][position][64 lanes].  A lane
// with limit T relaxes only from labels of time <= T; every label of time
// <= T is exact, the rest are masked to INF on output.
// ---------------------------------------------------------------------------
constexpr uint64_t INF2 = ~0ull;
constexpr int T2_SEG_WAVES = 16;  // wavefronts sharing one segment tile
// Levels whose segments have at most this many outside arcs run 8-wavefront
// tiles: their time goes to wavefront 0's chain walk, and three such tiles
// fit a CU (two of 16 wavefronts).
constexpr uint32_t T2_SMALL_ARCS = 4096;

uint32_t t2_small_arcs() {
    if (const char *e = std::getenv("BFG_T2_SMALL_ARCS")) {  // tuning override
        const long v = std::strtol(e, nullptr, 10);
        if (v >= 0 && v < (1L << 31)) return (uint32_t)v;
    }
    return T2_SMALL_ARCS;
}

__device__ __forceinline__ uint64_t add2(uint64_t a, uint64_t w) {
    const uint32_t t = sat_add((uint32_t)(a >> 32), (uint32_t)(w >> 32));
    const uint32_t l = sat_add((uint32_t)a, (uint32_t)w);
    return (uint64_t)t << 32 | l;
}

// a label relaxes its arcs only when reached within the lane's limit
__device__ __forceinline__ bool live2(uint64_t x, uint32_t thr) {
    const uint32_t t = (uint32_t)(x >> 32);
    return t != INF && t <= thr;
}

__global__ __launch_bounds__(NT) void t2_fill_kernel(uint64_t *__restrict__ p, size_t count) {
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < count; i += (size_t)gridDim.x * NT) p[i] = INF2;
}

// One wavefront per lane: the lane's seeds (distinct positions), then its
// upward phase (isochrone_handler.rs:527-568) as a walk over the union of the
// seeds' root paths in descending position (an ancestor always has a smaller
// position than its descendants, so every push into a node precedes its own
// pushes).  wk[k] is the walker of seed k; walkers meeting at a node merge.
// The wavefront's threads split a node's UP arcs (their heads are distinct);
// a barrier orders one node's pushes before the next node's label is read.
// Top-region labels go to the lane's own row of `top` (a node's ancestors in
// one segment are 64 adjacent words there, so a node's pushes touch few
// lines and the walk's labels stay in cache); bottom labels to the arena.
__global__ __launch_bounds__(64) void t2_up_kernel(uint32_t nl, uint32_t n, uint32_t n_top,
                                                   const uint32_t *__restrict__ s_off,
                                                   const uint32_t *__restrict__ s_pos,
                                                   const uint64_t *__restrict__ s_lab,
                                                   const uint32_t *__restrict__ lim,
                                                   const uint32_t *__restrict__ pos_parent,
                                                   const uint32_t *__restrict__ u_off,
                                                   const uint32_t *__restrict__ u_tgt,
                                                   const uint64_t *__restrict__ u_w, uint64_t *__restrict__ lab,
                                                   uint64_t *__restrict__ top) {
    __shared__ uint32_t wk[64];
    const uint32_t i = blockIdx.x, tid = threadIdx.x;
    if (i >= nl) return;
    uint64_t *col = lab + (size_t)(i / 64) * n * 64 + (i % 64);
    uint64_t *row = top + (size_t)i * n_top;
    auto at = [&](uint32_t p) -> uint64_t & { return p < n_top ? row[p] : col[(size_t)p * 64]; };
    const uint32_t a = s_off[i], ns = s_off[i + 1] - a, thr = lim[i];
    if (tid < ns) {  // at most 64 seeds per lane (checked on the host)
        at(s_pos[a + tid]) = s_lab[a + tid];
        wk[tid] = s_pos[a + tid];
    }
    __syncthreads();
    // the same for every thread: all read the same walkers
    auto next = [&](uint32_t at_u, uint32_t to) {  // the walk's next node once walkers at at_u move to `to`
        uint32_t m = NONE;
        for (uint32_t k = 0; k < ns; ++k) {
            uint32_t ck = wk[k];
            if (ck == at_u) ck = to;
            if (ck != NONE && (m == NONE || ck > m)) m = ck;
        }
        return m;
    };
    uint32_t u = next(NONE, NONE);
    if (u == NONE) return;
    uint32_t par = pos_parent[u], e0 = u_off[u], e1 = u_off[u + 1];
    constexpr uint32_t U = 8;  // arcs per thread and step, their loads in flight together
    while (u != NONE) {
        // the next node and its tables do not depend on labels: load them now
        const uint32_t un = next(u, par);
        uint32_t parn = NONE, e0n = 0, e1n = 0;
        if (un != NONE) {
            parn = pos_parent[un];
            e0n = u_off[un];
            e1n = u_off[un + 1];
        }
        for (uint32_t e = e0 + tid; e < e1; e += 64 * U) {
            uint32_t v[U];
            uint64_t w[U], x[U];
#pragma unroll
            for (uint32_t q = 0; q < U; ++q) {
                const uint32_t f = e + 64 * q;
                v[q] = f < e1 ? u_tgt[f] : NONE;
                w[q] = f < e1 ? u_w[f] : INF2;
            }
            const uint64_t lu = at(u);
            if (!live2(lu, thr)) break;
#pragma unroll
            for (uint32_t q = 0; q < U; ++q) x[q] = v[q] != NONE ? at(v[q]) : INF2;
#pragma unroll
            for (uint32_t q = 0; q < U; ++q)
                if (v[q] != NONE) at(v[q]) = min(x[q], add2(lu, w[q]));
        }
        __syncthreads();
        if (tid < ns && wk[tid] == u) wk[tid] = par;
        __syncthreads();
        u = un;
        par = parn;
        e0 = e0n;
        e1 = e1n;
    }
}

// The upward phase's top-region rows into the arena: a workgroup per (64 top
// positions, lane group), transposed through LDS.
__global__ __launch_bounds__(NT) void t2_top_kernel(uint32_t n, uint32_t n_top, const uint64_t *__restrict__ top,
                                                    uint64_t *__restrict__ lab) {
    __shared__ uint64_t tile[64][65];
    const uint32_t p0 = blockIdx.x * 64, g = blockIdx.y;
    for (uint32_t idx = threadIdx.x; idx < 64 * 64; idx += NT) {
        const uint32_t l = idx / 64, j = idx % 64;
        if (p0 + j < n_top) tile[j][l] = top[(size_t)(g * 64 + l) * n_top + p0 + j];
    }
    __syncthreads();
    uint64_t *base = lab + (size_t)g * n * 64;
    for (uint32_t idx = threadIdx.x; idx < 64 * 64; idx += NT) {
        const uint32_t j = idx / 64, l = idx % 64;
        if (p0 + j < n_top) base[(size_t)(p0 + j) * 64 + l] = tile[j][l];
    }
}

// Downward phase over one segment level of the top region: a workgroup per
// (segment, lane group).  Its NW wavefronts split the segment's arcs from
// outside (grouped by source, so a source label is loaded once per run) and
// merge into the LDS tile with atomic min; wavefront 0 then walks the chain,
// parent first, adding the arcs inside the segment.
template <int NW>
__global__ __launch_bounds__(NW * 64) void t2_seg_kernel(uint32_t s0, uint32_t n, uint32_t nl,
                                                    const uint32_t *__restrict__ seg_pos,
                                                    const uint32_t *__restrict__ x_off,
                                                    const uint32_t *__restrict__ x_sj,
                                                    const uint64_t *__restrict__ x_w,
                                                    const uint32_t *__restrict__ i_off,
                                                    const uint32_t *__restrict__ i_k,
                                                    const uint64_t *__restrict__ i_w,
                                                    const uint32_t *__restrict__ lim, uint64_t *__restrict__ lab) {
    constexpr uint32_t MAXI = SEG * (SEG - 1) / 2;  // arcs inside a segment (one per ordered pair)
    __shared__ uint64_t acc[SEG][64];
    __shared__ uint64_t iw[MAXI];
    __shared__ uint32_t io[SEG + 1];
    __shared__ uint8_t ik[MAXI];  // tail offsets (< 64): 51 KB of LDS in all, three tiles per CU
    const uint32_t s = s0 + blockIdx.x, g = blockIdx.y, lane = threadIdx.x & 63;
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t gl = g * 64 + lane, thr = gl < nl ? lim[gl] : 0;
    const uint32_t p0 = seg_pos[s], len = seg_pos[s + 1] - p0;
    uint64_t *base = lab + (size_t)g * n * 64 + lane;
    for (uint32_t j = w; j < len; j += NW) acc[j][lane] = base[(size_t)(p0 + j) * 64];
    // the chain's arcs into LDS, so wavefront 0 walks it without memory waits
    const uint32_t i0 = i_off[p0], ni = i_off[p0 + len] - i0;
    for (uint32_t k = threadIdx.x; k < ni; k += NW * 64) {
        ik[k] = (uint8_t)i_k[i0 + k];
        iw[k] = i_w[i0 + k];
    }
    for (uint32_t k = threadIdx.x; k <= len; k += NW * 64) io[k] = i_off[p0 + k] - i0;
    __syncthreads();
    const uint32_t a0 = x_off[s], a1 = x_off[s + 1], per = (a1 - a0 + NW - 1) / NW;
    const uint32_t e0 = min(a1, a0 + w * per), e1 = min(a1, e0 + per);
    // 64 arcs' words per vector load, broadcast by v_readlane; eight arcs per
    // step with their label loads in flight together (arcs of one source hit
    // the same line)
    constexpr uint32_t U = 8;
    for (uint32_t c0 = e0; c0 < e1; c0 += 64) {
        const uint32_t cl = min(64u, e1 - c0);
        const uint32_t my_sj = lane < cl ? x_sj[c0 + lane] : 0u;
        const uint64_t my_w = lane < cl ? x_w[c0 + lane] : INF2;
        const uint32_t my_wt = (uint32_t)(my_w >> 32), my_wl = (uint32_t)my_w;
        uint32_t j = 0;
        for (; j + U <= cl; j += U) {
            uint32_t sj[U];
            uint64_t ww[U], ls[U];
#pragma unroll
            for (uint32_t q = 0; q < U; ++q) {
                sj[q] = rl(my_sj, j + q);
                ww[q] = (uint64_t)rl(my_wt, j + q) << 32 | rl(my_wl, j + q);
            }
            if ((sj[0] >> 6) == (sj[U - 1] >> 6)) {  // one source (arcs come grouped by source): one load
                const uint64_t l1 = base[(size_t)(sj[0] >> 6) * 64];
                if (live2(l1, thr)) {
#pragma unroll
                    for (uint32_t q = 0; q < U; ++q)
                        atomicMin((unsigned long long *)&acc[sj[q] & 63][lane], (unsigned long long)add2(l1, ww[q]));
                }
                continue;
            }
#pragma unroll
            for (uint32_t q = 0; q < U; ++q) ls[q] = base[(size_t)(sj[q] >> 6) * 64];
#pragma unroll
            for (uint32_t q = 0; q < U; ++q)
                if (live2(ls[q], thr))
                    atomicMin((unsigned long long *)&acc[sj[q] & 63][lane], (unsigned long long)add2(ls[q], ww[q]));
        }
        for (; j < cl; ++j) {
            const uint32_t sj = rl(my_sj, j);
            const uint64_t ww = (uint64_t)rl(my_wt, j) << 32 | rl(my_wl, j);
            const uint64_t ls = base[(size_t)(sj >> 6) * 64];
            if (live2(ls, thr)) atomicMin((unsigned long long *)&acc[sj & 63][lane], (unsigned long long)add2(ls, ww));
        }
    }
    __syncthreads();
    if (w != 0) return;
    for (uint32_t i = 0; i < len; ++i) {
        uint64_t v = acc[i][lane];
        const uint32_t f1 = io[i + 1];
        uint32_t f = io[i];
        for (; f + 4 <= f1; f += 4) {
            uint64_t lk[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) lk[q] = acc[ik[f + q]][lane];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (live2(lk[q], thr)) v = min(v, add2(lk[q], iw[f + q]));
        }
        for (; f < f1; ++f) {
            const uint64_t lk = acc[ik[f]][lane];
            if (live2(lk, thr)) v = min(v, add2(lk, iw[f]));
        }
        acc[i][lane] = v;
        base[(size_t)(p0 + i) * 64] = v;
    }
}

// Downward phase over one height level of the bottom region: a wavefront per
// (node, lane group) pulls the node's DOWN in-arcs.  The level's nodes are
// positions [p0, p1), or plist[p0 .. p1) when a table needs only the bottom
// nodes above its target seeds.
__global__ __launch_bounds__(NT) void t2_pull_kernel(uint32_t p0, uint32_t p1, const uint32_t *__restrict__ plist,
                                                     uint32_t n, uint32_t n_top, uint32_t nl,
                                                     const uint32_t *__restrict__ b_off,
                                                     const uint32_t *__restrict__ b_src,
                                                     const uint64_t *__restrict__ b_w,
                                                     const uint32_t *__restrict__ lim, uint64_t *__restrict__ lab) {
    const uint32_t k = p0 + blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    if (k >= p1) return;
    const uint32_t p = plist ? plist[k] : k;
    const uint32_t lane = threadIdx.x & 63, g = blockIdx.y, gl = g * 64 + lane, thr = gl < nl ? lim[gl] : 0;
    uint64_t *base = lab + (size_t)g * n * 64 + lane;
    uint64_t v = base[(size_t)p * 64];
    for (uint32_t e = b_off[p - n_top]; e < b_off[p - n_top + 1]; ++e) {
        const uint64_t ls = base[(size_t)b_src[e] * 64];
        if (live2(ls, thr)) v = min(v, add2(ls, b_w[e]));
    }
    base[(size_t)p * 64] = v;
}

// Output rows by rank: a workgroup per (64 ranks, lane group) transposes
// through LDS; labels of time above the lane's limit (or unreached) are INF
// in both channels (isochrone_handler.rs:605-622 keeps time <= threshold).
__global__ __launch_bounds__(NT) void t2_out_kernel(uint32_t n, uint32_t nl, const uint32_t *__restrict__ pos_of_rank,
                                                    const uint32_t *__restrict__ lim, const uint64_t *__restrict__ lab,
                                                    uint32_t *__restrict__ o_t, uint32_t *__restrict__ o_l) {
    __shared__ uint64_t tile[64][65];
    const uint32_t v0 = blockIdx.x * 64, g = blockIdx.y, t = threadIdx.x;
    const uint64_t *base = lab + (size_t)g * n * 64;
    for (uint32_t r = t >> 6; r < 64; r += NT / 64)
        if (v0 + r < n) tile[r][t & 63] = base[(size_t)pos_of_rank[v0 + r] * 64 + (t & 63)];
    __syncthreads();
    const uint32_t r = t & 63;
    if (v0 + r >= n) return;
    for (uint32_t l = t >> 6; l < 64; l += NT / 64) {
        const uint32_t gl = g * 64 + l;
        if (gl >= nl) break;
        const uint64_t x = tile[r][l];
        const uint32_t tm = (uint32_t)(x >> 32);
        const bool ok = tm != INF && tm <= lim[gl];
        o_t[(size_t)gl * n + v0 + r] = ok ? tm : INF;
        o_l[(size_t)gl * n + v0 + r] = ok ? (uint32_t)x : INF;
    }
}

// Two-channel seeded table (bucket_ch.rs:6454-6481): a workgroup per (64
// targets, lane group).  Target j's cell for source lane i takes, over j's
// seeds (rank, part_time, part_len) in order, the first seed of least
// field time - part_time (saturating) whose rank the field settled (time <=
// the lane's field bound), and that seed's field length - part_len; above
// `thr` both are INF.  A target sharing a seed rank with the source is a
// conflict: flagged, cells INF.
__global__ __launch_bounds__(NT) void t2_table_kernel(uint32_t n, uint32_t T, uint32_t nl,
                                                      const uint32_t *__restrict__ toff,
                                                      const uint32_t *__restrict__ tpos,
                                                      const uint32_t *__restrict__ trank,
                                                      const uint32_t *__restrict__ tpt,
                                                      const uint32_t *__restrict__ tpl,
                                                      const uint32_t *__restrict__ soff,
                                                      const uint32_t *__restrict__ srank,
                                                      const uint32_t *__restrict__ lim, uint32_t thr,
                                                      const uint64_t *__restrict__ lab, uint32_t *__restrict__ o_t,
                                                      uint32_t *__restrict__ o_l, uint8_t *__restrict__ conf) {
    __shared__ uint32_t tt[64][65], tl[64][65];
    const uint32_t t0 = blockIdx.x * 64, g = blockIdx.y;
    const uint64_t *base = lab + (size_t)g * n * 64;
    for (uint32_t idx = threadIdx.x; idx < 64 * 64; idx += NT) {
        const uint32_t ti = idx / 64, lane = idx % 64, t = t0 + ti, gl = g * 64 + lane;
        uint32_t bt = INF, bl = INF, clash = 0;
        if (t < T && gl < nl) {
            const uint32_t lm = lim[gl];
            for (uint32_t q = toff[t]; q < toff[t + 1]; ++q) {
                const uint64_t x = base[(size_t)tpos[q] * 64 + lane];
                const uint32_t ft = (uint32_t)(x >> 32), fl = (uint32_t)x;
                if (ft != INF && ft <= lm) {
                    const uint32_t at = ft > tpt[q] ? ft - tpt[q] : 0u;
                    if (at < bt) {
                        bt = at;
                        bl = fl > tpl[q] ? fl - tpl[q] : 0u;
                    }
                }
                for (uint32_t k = soff[gl]; k < soff[gl + 1]; ++k) clash |= srank[k] == trank[q];
            }
            if (clash || (thr != INF && bt != INF && bt > thr)) bt = bl = INF;
            if (conf) conf[(size_t)gl * T + t] = (uint8_t)clash;
        }
        tt[ti][lane] = bt;
        tl[ti][lane] = bl;
    }
    __syncthreads();
    for (uint32_t idx = threadIdx.x; idx < 64 * 64; idx += NT) {
        const uint32_t lane = idx / 64, ti = idx % 64, gl = g * 64 + lane, t = t0 + ti;
        if (gl < nl && t < T) {
            o_t[(size_t)gl * T + t] = tt[ti][lane];
            o_l[(size_t)gl * T + t] = tl[ti][lane];
        }
    }
}

// Builds the two-channel arc lists from the CSR the context was created from
// and length weights aligned with it (arcs of INF time are dropped, as the
// time flats drop them).
int attach_length(bfg_ctx *c, hipStream_t s, const bfg_graph_desc *gd, const uint32_t *up_len,
                  const uint32_t *dn_len) {
    const bfg::Layout &L = c->L;
    const uint32_t n = L.n, nt = L.n_top;
    if (n >= (1u << 26)) return fail(BFG_NOT_SUPPORTED, "two-channel lists pack positions in 26 bits");
    // the upward walk visits a lane's nodes in descending position: every
    // ancestor must sit at a smaller position than its descendants
    for (uint32_t p = 0; p < n; ++p)
        if (L.pos_parent[p] != NONE && L.pos_parent[p] >= p)
            return fail(BFG_INTERNAL_ERROR, "a parent follows its child in position order");
    const uint32_t nseg = (uint32_t)L.seg_pos.size() - 1;
    for (uint32_t s = 0; s < nseg; ++s)
        if (L.seg_pos[s + 1] - L.seg_pos[s] > SEG) return fail(BFG_INTERNAL_ERROR, "segment longer than a tile");
    std::vector<uint32_t> seg_of(nt);
    for (uint32_t s = 0; s < nseg; ++s)
        for (uint32_t p = L.seg_pos[s]; p < L.seg_pos[s + 1]; ++p) seg_of[p] = s;
    auto pack = [](uint32_t t, uint32_t l) { return (uint64_t)t << 32 | l; };
    if (!gd->up_offsets || !gd->up_targets || !gd->up_weights || !gd->down_offsets || !gd->down_targets ||
        !gd->down_weights)
        return fail(BFG_INVALID_PARAMETER, "null graph array");
    // the graph must be the context's: every finite arc in range and in its
    // direction, and as many finite UP arcs per tail and DOWN arcs per head as
    // the context's layout holds
    std::vector<uint32_t> cu(n, 0), cd(n, 0);
    uint64_t mu = 0, md = 0;
    for (uint32_t r = 0; r < n; ++r) {
        if (gd->up_offsets[r + 1] < gd->up_offsets[r] || gd->down_offsets[r + 1] < gd->down_offsets[r])
            return fail(BFG_INVALID_PARAMETER, "offsets not monotone");
        for (uint64_t i = gd->up_offsets[r]; i < gd->up_offsets[r + 1]; ++i) {
            if (gd->up_weights[i] == INF) continue;
            if (gd->up_targets[i] >= n || gd->up_targets[i] <= r) return fail(BFG_INVALID_PARAMETER, "bad UP arc");
            ++cu[L.pos_of_rank[r]];
            ++mu;
        }
        for (uint64_t i = gd->down_offsets[r]; i < gd->down_offsets[r + 1]; ++i) {
            if (gd->down_weights[i] == INF) continue;
            if (gd->down_targets[i] >= r) return fail(BFG_INVALID_PARAMETER, "bad DOWN arc");
            ++cd[L.pos_of_rank[gd->down_targets[i]]];
            ++md;
        }
    }
    if (mu != L.m_up || md != L.m_down) return fail(BFG_INVALID_PARAMETER, "graph differs from the context's");
    for (uint32_t p = 0; p < n; ++p)
        if (cu[p] != L.up_off[p + 1] - L.up_off[p] || cd[p] != L.in_off[p + 1] - L.in_off[p])
            return fail(BFG_INVALID_PARAMETER, "graph differs from the context's");
    if ((mu && !up_len) || (md && !dn_len)) return fail(BFG_INVALID_PARAMETER, "null length weights");
    // DOWN arcs: count per list, then fill with tails in ascending position so
    // each segment's outside arcs come grouped by source
    std::vector<uint32_t> x_off(nseg + 1, 0), i_off(nt + 1, 0), b_off(n - nt + 1, 0);
    for (uint32_t r = 0; r < n; ++r) {
        const uint32_t ps = L.pos_of_rank[r];
        for (uint64_t i = gd->down_offsets[r]; i < gd->down_offsets[r + 1]; ++i) {
            if (gd->down_weights[i] == INF) continue;
            const uint32_t ph = L.pos_of_rank[gd->down_targets[i]];
            if (ph >= nt) ++b_off[ph - nt + 1];
            else if (ps < nt && seg_of[ps] == seg_of[ph]) ++i_off[ph + 1];
            else ++x_off[seg_of[ph] + 1];
        }
    }
    for (size_t k = 1; k < x_off.size(); ++k) x_off[k] += x_off[k - 1];
    for (size_t k = 1; k < i_off.size(); ++k) i_off[k] += i_off[k - 1];
    for (size_t k = 1; k < b_off.size(); ++k) b_off[k] += b_off[k - 1];
    for (uint32_t sg = 0; sg < nseg; ++sg)
        if (i_off[L.seg_pos[sg + 1]] - i_off[L.seg_pos[sg]] > SEG * (SEG - 1) / 2)
            return fail(BFG_INTERNAL_ERROR, "more arcs inside a segment than node pairs");
    std::vector<uint32_t> x_sj(x_off[nseg]), i_k(i_off[nt]), b_src(b_off[n - nt]);
    std::vector<uint64_t> x_w(x_sj.size()), i_w(i_k.size()), b_w(b_src.size());
    {
        std::vector<uint32_t> xc(x_off.begin(), x_off.end() - 1), ic(i_off.begin(), i_off.end() - 1),
            bc(b_off.begin(), b_off.end() - 1);
        for (uint32_t ps = 0; ps < n; ++ps) {
            const uint32_t r = L.rank_of_pos[ps];
            for (uint64_t i = gd->down_offsets[r]; i < gd->down_offsets[r + 1]; ++i) {
                if (gd->down_weights[i] == INF) continue;
                const uint32_t ph = L.pos_of_rank[gd->down_targets[i]];
                const uint64_t w2 = pack(gd->down_weights[i], dn_len[i]);
                if (ph >= nt) {
                    const uint32_t k = bc[ph - nt]++;
                    b_src[k] = ps;
                    b_w[k] = w2;
                } else if (ps < nt && seg_of[ps] == seg_of[ph]) {
                    const uint32_t k = ic[ph]++;
                    i_k[k] = ps - L.seg_pos[seg_of[ph]];
                    i_w[k] = w2;
                } else {
                    const uint32_t k = xc[seg_of[ph]]++;
                    x_sj[k] = ps << 6 | (ph - L.seg_pos[seg_of[ph]]);
                    x_w[k] = w2;
                }
            }
        }
    }
    // UP arcs by tail position
    std::vector<uint32_t> u_off(n + 1, 0);
    for (uint32_t r = 0; r < n; ++r) {
        uint32_t k = 0;
        for (uint64_t i = gd->up_offsets[r]; i < gd->up_offsets[r + 1]; ++i) k += gd->up_weights[i] != INF;
        u_off[L.pos_of_rank[r] + 1] = k;
    }
    for (size_t k = 1; k < u_off.size(); ++k) u_off[k] += u_off[k - 1];
    std::vector<uint32_t> u_tgt(u_off[n]);
    std::vector<uint64_t> u_w(u_tgt.size());
    for (uint32_t r = 0; r < n; ++r) {
        uint32_t k = u_off[L.pos_of_rank[r]];
        for (uint64_t i = gd->up_offsets[r]; i < gd->up_offsets[r + 1]; ++i) {
            if (gd->up_weights[i] == INF) continue;
            u_tgt[k] = L.pos_of_rank[gd->up_targets[i]];
            u_w[k++] = pack(gd->up_weights[i], up_len[i]);
        }
    }
    std::unique_ptr<TwoChannel> t(new TwoChannel());
    const StreamDrain drain{s};
    HIP_TRY(t->x_off.upload(x_off, s));
    HIP_TRY(t->x_sj.upload(x_sj, s));
    HIP_TRY(t->x_w.upload(x_w, s));
    HIP_TRY(t->i_off.upload(i_off, s));
    HIP_TRY(t->i_k.upload(i_k, s));
    HIP_TRY(t->i_w.upload(i_w, s));
    HIP_TRY(t->b_off.upload(b_off, s));
    HIP_TRY(t->b_src.upload(b_src, s));
    HIP_TRY(t->b_w.upload(b_w, s));
    HIP_TRY(t->u_off.upload(u_off, s));
    HIP_TRY(t->u_tgt.upload(u_tgt, s));
    HIP_TRY(t->u_w.upload(u_w, s));
    HIP_TRY(t->pos_of_rank.upload(L.pos_of_rank, s));
    HIP_TRY(hipStreamSynchronize(s));  // the host vectors die with this frame
    t->bytes = 4 * (x_off.size() + x_sj.size() + i_off.size() + i_k.size() + b_off.size() + b_src.size() +
                    u_off.size() + u_tgt.size() + n) +
               8 * (x_w.size() + i_w.size() + b_w.size() + u_w.size());
    for (size_t lv = 0; lv + 1 < L.seglvl_seg.size(); ++lv) {
        uint32_t m = 0;
        for (uint32_t sg = L.seglvl_seg[lv]; sg < L.seglvl_seg[lv + 1]; ++sg) m = std::max(m, x_off[sg + 1] - x_off[sg]);
        t->lvl_xmax.push_back(m);
    }
    c->t2 = std::move(t);
    return BFG_SUCCESS;
}

// Runs lanes [0, nl) of a two-channel request in batches of 64-lane groups;
// seeds per lane are distinct positions with their (time, length) labels.
// After each batch's sweep, emit(first lane, lanes, groups) reads the labels
// (t.lab, limits in t.lim) on the context stream.
// The bottom nodes a sweep covers: all of them, or per bottom level l the
// positions d_pos[off[l] .. off[l+1]) (off on the host, d_pos on the device).
struct BottomSet {
    const std::vector<uint32_t> *off = nullptr;
    const uint32_t *d_pos = nullptr;
};

template <class Emit>
int phast_2ch(bfg_ctx *c, hipStream_t s, const std::vector<uint32_t> &soff, const std::vector<uint32_t> &spos,
              const std::vector<uint64_t> &slab, const uint32_t *limits, uint32_t nl, bool dense, BottomSet bs,
              Emit emit) {
    TwoChannel &t = *c->t2;
    const bfg::Layout &L = c->L;
    const uint32_t n = L.n;
    // groups per batch: the labels (and for dense rows both output channels)
    // fit a share of free memory; a sweep's time barely grows with its groups
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)8 << 30;
    const size_t per_group = (size_t)n * 64 * (dense ? 16 : 8) + (size_t)L.n_top * 64 * 8;
    const size_t held = (t.lab.n + t.top.n) * 8 + (t.o_t.n + t.o_l.n) * 4;
    const uint32_t want = (nl + 63) / 64;
    uint32_t gmax = (uint32_t)std::max<size_t>(1, std::min<size_t>(64, (free_b + held) / 2 / per_group));
    if (const char *e = std::getenv("BFG_T2_GROUPS")) {  // tuning override: groups per batch
        const long v = std::strtol(e, nullptr, 10);
        if (v >= 1 && v <= 64) gmax = std::min<uint32_t>(gmax, (uint32_t)v);
    }
    const uint32_t G = std::min(want, gmax);
    if (t.lab.n < (size_t)G * n * 64) {
        t.lab.release();
        t.top.release();
        t.o_t.release();
        t.o_l.release();
        HIP_TRY(t.lab.alloc((size_t)G * n * 64));
        HIP_TRY(t.top.alloc((size_t)G * 64 * std::max<uint32_t>(L.n_top, 1)));
    }
    if (dense && t.o_t.n < (size_t)G * 64 * n) {
        t.o_t.release();
        t.o_l.release();
        HIP_TRY(t.o_t.alloc((size_t)G * 64 * n));
        HIP_TRY(t.o_l.alloc((size_t)G * 64 * n));
    }
    for (uint32_t l0 = 0; l0 < nl; l0 += G * 64) {
        const uint32_t bl = std::min(nl - l0, G * 64), bg = (bl + 63) / 64;
        // this batch's seeds, offsets rebased
        std::vector<uint32_t> off(bl + 1);
        for (uint32_t i = 0; i <= bl; ++i) off[i] = soff[l0 + i] - soff[l0];
        const uint32_t k0 = soff[l0], k1 = soff[l0 + bl];
        std::vector<uint32_t> pos(spos.begin() + k0, spos.begin() + k1), lim(limits + l0, limits + l0 + bl);
        std::vector<uint64_t> lb(slab.begin() + k0, slab.begin() + k1);
        const StreamDrain drain{s};
        HIP_TRY(t.s_off.upload(off, s));
        HIP_TRY(t.s_pos.upload(pos, s));
        HIP_TRY(t.s_lab.upload(lb, s));
        HIP_TRY(t.lim.upload(lim, s));
        const size_t cells = (size_t)bg * n * 64;
        const size_t tcells = (size_t)bg * 64 * L.n_top;
        t2_fill_kernel<<<(uint32_t)std::min<size_t>((cells + NT - 1) / NT, 1u << 20), NT, 0, s>>>(t.lab.p, cells);
        if (tcells)
            t2_fill_kernel<<<(uint32_t)std::min<size_t>((tcells + NT - 1) / NT, 1u << 20), NT, 0, s>>>(t.top.p,
                                                                                                     tcells);
        t2_up_kernel<<<bl, 64, 0, s>>>(bl, n, L.n_top, t.s_off.p, t.s_pos.p, t.s_lab.p, t.lim.p, c->pos_parent.p,
                                       t.u_off.p, t.u_tgt.p, t.u_w.p, t.lab.p, t.top.p);
        if (L.n_top)
            t2_top_kernel<<<dim3((L.n_top + 63) / 64, bg), NT, 0, s>>>(n, L.n_top, t.top.p, t.lab.p);
        const uint32_t small = t2_small_arcs();
        for (size_t lv = 0; lv + 1 < L.seglvl_seg.size(); ++lv) {
            const uint32_t s0 = L.seglvl_seg[lv], s1 = L.seglvl_seg[lv + 1];
            if (s1 == s0) continue;
            if (t.lvl_xmax[lv] <= small)
                t2_seg_kernel<8><<<dim3(s1 - s0, bg), 8 * 64, 0, s>>>(s0, n, bl, c->seg_pos.p, t.x_off.p, t.x_sj.p,
                                                                     t.x_w.p, t.i_off.p, t.i_k.p, t.i_w.p, t.lim.p,
                                                                     t.lab.p);
            else
                t2_seg_kernel<T2_SEG_WAVES><<<dim3(s1 - s0, bg), T2_SEG_WAVES * 64, 0, s>>>(
                    s0, n, bl, c->seg_pos.p, t.x_off.p, t.x_sj.p, t.x_w.p, t.i_off.p, t.i_k.p, t.i_w.p, t.lim.p,
                    t.lab.p);
        }
        for (size_t lv = 0; lv + 1 < L.botlvl_pos.size(); ++lv) {
            const uint32_t p0 = bs.off ? (*bs.off)[lv] : L.botlvl_pos[lv];
            const uint32_t p1 = bs.off ? (*bs.off)[lv + 1] : L.botlvl_pos[lv + 1];
            if (p1 > p0)
                t2_pull_kernel<<<dim3((p1 - p0 + NT / 64 - 1) / (NT / 64), bg), NT, 0, s>>>(
                    p0, p1, bs.d_pos, n, L.n_top, bl, t.b_off.p, t.b_src.p, t.b_w.p, t.lim.p, t.lab.p);
        }
        HIP_TRY(hipGetLastError());
        const int rc = emit(l0, bl, bg);
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(s));
    }
    return BFG_SUCCESS;
}

// Per lane, its seeds as the reference applies them (isochrone_handler.rs:
// 522-526: a seed applies while strictly faster than what its rank holds, so
// per rank the first of the fastest wins; an INF seed never applies).
int seeds_2ch(const bfg::Layout &L, const uint32_t *off, const uint32_t *rank, const uint32_t *tm,
              const uint32_t *ln, uint32_t nl, std::vector<uint32_t> &soff, std::vector<uint32_t> &spos,
              std::vector<uint64_t> &slab) {
    soff.assign(1, 0);
    for (uint32_t i = 0; i < nl; ++i) {
        const size_t first = spos.size();
        for (uint32_t k = off[i]; k < off[i + 1]; ++k) {
            if (tm[k] == INF) continue;
            const uint32_t p = L.pos_of_rank[rank[k]];
            const uint64_t x = (uint64_t)tm[k] << 32 | ln[k];
            size_t j = first;
            while (j < spos.size() && spos[j] != p) ++j;
            if (j == spos.size()) {
                spos.push_back(p);
                slab.push_back(x);
            } else if (tm[k] < (uint32_t)(slab[j] >> 32)) {
                slab[j] = x;
            }
        }
        if (spos.size() - first > 64) return fail(BFG_NOT_SUPPORTED, "more than 64 seed ranks in one lane");
        soff.push_back((uint32_t)spos.size());
    }
    return BFG_SUCCESS;
}

void fill_info(const bfg::Layout &L, bfg_ctx_info *info) {
    memset(info, 0, sizeof(*info));
    info->n_nodes = L.n;
    info->m_up = L.m_up;
    info->m_down = L.m_down;
    info->etree_depth = L.max_depth;
    info->n_top = L.n_top;
    info->n_segments = (uint32_t)L.seg_pos.size() - 1;
    info->n_segment_levels = (uint32_t)L.seglvl_seg.size() - 1;
    info->n_bottom_levels = (uint32_t)L.botlvl_pos.size() - 1;
}

bfg::CchInput input_of(const bfg_graph_desc *gd) {
    bfg::CchInput in;
    in.n = gd->n_nodes;
    in.up_off = gd->up_offsets; in.up_tgt = gd->up_targets; in.up_w = gd->up_weights;
    in.dn_off = gd->down_offsets; in.dn_tgt = gd->down_targets; in.dn_w = gd->down_weights;
    return in;
}

// The transposed graph of gd (the reverse metric's CCH): UP arcs at tail y
// are the DOWN arcs x -> y reversed, DOWN arcs at tail v the UP arcs u -> v
// reversed, each tail's targets ascending (the old tails are visited in
// ascending order).  Returns false on a target out of range.
struct Transposed {
    std::vector<uint64_t> up_off, dn_off;
    std::vector<uint32_t> up_tgt, up_w, dn_tgt, dn_w;
};

bool flip_csr(uint32_t n, const uint64_t *off, const uint32_t *tgt, const uint32_t *w, std::vector<uint64_t> &noff,
              std::vector<uint32_t> &ntgt, std::vector<uint32_t> &nw) {
    const uint64_t m = off[n];
    noff.assign((size_t)n + 1, 0);
    for (uint64_t i = 0; i < m; ++i) {
        if (tgt[i] >= n) return false;
        noff[tgt[i] + 1]++;
    }
    for (uint32_t v = 0; v < n; ++v) noff[v + 1] += noff[v];
    ntgt.resize(m);
    nw.resize(m);
    std::vector<uint64_t> fill(noff.begin(), noff.end() - 1);
    for (uint32_t r = 0; r < n; ++r)
        for (uint64_t i = off[r]; i < off[r + 1]; ++i) {
            const uint64_t k = fill[tgt[i]]++;
            ntgt[k] = r;
            nw[k] = w[i];
        }
    return true;
}

// Host tables of the block gate (see bfg_ctx): segment parents, the bottom
// trees in pre-order (ch_pos lists the bottom positions in pre-order, and a
// tree's nodes are consecutive there), the tree of each bottom position and
// the trees each bottom chunk touches.
struct GateTables {
    std::vector<uint32_t> xr_min, cs_min, pos_bt, bt_ch, xr_w16, xr_cap;
};

// Host tables of the block gate (see bfg_ctx): the least weight of every DOWN
// row and of every bottom chunk source's arcs; the bottom trees in pre-order
// (ch_pos lists the bottom positions in pre-order, and a tree's nodes are
// consecutive there), the tree of each bottom position and the chunks each
// tree touches.
int gate_tables(const bfg::Layout &L, GateTables &G) {
    const uint32_t nseg = (uint32_t)L.seg_pos.size() - 1, nt = L.n_top;
    G.xr_min.assign(L.xr_off.empty() ? 0 : L.xr_off.back(), INF);
    G.xr_cap.assign(G.xr_min.size(), 0);
    for (uint32_t sg = 0; sg < nseg; ++sg) {
        uint64_t w = L.xw_off[sg];
        for (uint32_t k = 0, st = 0; k < 8; st = L.xr_run[8 * (size_t)sg + k++]) {
            const uint32_t W = 8 * (8 - k), end = L.xr_run[8 * (size_t)sg + k];
            for (uint32_t i = st; i < end; ++i, w += W) {
                uint32_t m = INF, mx = 0;
                for (uint32_t j = 0; j < W; ++j) {
                    const uint32_t x = L.xr_w[w + j];
                    m = std::min(m, x);
                    if (x != INF) mx = std::max(mx, x);
                }
                G.xr_min[L.xr_off[sg] + i] = m;
                G.xr_cap[L.xr_off[sg] + i] = mx >= 0xFFFFu ? 0u : 0xFFFFu - mx;
            }
        }
        if (w != L.xw_off[sg + 1]) return fail(BFG_INTERNAL_ERROR, "row runs disagree");
    }
    if (L.xr_w.size() % 2) return fail(BFG_INTERNAL_ERROR, "row weights not in whole octets");
    G.xr_w16.resize(L.xr_w.size() / 2);
    for (size_t i = 0; i < G.xr_w16.size(); ++i)
        G.xr_w16[i] = std::min(L.xr_w[2 * i], 0xFFFFu) | std::min(L.xr_w[2 * i + 1], 0xFFFFu) << 16;
    G.cs_min.assign(L.cs_src.size(), INF);
    for (size_t k = 0; k < L.cs_src.size(); ++k)
        for (uint32_t a = L.cs_aoff[k]; a < L.cs_aoff[k + 1]; ++a) G.cs_min[k] = std::min(G.cs_min[k], L.ca_w[a]);
    G.pos_bt.assign(L.n - nt, NONE);
    const size_t nch = L.ch_off.size() - 1;
    uint32_t nbt = 0;
    for (size_t c = 0; c < nch; ++c)
        for (uint32_t i = L.ch_off[c]; i < L.ch_off[c + 1]; ++i) {
            const uint32_t p = L.ch_pos[i], pp = L.pos_parent[p];
            if (pp == NONE || pp < nt) {  // a bottom tree's root
                G.bt_ch.push_back((uint32_t)c);
                G.bt_ch.push_back((uint32_t)c);
                ++nbt;
            } else if (nbt == 0) {
                return fail(BFG_INTERNAL_ERROR, "bottom positions not in pre-order");
            }
            G.pos_bt[p - nt] = nbt - 1;
            G.bt_ch[2 * (nbt - 1) + 1] = (uint32_t)c;
        }
    for (uint32_t p = nt; p < L.n; ++p) {  // a node's tree is its parent's
        const uint32_t pp = L.pos_parent[p];
        if (G.pos_bt[p - nt] == NONE || (pp != NONE && pp >= nt && G.pos_bt[p - nt] != G.pos_bt[pp - nt]))
            return fail(BFG_INTERNAL_ERROR, "a bottom tree is not consecutive in pre-order");
    }
    return BFG_SUCCESS;
}

// ---------------------------------------------------------------------------
// The allocation-free matrix call.  Its device memory is one file-scope
// array per GPU, `scratch` (SCRATCH_BYTES), which every context on that GPU
// carves the same way for its own graph.  Its address on a device comes from
// hipGetSymbolAddress, once per device, at context creation with that device
// current; host code uses only that address, never the array's name.  Calls
// on one GPU are serialised by a per-device lock; each call leaves its own
// contents behind and clears or overwrites everything it reads.
// ---------------------------------------------------------------------------
constexpr size_t SCRATCH_BYTES = (size_t)512 << 20;
constexpr uint32_t NOALLOC_LANES = 64;               // one wave group
constexpr uint32_t NOALLOC_MAX_TARGETS = 1u << 18;
constexpr uint32_t NOALLOC_MAX_PATH = 1u << 21;      // bottom path elements above the targets
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) uint32_t scratch[SCRATCH_BYTES / 4];

std::mutex g_scratch_init_mu;
uint32_t *g_scratch_ptr[MAX_DEVICES] = {};
std::mutex g_scratch_mu[MAX_DEVICES];

// The scratch array's device address on `device` (current), looked up once.
int scratch_setup(int device, uint32_t **out) {
    *out = nullptr;
    if (device >= MAX_DEVICES) return BFG_SUCCESS;  // the allocation-free call is then unavailable
    std::lock_guard<std::mutex> g(g_scratch_init_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return fail(BFG_DEVICE_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = (uint32_t *)p;
    }
    *out = g_scratch_ptr[device];
    return BFG_SUCCESS;
}

// The scratch Work of a context: its buffers borrowed from the array at
// `base`, the label arena taking what the fixed tables leave; its host
// staging reserved to the call's limits.  Leaves scratch_work null when the
// graph's tables do not fit (the call then reports it).
int scratch_work_init(bfg_ctx *c, uint32_t *base) {
    if (!base) return BFG_SUCCESS;
    const size_t nseg = c->L.seg_pos.size() - 1, D1 = d1_of(c);
    std::unique_ptr<Work> w(new Work());
    int rc = work_init(w.get());
    if (rc) return rc;
    w->fixed = true;
    w->time_dom = false;
    size_t at = 0;                          // in u32 words, each piece 256-byte aligned
    const size_t total = SCRATCH_BYTES / 4;
    bool fits = true;
    auto take = [&](size_t words) -> uint32_t * {
        at = (at + 63) / 64 * 64;
        if (at + words > total) {
            fits = false;
            return base;
        }
        uint32_t *p = base + at;
        at += words;
        return p;
    };
    auto take32 = [&](DevBuf<uint32_t> &b, size_t count) { b.borrow(take(count), count); };
    auto take8 = [&](DevBuf<uint8_t> &b, size_t count) { b.borrow((uint8_t *)take((count + 3) / 4), count); };
    take32(w->seed_off, NOALLOC_LANES + 1);
    take32(w->lane_limit, NOALLOC_LANES);
    take32(w->row_map, NOALLOC_LANES);
    for (DevBuf<uint32_t> *b : {&w->seed_rank, &w->seed_pos, &w->seed_cost, &w->seed_depth, &w->seed_pre})
        take32(*b, NOALLOC_LANES);
    take32(w->pathbuf, NOALLOC_LANES * D1);
    take32(w->tgt_pos, NOALLOC_MAX_TARGETS);
    take32(w->path_off, NOALLOC_MAX_TARGETS + 1);
    take32(w->path_pos, NOALLOC_MAX_PATH);
    // gate flags of bounded calls: one wave group, at most 64 / 16 lane groups
    take8(w->seg_live, std::max<size_t>(nseg, 1));
    take8(w->chunk_ok, 4 * std::max<size_t>(c->L.ch_off.size() - 1, 1));
    take8(w->open_b, 4 * (size_t)std::max<uint32_t>(c->n_bt, 1));
    at = (at + 63) / 64 * 64;
    if (!fits || at >= total) return BFG_SUCCESS;
    w->arena.borrow(base + at, total - at);
    // host staging: capacity for the largest call, so the call never grows it
    for (auto *v : {&w->h_soff, &w->h_lim, &w->h_rows, &w->h_srank, &w->h_scost, &w->h_sdepth, &w->h_spre,
                    &w->h_spos, &w->h_src, &w->req.soff, &w->req.srank, &w->req.scost, &w->req.limit})
        v->reserve(NOALLOC_LANES + 1);
    w->h_key.reserve(NOALLOC_LANES);
    for (auto *v : {&w->h_tgt, &w->h_bt, &w->h_pending, &w->path_off_h, &w->path_leaf, &w->tgt_pos_h})
        v->reserve(NOALLOC_MAX_TARGETS + 1);
    w->h_elem.reserve(NOALLOC_MAX_TARGETS);
    w->path_pos_h.reserve(NOALLOC_MAX_PATH);
    c->scratch_work = std::move(w);
    return BFG_SUCCESS;
}

int matrix_impl(bfg_ctx *c, Work *w, Req &req, const uint32_t *src, uint32_t S, const uint32_t *tgt, uint32_t T,
                uint32_t thr, uint32_t *d_out, uint32_t *host_out) {
    single_source_req(req, src, S, thr);
    int rc = stage_target_paths(c, w, tgt, T);
    if (rc) return rc;
    PathsOn paths_on(w);  // restrict the bottom sweep for this call only
    DevBuf<uint32_t> tmp;
    if (!d_out) {
        HIP_TRY(tmp.alloc((size_t)S * T));
        d_out = tmp.p;
    }
    rc = run_all(c, w, req, 0, [&](uint32_t, uint32_t nl, const std::vector<uint32_t> &, float *oms) -> int {
        return gather_out(c, w, nl, w->tgt_pos.p, T, w->row_map.p, d_out, T, nullptr, 0, oms);
    });
    if (rc) return rc;
    if (host_out) HIP_TRY(hipMemcpy(host_out, d_out, (size_t)S * T * 4, hipMemcpyDeviceToHost));
    return BFG_SUCCESS;
}

// Several matrix requests in one sweep: request r has sources
// src[src_off[r] .. src_off[r+1]) and targets tgt[tgt_off[r] .. tgt_off[r+1])
// and limit thr[r]; its S_r x T_r row-major block goes to d_out from
// sum_{q < r} S_q T_q on.  Lanes are sorted by request, then by pre-order, so
// each request's lanes are consecutive; the bottom sweep covers the paths
// above the union of all targets, and each request gathers its own columns.
int matrix_multi(bfg_ctx *c, Work *w, uint32_t nr, const uint32_t *src_off, const uint32_t *src,
                 const uint32_t *tgt_off, const uint32_t *tgt, const uint32_t *thr, uint32_t *d_out) {
    const uint32_t S = src_off[nr], NT_ = tgt_off[nr];
    Req req;
    single_source_req(req, src, S, INF);
    req.major.resize(S);
    std::vector<uint64_t> out_off(nr + 1, 0);
    for (uint32_t r = 0; r < nr; ++r) {
        for (uint32_t l = src_off[r]; l < src_off[r + 1]; ++l) {
            req.major[l] = r;
            req.limit[l] = thr[r];
        }
        out_off[r + 1] = out_off[r] + (uint64_t)(src_off[r + 1] - src_off[r]) * (tgt_off[r + 1] - tgt_off[r]);
    }
    std::vector<uint32_t> U(tgt, tgt + NT_);
    std::sort(U.begin(), U.end());
    U.erase(std::unique(U.begin(), U.end()), U.end());
    int rc = stage_target_paths(c, w, U.data(), (uint32_t)U.size());
    if (rc) return rc;
    // each request's targets as label indices of the union's paths
    std::vector<uint32_t> tlab(NT_);
    for (uint32_t q = 0; q < NT_; ++q)
        tlab[q] = w->tgt_pos_h[std::lower_bound(U.begin(), U.end(), tgt[q]) - U.begin()];
    DevBuf<uint32_t> d_tlab;
    const StreamDrain drain{w->stream};
    HIP_TRY(d_tlab.upload(tlab, w->stream));
    PathsOn paths_on(w);
    return run_all(c, w, req, 0, [&](uint32_t b0, uint32_t nl, const std::vector<uint32_t> &, float *oms) -> int {
        // request r holds sorted lanes [src_off[r], src_off[r+1]); this batch [b0, b0 + nl)
        for (uint32_t r = 0; r < nr; ++r) {
            const uint32_t lo = std::max(src_off[r], b0), hi = std::min(src_off[r + 1], b0 + nl);
            const uint32_t T = tgt_off[r + 1] - tgt_off[r];
            if (lo >= hi || T == 0) continue;
            const int rc2 = gather_out(c, w, hi - b0, d_tlab.p + tgt_off[r], T, w->row_map.p, d_out + out_off[r], T,
                                       nullptr, 0, oms, lo - b0, src_off[r]);
            if (rc2) return rc2;
        }
        return BFG_SUCCESS;
    });
}

// The root position of the bottom tree (maximal bottom-region subtree) that
// holds bottom position p.
uint32_t bottom_root(const bfg_ctx *c, uint32_t p) { return c->bt_root[p - c->L.n_top]; }

// Runs one Work's batches through the upward phase only; always switched off
// again.
struct UpOnly {
    Work *w;
    explicit UpOnly(Work *ww) : w(ww) { w->up_only = true; }
    ~UpOnly() {
        w->up_only = false;
        w->cap_hint = 0;
    }
};

// Hand-over between the two sides of a join run at once: the target side
// signals its labels ready and waits until the join is done with them.
struct MeetSync {
    std::mutex mu;
    std::condition_variable cv;
    bool ready = false, failed = false, done = false;
    void set(bool MeetSync::*flag) {
        {
            std::lock_guard<std::mutex> g(mu);
            this->*flag = true;
        }
        cv.notify_all();
    }
};

// Wavefronts per tile of the join: enough to give a small matrix's few tiles
// about 4096 wavefronts (four per SIMD), one for large matrices.
uint32_t join_splits(uint32_t S, uint32_t T) {
    const uint64_t tiles = (uint64_t)((S + 63) / 64) * ((T + 63) / 64);
    return (uint32_t)std::min<uint64_t>(16, std::max<uint64_t>(1, 4096 / std::max<uint64_t>(tiles, 1)));
}

// The bucket many-to-many (table_bucket_parallel_bounded, bucket_ch.rs:
// 3436-3520) on the elimination tree: the target lanes' upward phase on the
// transposed context (the backward searches), the source lanes' upward phase
// on this one (the forward searches), then the join over their common
// meeting nodes: the tiles of top-region segments both wave groups reach
// (bfg_meet::launch_join) and the pairs that meet inside one bottom tree.
// Cells above thr are INF.  Target batches are the outer loop, so a matrix
// too large for two arenas reruns the sources once per target batch.
int bucket_matrix(bfg_ctx *c, const uint32_t *src, uint32_t S, const uint32_t *tgt, uint32_t T, uint32_t thr,
                  uint32_t *d_out, uint32_t *host_out) {
    bfg_ctx *r = c->rev;
    if (!r) return fail(BFG_NOT_SUPPORTED, "no reverse context is paired (bfg_ctx_pair_reverse)");
    std::shared_lock<std::shared_mutex> rcfg(r->cfg);
    WorkLease rlease;
    int rc = acquire_work(r, rlease);
    if (rc) return rc;
    CallScope cs(c);
    rc = begin_call(c, cs);
    if (rc) return rc;
    Work *wf = cs.lease.w, *wb = rlease.w;
    UpOnly uf(wf), ub(wb);
    const bfg::Layout &L = c->L;
    Req rs, rt;
    single_source_req(rs, src, S, thr);
    single_source_req(rt, tgt, T, thr);
    DevBuf<uint32_t> tmp;
    if (!d_out) {
        HIP_TRY(tmp.alloc((size_t)S * T));
        d_out = tmp.p;
    }
    const uint32_t nseg = (uint32_t)L.seg_pos.size() - 1, words = std::max<uint32_t>((nseg + 63) / 64, 1);
    const auto t_start = std::chrono::steady_clock::now();
    std::vector<uint64_t> tree_tgts;  // (bottom tree root << 32 | target lane), ascending
    float join_ms = 0, kernel_ms = 0;
    uint32_t joins = 0;
    uint64_t pairs = 0, visits = 0, redone = 0;
    bool bits16 = true;
    const char *u32_only = std::getenv("BFG_JOIN_U32");  // override: the u32 join alone
    // Lane capacities: the targets' arena first, the sources' from what is
    // left.  When both sides fit one batch, their upward phases run at once
    // (the target side on a thread of its own) and the join splits its tiles.
    const uint32_t cap_t = r->cap_fixed ? r->lane_cap : auto_lane_cap(r, wb, T, 0);
    const char *serial = std::getenv("BFG_BUCKET_SERIAL");  // profiling override: one side after the other
    bool together = cap_t >= T && !(serial && serial[0] == '1');
    if (together && !r->cap_fixed) {
        rc = ensure_arena(r, wb, (size_t)cap_t * label_stride(r, wb));
        if (rc) return rc;
        wb->cap_hint = cap_t;
    }
    const uint32_t cap_s = c->cap_fixed ? c->lane_cap : auto_lane_cap(c, wf, S, 0);
    together &= cap_s >= S;
    if (together && !c->cap_fixed) wf->cap_hint = cap_s;
    const uint32_t splits = together ? join_splits(S, T) : 1;

    // target batch ready on wb: its reach bits and its bottom targets by tree
    auto targets_ready = [&](uint32_t tnl, const std::vector<uint32_t> &trows) -> int {
        const uint32_t wgb = (tnl + 63) / 64;
        ENSURE(wb, reach, (size_t)wgb * words, "reach bits");
        HIP_TRY(bfg_meet::launch_reach(wb->stream, wb->seed_off.p, wb->seed_pre.p, tnl, r->pos_pre.p, r->pos_sub.p,
                                       r->seg_pos.p, nseg, words, wb->reach.p));
        HIP_TRY(hipStreamSynchronize(wb->stream));
        tree_tgts.clear();
        for (uint32_t b = 0; b < tnl; ++b) {
            const uint32_t p = L.pos_of_rank[tgt[trows[b]]];
            if (p >= L.n_top) tree_tgts.push_back((uint64_t)bottom_root(c, p) << 32 | b);
        }
        std::sort(tree_tgts.begin(), tree_tgts.end());
        return BFG_SUCCESS;
    };
    // the join of source batch snl (on wf) with the ready target batch tnl
    auto join = [&](uint32_t snl, const std::vector<uint32_t> &srows, uint32_t tnl) -> int {
        const uint32_t wga = (snl + 63) / 64;
        ENSURE(wf, reach, (size_t)wga * words, "reach bits");
        HIP_TRY(hipEventRecord(wf->ev[0], wf->stream));
        HIP_TRY(bfg_meet::launch_reach(wf->stream, wf->seed_off.p, wf->seed_pre.p, snl, c->pos_pre.p, c->pos_sub.p,
                                       c->seg_pos.p, nseg, words, wf->reach.p));
        if (splits > 1) HIP_TRY(bfg_meet::launch_fill(wf->stream, d_out, (uint64_t)S * T, INF));
        ENSURE(wf, visits, 2, "join counters");
        HIP_TRY(bfg_meet::launch_fill(wf->stream, (uint32_t *)wf->visits.p, 4, 0));
        bfg_meet::Side A{wf->arena.p, L.n_top, c->lanes, snl, wf->reach.p, wf->row_map.p, wf->segmin.p};
        bfg_meet::Side B{wb->arena.p, L.n_top, r->lanes, tnl, wb->reach.p, wb->row_map.p, wb->segmin.p};
        // one wavefront a tile on both sides' 16-bit labels, tiles with a cell
        // past 16 bits again on the u32 labels
        const bool use16 = splits == 1 && c->lanes == 64 && r->lanes == 64 && wf->lab16_ok && wb->lab16_ok &&
                           !(u32_only && u32_only[0] == '1');
        if (use16) {
            A.lab16 = wf->lab16.p;
            A.slot = wf->slot.p;
            A.lmin16 = wf->lmin16.p;
            B.lab16 = wb->lab16.p;
            B.slot = wb->slot.p;
            B.lmin16 = wb->lmin16.p;
            const size_t tiles = (size_t)((snl + 63) / 64) * ((tnl + 63) / 64);
            ENSURE(wf, redo, std::max<size_t>(tiles, 1), "join tiles");
            HIP_TRY(bfg_meet::launch_fill(wf->stream, wf->redo.p, tiles, 0));
        }
        HIP_TRY(hipEventRecord(wf->ev[1], wf->stream));
        if (use16) {
            HIP_TRY(bfg_meet::launch_join16(wf->stream, A, B, c->seg_pos.p, words, nseg, thr, d_out, T, wf->visits.p,
                                            wf->redo.p));
            HIP_TRY(bfg_meet::launch_join(wf->stream, A, B, c->seg_pos.p, words, nseg, thr, 1, d_out, T, nullptr,
                                          wf->redo.p));
        } else {
            HIP_TRY(bfg_meet::launch_join(wf->stream, A, B, c->seg_pos.p, words, nseg, thr, splits, d_out, T,
                                          wf->visits.p));
        }
        HIP_TRY(hipEventRecord(wf->ev[2], wf->stream));
        // sources and targets in one bottom tree also meet below the top region
        std::vector<uint32_t> &pr = wf->h_pairs;
        pr.clear();
        if (!tree_tgts.empty())
            for (uint32_t a = 0; a < snl; ++a) {
                const uint32_t p = L.pos_of_rank[src[srows[a]]];
                if (p < L.n_top) continue;
                const uint64_t root = bottom_root(c, p);
                for (auto it = std::lower_bound(tree_tgts.begin(), tree_tgts.end(), root << 32);
                     it != tree_tgts.end() && (*it >> 32) == root; ++it) {
                    pr.push_back(a);
                    pr.push_back((uint32_t)*it);
                }
            }
        const StreamDrain drain{wf->stream};  // the pair list is read by an asynchronous copy
        if (!pr.empty()) {
            ENSURE(wf, pairs, pr.size(), "bottom pairs");
            HIP_TRY(hipMemcpyAsync(wf->pairs.p, pr.data(), pr.size() * 4, hipMemcpyHostToDevice, wf->stream));
            const bfg_meet::PathSide PA{wf->seed_off.p, wf->seed_pos.p, wf->pathbuf.p, wf->row_map.p};
            const bfg_meet::PathSide PB{wb->seed_off.p, wb->seed_pos.p, wb->pathbuf.p, wb->row_map.p};
            HIP_TRY(bfg_meet::launch_bottom_pairs(wf->stream, wf->pairs.p, (uint32_t)(pr.size() / 2), PA, PB,
                                                  d1_of(c), c->pos_parent.p, c->pos_depth.p, L.n_top, thr, d_out,
                                                  T));
            pairs += pr.size() / 2;
        }
        HIP_TRY(hipEventRecord(wf->ev[3], wf->stream));
        HIP_TRY(hipEventSynchronize(wf->ev[3]));
        float ms = 0, kms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, wf->ev[0], wf->ev[3]));
        HIP_TRY(hipEventElapsedTime(&kms, wf->ev[1], wf->ev[2]));
        join_ms += ms;
        kernel_ms += kms;
        unsigned long long cnt[2] = {0, 0};
        HIP_TRY(hipMemcpy(cnt, wf->visits.p, sizeof(cnt), hipMemcpyDeviceToHost));
        visits += cnt[0];
        redone += cnt[1];
        bits16 &= use16;
        ++joins;
        return BFG_SUCCESS;
    };

    float up_f = 0;
    if (together) {
        MeetSync ms;
        uint32_t t_lanes = 0;
        int rc_t = BFG_SUCCESS;
        std::string err_t;
        std::thread th([&] {
            if (hipSetDevice(c->device) != hipSuccess) rc_t = fail(BFG_DEVICE_ERROR, "hipSetDevice");
            if (!rc_t)
                rc_t = run_all(r, wb, rt, 0, [&](uint32_t, uint32_t tnl, const std::vector<uint32_t> &trows,
                                                 float *) -> int {
                    const int rc2 = targets_ready(tnl, trows);
                    if (rc2) return rc2;
                    t_lanes = tnl;
                    ms.set(&MeetSync::ready);
                    std::unique_lock<std::mutex> g(ms.mu);
                    ms.cv.wait(g, [&] { return ms.done; });  // the join reads this batch's labels
                    return BFG_SUCCESS;
                });
            if (rc_t) err_t = g_last_error;
            ms.set(&MeetSync::failed);  // wakes a join still waiting (no-op once ready)
        });
        rc = run_all(c, wf, rs, 0, [&](uint32_t, uint32_t snl, const std::vector<uint32_t> &srows,
                                       float *) -> int {
            {
                std::unique_lock<std::mutex> g(ms.mu);
                ms.cv.wait(g, [&] { return ms.ready || ms.failed; });
                if (!ms.ready) return BFG_INTERNAL_ERROR;  // the target side failed; its error is reported
            }
            return join(snl, srows, t_lanes);
        });
        const std::string err_s = g_last_error;
        ms.set(&MeetSync::done);
        th.join();
        if (rc_t) return fail(rc_t, err_t);
        if (rc) return fail(rc, err_s);
        up_f = wf->last.up_ms;
    } else {
        rc = run_all(r, wb, rt, 0, [&](uint32_t, uint32_t tnl, const std::vector<uint32_t> &trows, float *) -> int {
            int rc2 = targets_ready(tnl, trows);
            if (rc2) return rc2;
            rc2 = run_all(c, wf, rs, 0, [&](uint32_t, uint32_t snl, const std::vector<uint32_t> &srows,
                                            float *) -> int { return join(snl, srows, tnl); });
            if (rc2) return rc2;
            up_f += wf->last.up_ms;
            return BFG_SUCCESS;
        });
        if (rc) return rc;
    }
    if (host_out) HIP_TRY(hipMemcpy(host_out, d_out, (size_t)S * T * 4, hipMemcpyDeviceToHost));
    bfg_timings tm{};
    tm.up_ms = up_f + wb->last.up_ms;  // forward and backward searches (overlapping when run at once)
    tm.sweep_ms = join_ms;             // the join takes the place of the DOWN sweep
    tm.batches = joins;
    tm.sweep_launches = joins;
    tm.lanes_run = (uint64_t)S + T;
    tm.nodes_swept = pairs;  // pairs that met inside a bottom tree
    tm.join_ms = kernel_ms;
    tm.join_visits = visits;
    tm.join_label_bits = joins && bits16 ? 16 : 32;
    tm.join_tiles_u32 = redone;
    tm.total_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    std::lock_guard<std::mutex> g(c->last_mu);
    c->last = tm;
    return BFG_SUCCESS;
}

// Seeded lanes of a call from the caller's CSR (at most 64 seeds per lane).
int seeded_req(bfg_ctx *c, Req &req, const uint32_t *seed_offsets, const uint32_t *seed_ranks,
               const uint32_t *seed_costs, uint32_t L, const uint32_t *thresholds) {
    if (seed_offsets[0] != 0) return fail(BFG_INVALID_PARAMETER, "seed_offsets[0] must be 0");
    for (uint32_t i = 0; i < L; ++i) {
        if (seed_offsets[i + 1] < seed_offsets[i]) return fail(BFG_INVALID_PARAMETER, "seed_offsets not monotone");
        if (seed_offsets[i + 1] - seed_offsets[i] > 64) return fail(BFG_NOT_SUPPORTED, "more than 64 seeds in one lane");
    }
    const uint32_t ns = seed_offsets[L];
    if (ns && (!seed_ranks || !seed_costs)) return fail(BFG_INVALID_PARAMETER, "null seed buffer");
    const int rc = check_ranks(c, seed_ranks, ns, "seed");
    if (rc) return rc;
    req.soff.assign(seed_offsets, seed_offsets + L + 1);
    req.srank.assign(seed_ranks, seed_ranks + ns);
    req.scost.assign(seed_costs, seed_costs + ns);
    req.limit.assign(thresholds, thresholds + L);
    return BFG_SUCCESS;
}

}  // namespace

namespace bfg {
void set_last_error(const std::string &msg) { g_last_error = msg; }
}  // namespace bfg

extern "C" {

const char *bfg_last_error(void) { return g_last_error.c_str(); }

int bfg_device_allocations(uint64_t *count) {
    if (!count) return fail(BFG_INVALID_PARAMETER, "null argument");
    *count = g_device_allocs.load(std::memory_order_relaxed);
    return BFG_SUCCESS;
}

int bfg_ctx_create(const bfg_graph_desc *gd, int device, bfg_ctx **out) {
    if (!gd || !out) return fail(BFG_INVALID_PARAMETER, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(BFG_DEVICE_ERROR, "no HIP device");
    if (device < 0 || device >= ndev) return fail(BFG_INVALID_PARAMETER, "device index out of range");
    std::unique_ptr<bfg_ctx> c(new bfg_ctx());
    c->device = device;
    HIP_TRY(hipSetDevice(device));
    if (const char *e = std::getenv("BFG_STREAMS")) {  // calls in flight at once
        const long v = std::strtol(e, nullptr, 10);
        if (v >= 1 && v <= 16) c->max_works = (uint32_t)v;
    }
    const bfg::CchInput in = input_of(gd);
    std::string err;
    const bfg::LayoutParams lp = layout_params();
    if (!bfg::build_layout(in, lp, c->L, err)) return fail(BFG_INVALID_PARAMETER, err);
    // every slot, position and segment index the kernels derive rests on these
    // properties; refuse the graph rather than launch on a broken layout
    if (!bfg::check_layout(in, c->L, err)) return fail(BFG_INTERNAL_ERROR, err);
    GateTables gt;
    int rc = gate_tables(c->L, gt);
    if (rc) return rc;
    c->n_bt = (uint32_t)(gt.bt_ch.size() / 2);
    WorkLease lease;  // the first Work: its stream carries the uploads
    rc = acquire_work(c.get(), lease);
    if (rc) return rc;
    hipStream_t s = lease.w->stream;
    auto &L = c->L;
    HIP_TRY(c->parent.upload(L.parent, s));
    HIP_TRY(c->pos_parent.upload(L.pos_parent, s));
    HIP_TRY(c->pos_pre.upload(L.pos_pre, s));
    HIP_TRY(c->pos_sub.upload(L.pos_subtree, s));
    HIP_TRY(c->pos_depth.upload(L.pos_depth, s));
    HIP_TRY(c->seg_pos.upload(L.seg_pos, s));
    HIP_TRY(c->in_off.upload(L.in_off, s));
    c->top_in_arcs = L.in_off[L.n_top];
    // ranks descend from parents to children (a parent's rank is higher)
    c->bdep.assign(L.n, 0);
    for (uint32_t r = L.n; r-- > 0;) {
        const uint32_t p = L.pos_of_rank[r], q = L.pos_parent[p];
        if (p >= L.n_top && q != NONE && q >= L.n_top) c->bdep[p] = c->bdep[q] + 1;
    }
    for (size_t l = 0; l + 1 < L.seglvl_seg.size(); ++l) {
        const uint32_t p0 = L.seg_pos[L.seglvl_seg[l]], p1 = L.seg_pos[L.seglvl_seg[l + 1]];
        uint64_t m = 0;
        for (uint32_t p = p0; p < p1; ++p) m += L.in_split[p] - L.in_off[p];
        c->lvl_dom_bytes.push_back(8 * m + 8ull * 64 * (p1 - p0));
    }
    HIP_TRY(c->in_split.upload(L.in_split, s));
    HIP_TRY(c->in_src.upload(L.in_src, s));
    HIP_TRY(c->in_w.upload(L.in_w, s));
    HIP_TRY(c->in_dd.upload(L.in_dd, s));
    HIP_TRY(c->ch_off.upload(L.ch_off, s));
    HIP_TRY(c->ch_pos.upload(L.ch_pos, s));
    HIP_TRY(c->cs_off.upload(L.cs_off, s));
    HIP_TRY(c->cs_src.upload(L.cs_src, s));
    HIP_TRY(c->cs_aoff.upload(L.cs_aoff, s));
    HIP_TRY(c->ca_v.upload(L.ca_v, s));
    HIP_TRY(c->ca_w.upload(L.ca_w, s));
    HIP_TRY(c->up_off.upload(L.up_off, s));
    HIP_TRY(c->up_tgt.upload(L.up_tgt, s));
    HIP_TRY(c->up_w.upload(L.up_w, s));
    HIP_TRY(c->xr_off.upload(L.xr_off, s));
    HIP_TRY(c->xr_run.upload(L.xr_run, s));
    HIP_TRY(c->xw_off.upload(L.xw_off, s));
    HIP_TRY(c->pr_wid.upload(L.pr_wid, s));
    HIP_TRY(c->pw_off.upload(L.pw_off, s));
    HIP_TRY(c->xr_src.upload(L.xr_src, s));
    HIP_TRY(c->xr_w.upload(L.xr_w, s));
    HIP_TRY(c->tb_dn.upload(L.tb_dn, s));
    HIP_TRY(c->tb_up.upload(L.tb_up, s));
    if (L.tb_w.size() < 64) L.tb_w.resize(64, INF);  // tri_apply reads block 0 for absent blocks
    HIP_TRY(c->tb_w.upload(L.tb_w, s));
    HIP_TRY(c->pt_blk.upload(L.pt_blk, s));
    HIP_TRY(c->pt_tgt.upload(L.pt_tgt, s));
    HIP_TRY(c->pr_off.upload(L.pr_off, s));
    HIP_TRY(c->pr_x.upload(L.pr_x, s));
    HIP_TRY(c->pr_w.upload(L.pr_w, s));
    {
        std::vector<uint32_t> w16(L.pr_w.size() / 2);
        auto h16 = [](uint32_t w) { return w == INF ? 0xFFFFu : std::min(w, PUSH_CAP); };
        for (size_t q = 0; q < w16.size(); ++q) w16[q] = h16(L.pr_w[2 * q]) | h16(L.pr_w[2 * q + 1]) << 16;
        HIP_TRY(c->pr_w16.upload(w16, s));
        HIP_TRY(hipStreamSynchronize(s));  // before w16 goes
    }
    HIP_TRY(c->pk_seg.upload(L.pk_seg, s));
    HIP_TRY(c->xr_min.upload(gt.xr_min, s));
    HIP_TRY(c->xr_w16.upload(gt.xr_w16, s));
    HIP_TRY(c->xr_cap.upload(gt.xr_cap, s));
    HIP_TRY(c->cs_min.upload(gt.cs_min, s));
    HIP_TRY(c->pos_bt.upload(gt.pos_bt, s));
    HIP_TRY(c->bt_ch.upload(gt.bt_ch, s));
    HIP_TRY(c->pos_of_rank_d.upload(L.pos_of_rank, s));
    HIP_TRY(hipStreamSynchronize(s));
    c->graph_bytes = 0;
    for (auto *b : {&c->parent, &c->pos_parent, &c->pos_pre, &c->pos_sub, &c->pos_depth, &c->seg_pos, &c->in_off,
                    &c->in_split, &c->in_src, &c->in_w, &c->ch_off, &c->ch_pos, &c->cs_off, &c->cs_src, &c->cs_aoff,
                    &c->ca_w, &c->up_off, &c->up_tgt, &c->up_w, &c->xr_off, &c->xr_run, &c->xw_off, &c->pr_wid,
                    &c->pw_off, &c->xr_src, &c->xr_w, &c->tb_dn, &c->tb_up, &c->tb_w, &c->pt_blk, &c->pt_tgt,
                    &c->pr_off, &c->pr_x, &c->pr_w, &c->pr_w16, &c->pk_seg, &c->xr_min, &c->cs_min, &c->pos_bt, &c->bt_ch, &c->xr_w16, &c->xr_cap,
                    &c->pos_of_rank_d})
        c->graph_bytes += b->n * sizeof(uint32_t);
    c->graph_bytes += (c->ca_v.n + c->in_dd.n) * sizeof(uint16_t);
    // host copies the kernels alone need
    for (auto *v : {&L.in_src, &L.in_w, &L.up_tgt, &L.up_w, &L.cs_src, &L.cs_aoff, &L.ca_w, &L.ch_pos,
                    &L.xr_src, &L.xr_w, &L.tb_dn, &L.tb_up, &L.tb_w, &L.pt_tgt, &L.pr_off, &L.pr_x, &L.pr_w,
                    &L.pk_seg})
        std::vector<uint32_t>().swap(*v);
    std::vector<uint16_t>().swap(L.ca_v);
    std::vector<uint16_t>().swap(L.in_dd);
    // the one-wave-group DOWN plan and the scratch Work of the allocation-free call
    std::shared_ptr<Plan> p1(new Plan());
    rc = build_plan(c.get(), s, 1, *p1);
    if (rc) return rc;
    c->plan1 = std::move(p1);
    uint32_t *sp = nullptr;
    rc = scratch_setup(device, &sp);
    if (rc) return rc;
    rc = scratch_work_init(c.get(), sp);
    if (rc) return rc;
    lease.w = nullptr;  // hand the first Work to the pool by hand: the context is about to be released
    c->idle.push_back(c->works.front().get());
    *out = c.release();
    return BFG_SUCCESS;
}

int bfg_ctx_create_reverse(const bfg_graph_desc *gd, int device, bfg_ctx **out) {
    if (!gd || !out) return fail(BFG_INVALID_PARAMETER, "null argument");
    *out = nullptr;
    const uint32_t n = gd->n_nodes;
    if (!gd->up_offsets || !gd->down_offsets) return fail(BFG_INVALID_PARAMETER, "null offsets");
    for (uint32_t r = 0; r < n; ++r)
        if (gd->up_offsets[r + 1] < gd->up_offsets[r] || gd->down_offsets[r + 1] < gd->down_offsets[r])
            return fail(BFG_INVALID_PARAMETER, "offsets not monotone");
    if (gd->up_offsets[0] != 0 || gd->down_offsets[0] != 0) return fail(BFG_INVALID_PARAMETER, "offsets must start at 0");
    if ((gd->up_offsets[n] && (!gd->up_targets || !gd->up_weights)) ||
        (gd->down_offsets[n] && (!gd->down_targets || !gd->down_weights)))
        return fail(BFG_INVALID_PARAMETER, "null arc buffer");
    Transposed t;
    if (!flip_csr(n, gd->down_offsets, gd->down_targets, gd->down_weights, t.up_off, t.up_tgt, t.up_w) ||
        !flip_csr(n, gd->up_offsets, gd->up_targets, gd->up_weights, t.dn_off, t.dn_tgt, t.dn_w))
        return fail(BFG_INVALID_PARAMETER, "arc target out of range");
    const bfg_graph_desc td{n, t.up_off.data(), t.up_tgt.data(), t.up_w.data(), t.dn_off.data(), t.dn_tgt.data(),
                            t.dn_w.data()};
    return bfg_ctx_create(&td, device, out);
}

void bfg_ctx_destroy(bfg_ctx *c) { delete c; }

int bfg_validate_graph(const bfg_graph_desc *gd, bfg_ctx_info *info) {
    if (!gd || !info) return fail(BFG_INVALID_PARAMETER, "null argument");
    const bfg::CchInput in = input_of(gd);
    bfg::Layout L;
    std::string err;
    const bfg::LayoutParams lp = layout_params();
    if (!bfg::build_layout(in, lp, L, err)) return fail(BFG_INVALID_PARAMETER, err);
    if (!bfg::check_layout(in, L, err)) return fail(BFG_INTERNAL_ERROR, err);
    GateTables gt;
    const int rc = gate_tables(L, gt);
    if (rc) return rc;
    fill_info(L, info);
    return BFG_SUCCESS;
}

int bfg_ctx_get_info(bfg_ctx *c, bfg_ctx_info *info) {
    if (!c || !info) return fail(BFG_INVALID_PARAMETER, "null argument");
    std::shared_lock<std::shared_mutex> cfg(c->cfg);
    fill_info(c->L, info);
    info->lanes = c->lanes;
    std::lock_guard<std::mutex> g(c->pool_mu);
    uint64_t held = 0;
    uint32_t cap = 0;
    for (auto &w : c->works) {
        held += (w->arena.n + w->pathbuf.n) * 4;
        cap = std::max(cap, w->lane_cap);
    }
    info->lane_capacity = c->cap_fixed ? c->lane_cap : cap;
    info->device_bytes = c->graph_bytes + held;
    return BFG_SUCCESS;
}

int bfg_ctx_last_timings(bfg_ctx *c, bfg_timings *t) {
    if (!c || !t) return fail(BFG_INVALID_PARAMETER, "null argument");
    std::lock_guard<std::mutex> g(c->last_mu);
    *t = c->last;
    return BFG_SUCCESS;
}

int bfg_ctx_tune(bfg_ctx *c, uint32_t plan_rows, uint32_t tri_scalar_segs) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if (plan_rows != 0 && (plan_rows < PLAN_MIN_ROWS || plan_rows >= (1u << 24)))
        return fail(BFG_INVALID_PARAMETER, "plan_rows out of range");
    std::unique_lock<std::shared_mutex> cfg(c->cfg);
    HIP_TRY(hipSetDevice(c->device));
    c->plan_rows = plan_rows;
    c->tri_segs = tri_scalar_segs;
    std::lock_guard<std::mutex> g(c->plan_mu);
    c->plans.clear();  // rebuilt by the next calls
    c->plan_lru.clear();
    WorkLease lease;
    int rc = acquire_work(c, lease);
    if (rc) return rc;
    std::shared_ptr<Plan> p1(new Plan());
    rc = build_plan(c, lease.w->stream, 1, *p1);
    if (rc) return rc;
    c->plan1 = std::move(p1);
    return BFG_SUCCESS;
}

int bfg_ctx_configure(bfg_ctx *c, uint32_t lanes, uint32_t lane_capacity) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    std::unique_lock<std::shared_mutex> cfg(c->cfg);
    if (lanes != 0) {
        if (lanes != 16 && lanes != 32 && lanes != 64) return fail(BFG_INVALID_PARAMETER, "lanes must be 16, 32 or 64");
        c->lanes = lanes;
    }
    if (lane_capacity != 0) {
        c->lane_cap = (lane_capacity + 63) / 64 * 64;  // whole wavefronts of lanes
        c->cap_fixed = true;
    } else {
        c->cap_fixed = false;
    }
    return BFG_SUCCESS;
}

int bfg_ctx_set_label_width(bfg_ctx *c, uint32_t bits) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if (bits != 16 && bits != 32) return fail(BFG_INVALID_PARAMETER, "label width must be 16 or 32");
    std::unique_lock<std::shared_mutex> cfg(c->cfg);
    c->label_bits = bits;
    return BFG_SUCCESS;
}

int bfg_ctx_set_concurrency(bfg_ctx *c, uint32_t max_calls) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if (max_calls < 1 || max_calls > 16) return fail(BFG_INVALID_PARAMETER, "max_calls must be in 1..16");
    std::lock_guard<std::mutex> g(c->pool_mu);
    c->max_works = max_calls;  // existing Works beyond the limit stay until the context goes
    return BFG_SUCCESS;
}

int bfg_matrix_bounded(bfg_ctx *c, const uint32_t *src, uint32_t S, const uint32_t *tgt, uint32_t T, uint32_t thr,
                       uint32_t *out) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if ((S && !src) || (T && !tgt) || (S && T && !out)) return fail(BFG_INVALID_PARAMETER, "null buffer");
    int rc = check_ranks(c, src, S, "source");
    if (!rc) rc = check_ranks(c, tgt, T, "target");
    if (rc) return rc;
    if (S == 0 || T == 0) return BFG_SUCCESS;
    CallScope cs(c);
    rc = begin_call(c, cs);
    if (rc) return rc;
    Req req;
    return matrix_impl(c, cs.lease.w, req, src, S, tgt, T, thr, nullptr, out);
}

int bfg_matrix(bfg_ctx *c, const uint32_t *src, uint32_t S, const uint32_t *tgt, uint32_t T, uint32_t *out) {
    return bfg_matrix_bounded(c, src, S, tgt, T, INF, out);
}

int bfg_matrix_device(bfg_ctx *c, const uint32_t *d_src, uint32_t S, const uint32_t *d_tgt, uint32_t T, uint32_t thr,
                      uint32_t *d_out) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if ((S && !d_src) || (T && !d_tgt) || (S && T && !d_out)) return fail(BFG_INVALID_PARAMETER, "null buffer");
    if (S == 0 || T == 0) return BFG_SUCCESS;
    CallScope cs(c);
    int rc = begin_call(c, cs);
    if (rc) return rc;
    Work *w = cs.lease.w;
    w->h_src.resize(S);
    w->h_tgt.resize(T);
    HIP_TRY(hipMemcpy(w->h_src.data(), d_src, (size_t)S * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(w->h_tgt.data(), d_tgt, (size_t)T * 4, hipMemcpyDeviceToHost));
    rc = check_ranks(c, w->h_src.data(), S, "source");
    if (!rc) rc = check_ranks(c, w->h_tgt.data(), T, "target");
    if (rc) return rc;
    Req req;
    return matrix_impl(c, w, req, w->h_src.data(), S, w->h_tgt.data(), T, thr, d_out, nullptr);
}

int bfg_matrix_multi(bfg_ctx *c, uint32_t n_req, const uint32_t *src_offsets, const uint32_t *sources,
                     const uint32_t *tgt_offsets, const uint32_t *targets, const uint32_t *thresholds,
                     uint32_t *out) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if (n_req == 0) return BFG_SUCCESS;
    if (!src_offsets || !tgt_offsets || !thresholds) return fail(BFG_INVALID_PARAMETER, "null buffer");
    if (src_offsets[0] != 0 || tgt_offsets[0] != 0) return fail(BFG_INVALID_PARAMETER, "offsets must start at 0");
    uint64_t cells = 0;
    for (uint32_t r = 0; r < n_req; ++r) {
        if (src_offsets[r + 1] < src_offsets[r] || tgt_offsets[r + 1] < tgt_offsets[r])
            return fail(BFG_INVALID_PARAMETER, "offsets not monotone");
        cells += (uint64_t)(src_offsets[r + 1] - src_offsets[r]) * (tgt_offsets[r + 1] - tgt_offsets[r]);
    }
    const uint32_t S = src_offsets[n_req], T = tgt_offsets[n_req];
    if ((S && !sources) || (T && !targets) || (cells && !out)) return fail(BFG_INVALID_PARAMETER, "null buffer");
    int rc = check_ranks(c, sources, S, "source");
    if (!rc) rc = check_ranks(c, targets, T, "target");
    if (rc) return rc;
    if (cells == 0) return BFG_SUCCESS;
    CallScope cs(c);
    rc = begin_call(c, cs);
    if (rc) return rc;
    DevBuf<uint32_t> d_out;
    HIP_TRY(d_out.alloc(cells));
    rc = matrix_multi(c, cs.lease.w, n_req, src_offsets, sources, tgt_offsets, targets, thresholds, d_out.p);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(out, d_out.p, cells * 4, hipMemcpyDeviceToHost));
    return BFG_SUCCESS;
}

int bfg_matrix_device_noalloc(bfg_ctx *c, const uint32_t *d_src, uint32_t S, const uint32_t *d_tgt, uint32_t T,
                              uint32_t thr, uint32_t *d_out) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if ((S && !d_src) || (T && !d_tgt) || (S && T && !d_out)) return fail(BFG_INVALID_PARAMETER, "null buffer");
    if (S > NOALLOC_LANES)
        return fail(BFG_NOT_SUPPORTED, "the allocation-free matrix takes at most 64 sources per call");
    if (T > NOALLOC_MAX_TARGETS)
        return fail(BFG_NOT_SUPPORTED, "the allocation-free matrix takes at most 262144 targets per call");
    if (S == 0 || T == 0) return BFG_SUCCESS;
    std::shared_lock<std::shared_mutex> cfg(c->cfg);
    HIP_TRY(hipSetDevice(c->device));
    if (!c->scratch_work)
        return fail(BFG_NOT_SUPPORTED, "the graph's tables do not fit the scratch array of the allocation-free call");
    std::lock_guard<std::mutex> g(g_scratch_mu[c->device]);
    Work *w = c->scratch_work.get();
    w->h_src.resize(S);
    w->h_tgt.resize(T);
    HIP_TRY(hipMemcpy(w->h_src.data(), d_src, (size_t)S * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(w->h_tgt.data(), d_tgt, (size_t)T * 4, hipMemcpyDeviceToHost));
    int rc = check_ranks(c, w->h_src.data(), S, "source");
    if (!rc) rc = check_ranks(c, w->h_tgt.data(), T, "target");
    if (rc) return rc;
    rc = matrix_impl(c, w, w->req, w->h_src.data(), S, w->h_tgt.data(), T, thr, d_out, nullptr);
    const hipError_t e = hipStreamSynchronize(w->stream);  // the next call may reuse the array
    if (!rc && e != hipSuccess) return fail(BFG_DEVICE_ERROR, hipGetErrorString(e));
    return rc;
}

int bfg_ctx_pair_reverse(bfg_ctx *c, bfg_ctx *rev) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    std::unique_lock<std::shared_mutex> g(c->cfg);
    if (!rev) {  // unpair
        if (c->rev) c->rev->rev_of = nullptr;
        c->rev = nullptr;
        return BFG_SUCCESS;
    }
    if (rev == c) return fail(BFG_INVALID_PARAMETER, "a context cannot be its own reverse");
    if (rev->device != c->device) return fail(BFG_INVALID_PARAMETER, "the reverse context is on another device");
    if (rev->rev || c->rev_of) return fail(BFG_INVALID_PARAMETER, "a pair's reverse context cannot itself be paired");
    if (rev->rev_of && rev->rev_of != c) return fail(BFG_INVALID_PARAMETER, "the reverse context is paired already");
    const bfg::Layout &A = c->L, &B = rev->L;
    // the transposed graph has the same finite pattern, so the same
    // elimination tree and positions; its UP arcs are this graph's DOWN arcs
    // reversed, position by position
    if (A.n != B.n || A.n_top != B.n_top || A.pos_of_rank != B.pos_of_rank || A.seg_pos != B.seg_pos ||
        A.m_up != B.m_down || A.m_down != B.m_up)
        return fail(BFG_INVALID_PARAMETER, "the reverse context is not over the transposed graph");
    for (uint32_t p = 0; p < A.n; ++p)
        if (A.in_off[p + 1] - A.in_off[p] != B.up_off[p + 1] - B.up_off[p] ||
            B.in_off[p + 1] - B.in_off[p] != A.up_off[p + 1] - A.up_off[p])
            return fail(BFG_INVALID_PARAMETER, "the reverse context is not over the transposed graph");
    for (bfg_ctx *x : {c, rev}) upward_list_tables(x);
    if (c->rev) c->rev->rev_of = nullptr;
    c->rev = rev;
    rev->rev_of = c;
    return BFG_SUCCESS;
}

int bfg_table_bucket(bfg_ctx *c, const uint32_t *src, uint32_t S, const uint32_t *tgt, uint32_t T, uint32_t thr,
                     uint32_t *out) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if ((S && !src) || (T && !tgt) || (S && T && !out)) return fail(BFG_INVALID_PARAMETER, "null buffer");
    int rc = check_ranks(c, src, S, "source");
    if (!rc) rc = check_ranks(c, tgt, T, "target");
    if (rc) return rc;
    if (!c->rev) return fail(BFG_NOT_SUPPORTED, "no reverse context is paired (bfg_ctx_pair_reverse)");
    if (S == 0 || T == 0) return BFG_SUCCESS;
    return bucket_matrix(c, src, S, tgt, T, thr, nullptr, out);
}

int bfg_table_bucket_device(bfg_ctx *c, const uint32_t *d_src, uint32_t S, const uint32_t *d_tgt, uint32_t T,
                            uint32_t thr, uint32_t *d_out) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if ((S && !d_src) || (T && !d_tgt) || (S && T && !d_out)) return fail(BFG_INVALID_PARAMETER, "null buffer");
    if (!c->rev) return fail(BFG_NOT_SUPPORTED, "no reverse context is paired (bfg_ctx_pair_reverse)");
    if (S == 0 || T == 0) return BFG_SUCCESS;
    HIP_TRY(hipSetDevice(c->device));
    std::vector<uint32_t> hs(S), ht(T);
    HIP_TRY(hipMemcpy(hs.data(), d_src, (size_t)S * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ht.data(), d_tgt, (size_t)T * 4, hipMemcpyDeviceToHost));
    int rc = check_ranks(c, hs.data(), S, "source");
    if (!rc) rc = check_ranks(c, ht.data(), T, "target");
    if (rc) return rc;
    return bucket_matrix(c, hs.data(), S, ht.data(), T, thr, d_out, nullptr);
}

int bfg_query_batch(bfg_ctx *c, const uint32_t *src, uint32_t S, uint32_t thr, uint32_t *dist) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if (S && (!src || !dist)) return fail(BFG_INVALID_PARAMETER, "null buffer");
    int rc = check_ranks(c, src, S, "source");
    if (rc) return rc;
    if (S == 0) return BFG_SUCCESS;
    CallScope cs(c);
    rc = begin_call(c, cs);
    if (rc) return rc;
    Req req;
    single_source_req(req, src, S, thr);
    return dense_rows(c, cs.lease.w, req, dist, nullptr);
}

int bfg_isochrone(bfg_ctx *c, const uint32_t *origins, const uint32_t *limits, uint32_t L, uint64_t *bits,
                  uint32_t *times) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if (L && (!origins || !limits || !bits)) return fail(BFG_INVALID_PARAMETER, "null buffer");
    int rc = check_ranks(c, origins, L, "origin");
    if (rc) return rc;
    if (L == 0) return BFG_SUCCESS;
    CallScope cs(c);
    rc = begin_call(c, cs);
    if (rc) return rc;
    Req req;
    single_source_req(req, origins, L, INF);
    req.limit.assign(limits, limits + L);
    return dense_rows(c, cs.lease.w, req, times, bits);
}

int bfg_phast_seeded(bfg_ctx *c, const uint32_t *seed_offsets, const uint32_t *seed_ranks, const uint32_t *seed_costs,
                     uint32_t L, const uint32_t *thresholds, uint32_t *dist) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if (L && (!seed_offsets || !thresholds || !dist)) return fail(BFG_INVALID_PARAMETER, "null buffer");
    if (L == 0) return BFG_SUCCESS;
    Req req;
    int rc = seeded_req(c, req, seed_offsets, seed_ranks, seed_costs, L, thresholds);
    if (rc) return rc;
    CallScope cs(c);
    rc = begin_call(c, cs);
    if (rc) return rc;
    return dense_rows(c, cs.lease.w, req, dist, nullptr);
}

int bfg_isochrone_lists(bfg_ctx *c, const uint32_t *seed_offsets, const uint32_t *seed_ranks,
                        const uint32_t *seed_costs, uint32_t L, const uint32_t *thresholds, uint64_t *reach_bits,
                        uint64_t *list_start, uint32_t *list_len, uint32_t *pairs, uint64_t capacity,
                        uint64_t *total) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if (L && (!seed_offsets || !thresholds || !list_start || !list_len)) return fail(BFG_INVALID_PARAMETER, "null buffer");
    if (capacity && !pairs) return fail(BFG_INVALID_PARAMETER, "null pair buffer");
    if (total) *total = 0;
    if (L == 0) return BFG_SUCCESS;
    Req req;
    int rc = seeded_req(c, req, seed_offsets, seed_ranks, seed_costs, L, thresholds);
    if (rc) return rc;
    CallScope cs(c);
    rc = begin_call(c, cs);
    if (rc) return rc;
    return iso_lists(c, cs.lease.w, req, reach_bits, nullptr, list_start, list_len, pairs, nullptr, capacity, total);
}

int bfg_isochrone_lists_device(bfg_ctx *c, const uint32_t *seed_offsets, const uint32_t *seed_ranks,
                               const uint32_t *seed_costs, uint32_t L, const uint32_t *thresholds,
                               uint64_t *d_reach_bits, uint64_t *list_start, uint32_t *list_len, uint32_t *d_pairs,
                               uint64_t capacity, uint64_t *total) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if (L && (!seed_offsets || !thresholds || !list_start || !list_len || !d_reach_bits))
        return fail(BFG_INVALID_PARAMETER, "null buffer");
    if (capacity && !d_pairs) return fail(BFG_INVALID_PARAMETER, "null pair buffer");
    if (total) *total = 0;
    if (L == 0) return BFG_SUCCESS;
    Req req;
    int rc = seeded_req(c, req, seed_offsets, seed_ranks, seed_costs, L, thresholds);
    if (rc) return rc;
    CallScope cs(c);
    rc = begin_call(c, cs);
    if (rc) return rc;
    return iso_lists(c, cs.lease.w, req, nullptr, d_reach_bits, list_start, list_len, nullptr, d_pairs, capacity,
                     total);
}

int bfg_ctx_attach_length(bfg_ctx *c, const bfg_graph_desc *gd, const uint32_t *up_len, const uint32_t *down_len) {
    if (!c || !gd) return fail(BFG_INVALID_PARAMETER, "null argument");
    if (gd->n_nodes != c->L.n) return fail(BFG_INVALID_PARAMETER, "graph differs from the context's");
    std::unique_lock<std::shared_mutex> cfg(c->cfg);
    HIP_TRY(hipSetDevice(c->device));
    WorkLease lease;
    int rc = acquire_work(c, lease);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(c->t2_mu);
    c->t2.reset();
    return attach_length(c, lease.w->stream, gd, up_len, down_len);
}

int bfg_phast_seeded_2ch(bfg_ctx *c, const uint32_t *seed_offsets, const uint32_t *seed_ranks,
                         const uint32_t *seed_times, const uint32_t *seed_lens, uint32_t L,
                         const uint32_t *thresholds, uint32_t *times, uint32_t *lens) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if (L && (!seed_offsets || !thresholds || !times || !lens)) return fail(BFG_INVALID_PARAMETER, "null buffer");
    CallScope cs(c);
    int rc = begin_call(c, cs);
    if (rc) return rc;
    std::lock_guard<std::mutex> g2(c->t2_mu);
    if (!c->t2) return fail(BFG_INVALID_PARAMETER, "no length weights attached (bfg_ctx_attach_length)");
    if (L == 0) return BFG_SUCCESS;
    if (seed_offsets[0] != 0) return fail(BFG_INVALID_PARAMETER, "seed_offsets[0] must be 0");
    for (uint32_t i = 0; i < L; ++i)
        if (seed_offsets[i + 1] < seed_offsets[i]) return fail(BFG_INVALID_PARAMETER, "seed_offsets not monotone");
    const uint32_t ns = seed_offsets[L];
    if (ns && (!seed_ranks || !seed_times || !seed_lens)) return fail(BFG_INVALID_PARAMETER, "null seed buffer");
    rc = check_ranks(c, seed_ranks, ns, "seed");
    if (rc) return rc;
    std::vector<uint32_t> soff, spos;
    std::vector<uint64_t> slab;
    rc = seeds_2ch(c->L, seed_offsets, seed_ranks, seed_times, seed_lens, L, soff, spos, slab);
    if (rc) return rc;
    TwoChannel &t = *c->t2;
    const uint32_t n = c->L.n;
    hipStream_t s = cs.lease.w->stream;
    return phast_2ch(c, s, soff, spos, slab, thresholds, L, true, BottomSet{},
                     [&](uint32_t l0, uint32_t bl, uint32_t bg) -> int {
        t2_out_kernel<<<dim3((n + 63) / 64, bg), NT, 0, s>>>(n, bl, t.pos_of_rank.p, t.lim.p, t.lab.p, t.o_t.p,
                                                            t.o_l.p);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(times + (size_t)l0 * n, t.o_t.p, (size_t)bl * n * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(lens + (size_t)l0 * n, t.o_l.p, (size_t)bl * n * 4, hipMemcpyDeviceToHost, s));
        return BFG_SUCCESS;
    });
}

int bfg_table_seeded_2ch(bfg_ctx *c, const uint32_t *src_offsets, const uint32_t *src_ranks,
                         const uint32_t *src_part_times, const uint32_t *src_part_lens, uint32_t S,
                         const uint32_t *tgt_offsets, const uint32_t *tgt_ranks, const uint32_t *tgt_part_times,
                         const uint32_t *tgt_part_lens, uint32_t T, uint32_t threshold, uint32_t *times,
                         uint32_t *lens, uint8_t *conflicts) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if ((S && !src_offsets) || (T && !tgt_offsets) || (S && T && (!times || !lens)))
        return fail(BFG_INVALID_PARAMETER, "null buffer");
    CallScope cs(c);
    int rc = begin_call(c, cs);
    if (rc) return rc;
    std::lock_guard<std::mutex> g2(c->t2_mu);
    if (!c->t2) return fail(BFG_INVALID_PARAMETER, "no length weights attached (bfg_ctx_attach_length)");
    if (S == 0 || T == 0) return BFG_SUCCESS;
    if (src_offsets[0] != 0 || tgt_offsets[0] != 0) return fail(BFG_INVALID_PARAMETER, "seed offsets must start at 0");
    for (uint32_t i = 0; i < S; ++i)
        if (src_offsets[i + 1] < src_offsets[i]) return fail(BFG_INVALID_PARAMETER, "src_offsets not monotone");
    for (uint32_t j = 0; j < T; ++j)
        if (tgt_offsets[j + 1] < tgt_offsets[j]) return fail(BFG_INVALID_PARAMETER, "tgt_offsets not monotone");
    const uint32_t ns = src_offsets[S], nt = tgt_offsets[T];
    if ((ns && (!src_ranks || !src_part_times || !src_part_lens)) ||
        (nt && (!tgt_ranks || !tgt_part_times || !tgt_part_lens)))
        return fail(BFG_INVALID_PARAMETER, "null seed buffer");
    rc = check_ranks(c, src_ranks, ns, "source seed");
    if (!rc) rc = check_ranks(c, tgt_ranks, nt, "target seed");
    if (rc) return rc;
    // one field per source, bounded by threshold + its largest time part + the
    // largest target time part (bucket_ch.rs:6400-6415)
    uint32_t max_tgt_part = 0;
    for (uint32_t q = 0; q < nt; ++q) max_tgt_part = std::max(max_tgt_part, tgt_part_times[q]);
    auto sat = [](uint32_t a, uint32_t b) { return a > INF - b ? INF : a + b; };
    std::vector<uint32_t> bound(S);
    for (uint32_t i = 0; i < S; ++i) {
        uint32_t mp = 0;
        for (uint32_t k = src_offsets[i]; k < src_offsets[i + 1]; ++k) mp = std::max(mp, src_part_times[k]);
        bound[i] = threshold == INF ? INF : sat(sat(threshold, mp), max_tgt_part);
    }
    std::vector<uint32_t> soff, spos;
    std::vector<uint64_t> slab;
    rc = seeds_2ch(c->L, src_offsets, src_ranks, src_part_times, src_part_lens, S, soff, spos, slab);
    if (rc) return rc;
    std::vector<uint32_t> tpos(nt);
    for (uint32_t q = 0; q < nt; ++q) tpos[q] = c->L.pos_of_rank[tgt_ranks[q]];
    // the field is read at the target seeds only: their labels depend on their
    // ancestors alone, so the bottom sweep covers the bottom nodes above them
    const bfg::Layout &Ly = c->L;
    std::vector<uint8_t> need(Ly.n - Ly.n_top, 0);
    for (uint32_t q = 0; q < nt; ++q)
        for (uint32_t p = tpos[q]; p != NONE && p >= Ly.n_top && !need[p - Ly.n_top]; p = Ly.pos_parent[p])
            need[p - Ly.n_top] = 1;
    std::vector<uint32_t> bpos, boff(1, 0);
    for (size_t lv = 0; lv + 1 < Ly.botlvl_pos.size(); ++lv) {
        for (uint32_t p = Ly.botlvl_pos[lv]; p < Ly.botlvl_pos[lv + 1]; ++p)
            if (need[p - Ly.n_top]) bpos.push_back(p);
        boff.push_back((uint32_t)bpos.size());
    }
    const std::vector<uint32_t> h_toff(tgt_offsets, tgt_offsets + T + 1), h_trank(tgt_ranks, tgt_ranks + nt),
        h_tpt(tgt_part_times, tgt_part_times + nt), h_tpl(tgt_part_lens, tgt_part_lens + nt),
        h_soff(src_offsets, src_offsets + S + 1), h_srank(src_ranks, src_ranks + ns);
    DevBuf<uint32_t> d_toff, d_tpos, d_trank, d_tpt, d_tpl, d_soff, d_srank, d_t, d_l, d_bpos;
    DevBuf<uint8_t> d_conf;
    hipStream_t st = cs.lease.w->stream;
    const StreamDrain drain{st};  // before the device buffers and host tables above go
    HIP_TRY(d_toff.upload(h_toff, st));
    HIP_TRY(d_tpos.upload(tpos, st));
    HIP_TRY(d_trank.upload(h_trank, st));
    HIP_TRY(d_tpt.upload(h_tpt, st));
    HIP_TRY(d_tpl.upload(h_tpl, st));
    HIP_TRY(d_soff.upload(h_soff, st));
    HIP_TRY(d_srank.upload(h_srank, st));
    HIP_TRY(d_bpos.upload(bpos, st));
    HIP_TRY(d_t.alloc((size_t)S * T));
    HIP_TRY(d_l.alloc((size_t)S * T));
    if (conflicts) HIP_TRY(d_conf.alloc((size_t)S * T));
    HIP_TRY(hipStreamSynchronize(st));  // before the host tables above may go
    rc = phast_2ch(c, st, soff, spos, slab, bound.data(), S, false, BottomSet{&boff, d_bpos.p},
                   [&](uint32_t l0, uint32_t bl, uint32_t bg) -> int {
        // this batch's lanes are sources [l0, l0 + bl): offset the per-source tables
        t2_table_kernel<<<dim3((T + 63) / 64, bg), NT, 0, st>>>(
            c->L.n, T, bl, d_toff.p, d_tpos.p, d_trank.p, d_tpt.p, d_tpl.p, d_soff.p + l0, d_srank.p,
            c->t2->lim.p, threshold, c->t2->lab.p, d_t.p + (size_t)l0 * T, d_l.p + (size_t)l0 * T,
            conflicts ? d_conf.p + (size_t)l0 * T : nullptr);
        HIP_TRY(hipGetLastError());
        return BFG_SUCCESS;
    });
    if (rc) return rc;
    HIP_TRY(hipMemcpy(times, d_t.p, (size_t)S * T * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(lens, d_l.p, (size_t)S * T * 4, hipMemcpyDeviceToHost));
    if (conflicts) HIP_TRY(hipMemcpy(conflicts, d_conf.p, (size_t)S * T, hipMemcpyDeviceToHost));
    return BFG_SUCCESS;
}

int bfg_table_seeded(bfg_ctx *c, const uint32_t *src_offsets, const uint32_t *src_ranks, const uint32_t *src_parts,
                     uint32_t S, const uint32_t *tgt_offsets, const uint32_t *tgt_ranks, const uint32_t *tgt_parts,
                     uint32_t T, uint32_t threshold, uint32_t *out, uint8_t *conflicts) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if ((S && !src_offsets) || (T && !tgt_offsets) || (S && T && !out))
        return fail(BFG_INVALID_PARAMETER, "null buffer");
    if (S == 0 || T == 0) return BFG_SUCCESS;
    CallScope cs(c);
    int rc = begin_call(c, cs);
    if (rc) return rc;
    Work *w = cs.lease.w;
    if (src_offsets[0] != 0 || tgt_offsets[0] != 0) return fail(BFG_INVALID_PARAMETER, "seed offsets must start at 0");
    for (uint32_t i = 0; i < S; ++i) {
        if (src_offsets[i + 1] < src_offsets[i]) return fail(BFG_INVALID_PARAMETER, "src_offsets not monotone");
        if (src_offsets[i + 1] - src_offsets[i] > 64) return fail(BFG_NOT_SUPPORTED, "more than 64 seeds in one source");
    }
    for (uint32_t j = 0; j < T; ++j)
        if (tgt_offsets[j + 1] < tgt_offsets[j]) return fail(BFG_INVALID_PARAMETER, "tgt_offsets not monotone");
    const uint32_t ns = src_offsets[S], nt = tgt_offsets[T];
    if ((ns && (!src_ranks || !src_parts)) || (nt && (!tgt_ranks || !tgt_parts)))
        return fail(BFG_INVALID_PARAMETER, "null seed buffer");
    rc = check_ranks(c, src_ranks, ns, "source seed");
    if (!rc) rc = check_ranks(c, tgt_ranks, nt, "target seed");
    if (rc) return rc;
    // one field per source, bounded by threshold + its largest part + the
    // largest target part (bucket_ch.rs:6738-6752)
    uint32_t max_tgt_part = 0;
    for (uint32_t q = 0; q < nt; ++q) max_tgt_part = std::max(max_tgt_part, tgt_parts[q]);
    auto sat = [](uint32_t a, uint32_t b) { return a > INF - b ? INF : a + b; };
    Req req;
    req.soff.assign(src_offsets, src_offsets + S + 1);
    req.srank.assign(src_ranks, src_ranks + ns);
    req.scost.assign(src_parts, src_parts + ns);
    req.limit.resize(S);
    for (uint32_t i = 0; i < S; ++i) {
        uint32_t mp = 0;
        for (uint32_t k = src_offsets[i]; k < src_offsets[i + 1]; ++k) mp = std::max(mp, src_parts[k]);
        req.limit[i] = threshold == INF ? INF : sat(sat(threshold, mp), max_tgt_part);
    }
    // the field is read at the target seeds' ranks only: they are the
    // matrix targets of the restricted sweep
    std::vector<uint32_t> U(tgt_ranks, tgt_ranks + nt);
    std::sort(U.begin(), U.end());
    U.erase(std::unique(U.begin(), U.end()), U.end());
    if (U.empty()) {  // no target has a seed: every cell is INF, no conflicts
        std::fill(out, out + (size_t)S * T, INF);
        if (conflicts) std::memset(conflicts, 0, (size_t)S * T);
        return BFG_SUCCESS;
    }
    rc = stage_target_paths(c, w, U.data(), (uint32_t)U.size());
    if (rc) return rc;
    std::vector<uint32_t> tlab(nt);
    for (uint32_t q = 0; q < nt; ++q)
        tlab[q] = w->tgt_pos_h[std::lower_bound(U.begin(), U.end(), tgt_ranks[q]) - U.begin()];
    DevBuf<uint32_t> d_toff, d_tlab, d_trank, d_tpart, d_out;
    DevBuf<uint8_t> d_conf;
    hipStream_t st = w->stream;
    const std::vector<uint32_t> h_toff(tgt_offsets, tgt_offsets + T + 1), h_trank(tgt_ranks, tgt_ranks + nt),
        h_tpart(tgt_parts, tgt_parts + nt);
    const StreamDrain drain{st};
    HIP_TRY(d_toff.upload(h_toff, st));
    HIP_TRY(d_tlab.upload(tlab, st));
    HIP_TRY(d_trank.upload(h_trank, st));
    HIP_TRY(d_tpart.upload(h_tpart, st));
    HIP_TRY(d_out.alloc((size_t)S * T));
    if (conflicts) HIP_TRY(d_conf.alloc((size_t)S * T));
    HIP_TRY(hipStreamSynchronize(st));  // before the host tables above may go
    PathsOn paths_on(w);  // restrict the bottom sweep for this call only
    rc = run_all(c, w, req, 0, [&](uint32_t, uint32_t nl, const std::vector<uint32_t> &, float *oms) -> int {
        const uint32_t groups = (nl + c->lanes - 1) / c->lanes;
        const dim3 grid((T + 63) / 64, groups);
        HIP_TRY(hipEventRecord(w->ev[0], st));
        switch (c->lanes) {
            case 16:
                hipLaunchKernelGGL(table_gather_kernel<16>, grid, dim3(NT), 0, st, w->arena.p, label_stride(c, w), T,
                                   d_toff.p, d_tlab.p, d_trank.p, d_tpart.p, w->seed_off.p, w->seed_rank.p,
                                   w->lane_limit.p, nl, w->row_map.p, threshold, d_out.p, d_conf.p);
                break;
            case 64:
                hipLaunchKernelGGL(table_gather_kernel<64>, grid, dim3(NT), 0, st, w->arena.p, label_stride(c, w), T,
                                   d_toff.p, d_tlab.p, d_trank.p, d_tpart.p, w->seed_off.p, w->seed_rank.p,
                                   w->lane_limit.p, nl, w->row_map.p, threshold, d_out.p, d_conf.p);
                break;
            default:
                hipLaunchKernelGGL(table_gather_kernel<32>, grid, dim3(NT), 0, st, w->arena.p, label_stride(c, w), T,
                                   d_toff.p, d_tlab.p, d_trank.p, d_tpart.p, w->seed_off.p, w->seed_rank.p,
                                   w->lane_limit.p, nl, w->row_map.p, threshold, d_out.p, d_conf.p);
                break;
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(w->ev[1], st));
        HIP_TRY(hipEventSynchronize(w->ev[1]));
        float a = 0;
        HIP_TRY(hipEventElapsedTime(&a, w->ev[0], w->ev[1]));
        *oms += a;
        return BFG_SUCCESS;
    });
    if (rc) return rc;
    HIP_TRY(hipMemcpy(out, d_out.p, (size_t)S * T * 4, hipMemcpyDeviceToHost));
    if (conflicts) HIP_TRY(hipMemcpy(conflicts, d_conf.p, (size_t)S * T, hipMemcpyDeviceToHost));
    return BFG_SUCCESS;
}

}  // extern "C"
