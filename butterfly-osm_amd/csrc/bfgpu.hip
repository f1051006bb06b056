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
//   fill_top     top-region labels of the batch := INF
//   bottom_walk  one wavefront per lane walks the bottom part of its root path
//                (depth-indexed slots), relaxing UP arcs into bottom ancestors
//                (slots) and top ancestors (atomic min into the arena)
//   up_tile      per (group, segment), deepest segment level first, as a
//                register tile (8 nodes x KL/8 lanes per thread): what lower
//                levels pushed, intra-segment UP arcs as 8x8 blocks from the
//                last octet to the first, then the UP arcs leaving the
//                segment as ELL rows reduced per target block and pushed with
//                one atomic min per (target, lane)
// Downward phase — the rank-descending DOWN scan (batched_phast.rs:205-250,
// :538-590) scheduled by the elimination tree:
//   down_tile    per (group, segment) of a segment level, as the same register
//                tile seeded with the upward labels: DOWN arcs from outside the
//                segment as ELL rows (min-plus product of the sources' labels
//                and the tile's weights), then the intra-segment blocks from the
//                first octet to the last
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
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/bfgpu.h"
#include "layout.hpp"

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
__global__ void fill_top_kernel(uint32_t *__restrict__ arena, uint64_t group_stride, uint64_t top_words,
                                uint32_t groups) {
    const uint64_t total = top_words * groups;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t g = i / top_words, r = i - g * top_words;
        arena[g * group_stride + r] = INF;
    }
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
// Dense tiles of the top region.  A workgroup of 256 threads owns one
// (segment, group) pair: thread t holds nodes 8*(t/8) .. +8 of the segment
// (its octet) for lanes TK*(t%8) .. +TK of the group, TK = KL/8, as an 8 x TK
// register tile.  Wavefront w thus covers the 64 nodes of wave block w.
// ---------------------------------------------------------------------------
constexpr int TILE_THREADS = 256;

template <int N>
__device__ __forceinline__ void load_u32(const uint32_t *__restrict__ p, uint32_t (&v)[N]) {
    static_assert(N == 2 || N % 4 == 0, "vector width");
    if constexpr (N == 2) {
        const uint2 x = *reinterpret_cast<const uint2 *>(p);
        v[0] = x.x;
        v[1] = x.y;
    } else {
#pragma unroll
        for (int q = 0; q < N / 4; ++q) {
            const uint4 x = reinterpret_cast<const uint4 *>(p)[q];
            v[4 * q] = x.x;
            v[4 * q + 1] = x.y;
            v[4 * q + 2] = x.z;
            v[4 * q + 3] = x.w;
        }
    }
}

template <int N>
__device__ __forceinline__ void store_u32(uint32_t *__restrict__ p, const uint32_t (&v)[N]) {
    if constexpr (N == 2) {
        *reinterpret_cast<uint2 *>(p) = make_uint2(v[0], v[1]);
    } else {
#pragma unroll
        for (int q = 0; q < N / 4; ++q)
            reinterpret_cast<uint4 *>(p)[q] = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    }
}

// acc[j][k] = min(acc[j][k], lab[key][k] + w[j]) over the ELL rows [r0, r1):
// this thread reads entry 8r + o (its octet o), the key's TK labels of its
// lanes and the entry's 8 weights.  Two rows per step: min(acc, x0, x1) is
// one v_min3.  key_base offsets the keys (segment-local tails).
template <int TK>
__device__ __forceinline__ void ell_rows(uint32_t (&acc)[8][TK], uint32_t r0, uint32_t r1, uint32_t o,
                                         uint32_t lane0, const uint32_t *__restrict__ key,
                                         const uint32_t *__restrict__ wts, const uint32_t *__restrict__ lab,
                                         uint32_t key_base, uint32_t KL) {
    uint32_t r = r0;
    for (; r + 2 <= r1; r += 2) {
        const size_t e0 = (size_t)r * 8 + o, e1 = e0 + 8;
        const uint32_t u0 = key[e0] + key_base, u1 = key[e1] + key_base;
        uint32_t w0[8], w1[8], a0[TK], a1[TK];
        load_u32<8>(wts + e0 * 8, w0);
        load_u32<8>(wts + e1 * 8, w1);
        load_u32<TK>(lab + (size_t)u0 * KL + lane0, a0);
        load_u32<TK>(lab + (size_t)u1 * KL + lane0, a1);
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int k = 0; k < TK; ++k)
                acc[j][k] = min(acc[j][k], min(sat_add(a0[k], w0[j]), sat_add(a1[k], w1[j])));
    }
    if (r < r1) {
        const size_t e0 = (size_t)r * 8 + o;
        const uint32_t u0 = key[e0] + key_base;
        uint32_t w0[8], a0[TK];
        load_u32<8>(wts + e0 * 8, w0);
        load_u32<TK>(lab + (size_t)u0 * KL + lane0, a0);
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int k = 0; k < TK; ++k) acc[j][k] = min(acc[j][k], sat_add(a0[k], w0[j]));
    }
}

// Intra-segment relaxation over 8x8 blocks, octet by octet (DOWN: first to
// last, UP: last to first).  Step a: the threads of octet a finish it with
// its diagonal block (a sequential pass over its 8 nodes in registers) and
// publish its labels in LDS; after one barrier every octet still to come
// applies block (a, its octet).  xb is double-buffered, so one barrier per
// step suffices.
template <int KL, bool DOWN>
__device__ __forceinline__ void tile_triangle(uint32_t (&acc)[8][KL / 8], uint32_t (*xb)[8][KL], uint32_t len,
                                              uint32_t ov, uint32_t lane0, const uint32_t *__restrict__ tab,
                                              const uint32_t *__restrict__ tw) {
    constexpr int TK = KL / 8;
    const uint32_t nb = (len + 7) / 8;
    for (uint32_t step = 0; step < nb; ++step) {
        const uint32_t a = DOWN ? step : nb - 1 - step;
        uint32_t(*x)[KL] = xb[step & 1];
        if (ov == a) {
            const uint32_t k = tab[32 * a + a];
            if (k != NONE) {
                const uint32_t *W = tw + (size_t)k * 64;
#pragma unroll
                for (int ii = 0; ii < 8; ++ii) {
                    const int i = DOWN ? ii : 7 - ii;
                    uint32_t w[8];
                    load_u32<8>(W + 8 * i, w);
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        if (DOWN ? j <= i : j >= i) continue;
#pragma unroll
                        for (int q = 0; q < TK; ++q) acc[j][q] = min(acc[j][q], sat_add(acc[i][q], w[j]));
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) store_u32<TK>(&x[i][lane0], acc[i]);
        }
        __syncthreads();
        if ((DOWN ? ov > a : ov < a) && ov < nb) {
            const uint32_t k = tab[32 * a + ov];
            if (k != NONE) {
                const uint32_t *W = tw + (size_t)k * 64;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    uint32_t w[8], d[TK];
                    load_u32<8>(W + 8 * i, w);
                    load_u32<TK>(&x[i][lane0], d);
#pragma unroll
                    for (int j = 0; j < 8; ++j)
#pragma unroll
                        for (int q = 0; q < TK; ++q) acc[j][q] = min(acc[j][q], sat_add(d[q], w[j]));
                }
            }
        }
    }
}

template <int KL>
__device__ __forceinline__ void tile_load(uint32_t (&acc)[8][KL / 8], const uint32_t *__restrict__ lab, uint32_t p0,
                                          uint32_t len, uint32_t ov, uint32_t lane0) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t i = ov * 8 + j;
        if (i < len) {
            load_u32<KL / 8>(lab + (size_t)(p0 + i) * KL + lane0, acc[j]);
        } else {
#pragma unroll
            for (int q = 0; q < KL / 8; ++q) acc[j][q] = INF;
        }
    }
}

template <int KL>
__device__ __forceinline__ void tile_store(const uint32_t (&acc)[8][KL / 8], uint32_t *__restrict__ lab, uint32_t p0,
                                           uint32_t len, uint32_t ov, uint32_t lane0) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t i = ov * 8 + j;
        if (i < len) store_u32<KL / 8>(lab + (size_t)(p0 + i) * KL + lane0, acc[j]);
    }
}

// down_tile: group blockIdx.x, segment seg0 + blockIdx.y of one segment level.
// The tile starts from the upward labels, takes the DOWN arcs from outside the
// segment (ELL rows of its wave block) and then the intra-segment blocks.
template <int KL>
__global__ __launch_bounds__(TILE_THREADS) void down_tile_kernel(
    uint32_t seg0, const uint32_t *__restrict__ seg_pos, const uint32_t *__restrict__ xr_off,
    const uint32_t *__restrict__ xr_src, const uint32_t *__restrict__ xr_w, const uint32_t *__restrict__ tb_dn,
    const uint32_t *__restrict__ tb_w, uint32_t *__restrict__ arena, uint32_t n) {
    constexpr int TK = KL / 8;
    __shared__ __align__(16) uint32_t xb[2][8][KL];
    const uint32_t g = blockIdx.x, s = seg0 + blockIdx.y;
    const uint32_t ov = threadIdx.x / 8, lane0 = (threadIdx.x % 8) * TK;
    const uint32_t o = ov % 8, wb = ov / 8;
    const uint32_t p0 = seg_pos[s], len = seg_pos[s + 1] - p0;
    uint32_t *lab = arena + (size_t)g * n * KL;
    uint32_t acc[8][TK];
    tile_load<KL>(acc, lab, p0, len, ov, lane0);
    ell_rows<TK>(acc, xr_off[4 * s + wb], xr_off[4 * s + wb + 1], o, lane0, xr_src, xr_w, lab, 0, KL);
    tile_triangle<KL, true>(acc, xb, len, ov, lane0, tb_dn + (size_t)s * 1024, tb_w);
    tile_store<KL>(acc, lab, p0, len, ov, lane0);
}

// up_tile: group blockIdx.x, segment seg0 + blockIdx.y of one segment level
// (levels run deepest first).  The tile holds what lower levels pushed; a
// group none of whose lanes reaches the segment stops here.  Intra-segment
// UP blocks run last to first; then, if push_out, the UP arcs leaving the
// segment: each wavefront takes target blocks of 64 targets, reduces their
// ELL rows in registers and pushes one atomic min per finite (target, lane).
template <int KL>
__global__ __launch_bounds__(TILE_THREADS) void up_tile_kernel(
    uint32_t seg0, int push_out, const uint32_t *__restrict__ seg_pos, const uint32_t *__restrict__ tb_up,
    const uint32_t *__restrict__ tb_w, const uint32_t *__restrict__ pt_blk, const uint32_t *__restrict__ pt_tgt,
    const uint32_t *__restrict__ pr_off, const uint32_t *__restrict__ pr_x, const uint32_t *__restrict__ pr_w,
    uint32_t *__restrict__ arena, uint32_t n) {
    constexpr int TK = KL / 8;
    __shared__ __align__(16) uint32_t xb[2][8][KL];
    const uint32_t g = blockIdx.x, s = seg0 + blockIdx.y;
    const uint32_t ov = threadIdx.x / 8, lane0 = (threadIdx.x % 8) * TK;
    const uint32_t o = ov % 8, wb = ov / 8;
    const uint32_t p0 = seg_pos[s], len = seg_pos[s + 1] - p0;
    uint32_t *lab = arena + (size_t)g * n * KL;
    uint32_t acc[8][TK];
    tile_load<KL>(acc, lab, p0, len, ov, lane0);
    int any = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
        for (int q = 0; q < TK; ++q) any |= acc[j][q] != INF;
    if (!__syncthreads_or(any)) return;  // no lane of this group reaches the segment
    tile_triangle<KL, false>(acc, xb, len, ov, lane0, tb_up + (size_t)s * 1024, tb_w);
    tile_store<KL>(acc, lab, p0, len, ov, lane0);
    if (!push_out) return;
    __syncthreads();  // the segment's final labels, written above, are read below by every wavefront
    for (uint32_t kb = pt_blk[s] + wb; kb < pt_blk[s + 1]; kb += TILE_THREADS / 64) {
        uint32_t red[8][TK];
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int q = 0; q < TK; ++q) red[j][q] = INF;
        ell_rows<TK>(red, pr_off[kb], pr_off[kb + 1], o, lane0, pr_x, pr_w, lab, p0, KL);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint32_t t = pt_tgt[(size_t)kb * 64 + o * 8 + j];
            if (t == NONE) continue;
#pragma unroll
            for (int q = 0; q < TK; ++q)
                if (red[j][q] != INF) atomicMin(lab + (size_t)t * KL + lane0 + q, red[j][q]);
        }
    }
}

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
    const uint32_t *__restrict__ pathbuf, uint32_t D1, uint32_t n_lanes, uint32_t *__restrict__ arena, uint32_t n) {
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
    __syncthreads();
    relax_sources<KL, NS>(L, lab, cs_off[c], cs_off[c + 1], cs_src, cs_aoff, ca_v, ca_w, slot, lane);
    __syncthreads();
    for (uint32_t i = slot; i < len; i += NS) lab[(size_t)ch_pos[i0 + i] * KL + lane] = L[i * KL + lane];
}

// bottom_pull: bottom nodes [p0, p1) of one height level, group blockIdx.y:
// the chunk result min the DOWN arcs from bottom sources (higher levels).
template <int KL>
__global__ __launch_bounds__(NT) void bottom_pull_kernel(uint32_t p0, uint32_t p1, const uint32_t *__restrict__ in_off,
                                                         const uint32_t *__restrict__ in_split,
                                                         const uint32_t *__restrict__ in_src,
                                                         const uint32_t *__restrict__ in_w,
                                                         uint32_t *__restrict__ arena, uint32_t n) {
    constexpr int NS = NT / KL;
    const uint32_t lane = threadIdx.x % KL, slot = slot_id<KL>();
    const uint32_t p = p0 + blockIdx.x * NS + slot;
    if (p >= p1) return;
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
    const uint32_t *__restrict__ lane_limit, uint32_t n_lanes, const uint32_t *__restrict__ row_map,
    uint32_t *__restrict__ out, uint64_t ld, unsigned long long *__restrict__ bits, uint64_t bits_ld) {
    __shared__ uint32_t tile[64][KL + 1];
    const uint32_t t0 = blockIdx.x * 64;
    const uint32_t g = blockIdx.y;
    const uint32_t *lab = arena + (size_t)g * n * KL;
    for (uint32_t idx = threadIdx.x; idx < 64 * KL; idx += NT) {
        const uint32_t ti = idx / KL, lane = idx % KL;
        const uint32_t t = t0 + ti;
        uint32_t v = INF;
        const uint32_t gl = g * KL + lane;
        if (t < T && gl < n_lanes) {
            v = lab[(size_t)tgt_pos[t] * KL + lane];
            if (v > lane_limit[gl]) v = INF;
        }
        tile[ti][lane] = v;
    }
    __syncthreads();
    if (out) {
        for (uint32_t idx = threadIdx.x; idx < 64 * KL; idx += NT) {
            const uint32_t lane = idx / 64, ti = idx % 64;
            const uint32_t gl = g * KL + lane, t = t0 + ti;
            if (gl < n_lanes && t < T) {
                const uint64_t row = row_map ? row_map[gl] : gl;
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
            if (ti == 0 && gl < n_lanes) {
                const uint64_t row = row_map ? row_map[gl] : gl;
                bits[row * bits_ld + (t0 / 64)] = word;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// host context
// ---------------------------------------------------------------------------
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e == hipSuccess) n = count;
        return e;
    }
    hipError_t upload(const std::vector<T> &v, hipStream_t s) {
        hipError_t e = alloc(v.size());
        if (e != hipSuccess) return e;
        if (v.empty()) return hipSuccess;
        return hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
    }
};

constexpr int SEG = 256;  // nodes of a segment tile and of a bottom chunk

}  // namespace

struct bfg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    bfg::Layout L;
    uint32_t lanes = 32;        // KL: sources per lane group (16, 32 or 64)
    uint32_t lane_cap = 0;      // lanes per batch (multiple of KL)
    bool cap_fixed = false;
    // resident graph
    DevBuf<uint32_t> parent, pos_parent, pos_pre, pos_sub, pos_depth, seg_pos;
    DevBuf<uint32_t> in_off, in_split, in_src, in_w;
    DevBuf<uint32_t> ch_off, ch_pos, cs_off, cs_src, cs_aoff, ca_w;
    DevBuf<uint16_t> ca_v;
    DevBuf<uint32_t> up_off, up_tgt, up_w;
    DevBuf<uint32_t> xr_off, xr_src, xr_w, tb_dn, tb_up, tb_w, pt_blk, pt_tgt, pr_off, pr_x, pr_w;
    // arenas (per batch)
    DevBuf<uint32_t> arena, pathbuf;
    DevBuf<uint32_t> seed_off, seed_rank, seed_pos, seed_cost, seed_depth, seed_pre, lane_limit, row_map;
    size_t seeds_cap = 0, lanes_alloc = 0;
    DevBuf<uint32_t> tgt_pos;
    size_t tgt_cap = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bfg_timings last{};
    uint64_t graph_bytes = 0;
};

namespace {

uint32_t d1_of(const bfg_ctx *c) { return c->L.max_height + 1; }

int ensure_lane_buffers(bfg_ctx *c, size_t n_lanes, size_t n_seeds) {
    if (n_lanes > c->lanes_alloc) {
        HIP_TRY(c->seed_off.alloc(n_lanes + 1));
        HIP_TRY(c->lane_limit.alloc(n_lanes));
        HIP_TRY(c->row_map.alloc(n_lanes));
        c->lanes_alloc = n_lanes;
    }
    if (n_seeds > c->seeds_cap) {
        HIP_TRY(c->seed_rank.alloc(n_seeds));
        HIP_TRY(c->seed_pos.alloc(n_seeds));
        HIP_TRY(c->seed_cost.alloc(n_seeds));
        HIP_TRY(c->seed_depth.alloc(n_seeds));
        HIP_TRY(c->seed_pre.alloc(n_seeds));
        HIP_TRY(c->pathbuf.alloc(n_seeds * (size_t)d1_of(c)));
        c->seeds_cap = n_seeds;
    }
    return BFG_SUCCESS;
}

int ensure_arena(bfg_ctx *c) {
    const size_t need = (size_t)c->lane_cap * c->L.n;
    if (c->arena.n < need) {
        c->arena.release();
        HIP_TRY(c->arena.alloc(need));
    }
    return BFG_SUCCESS;
}

uint32_t auto_lane_cap(bfg_ctx *c, uint32_t want) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)8 << 30;
    const size_t have = free_b + c->arena.n * sizeof(uint32_t);
    const size_t per_lane = (size_t)c->L.n * 4 + (size_t)d1_of(c) * 4 + 64;
    const size_t budget = (size_t)(have * 0.6);
    size_t cap = std::max<size_t>(budget / std::max<size_t>(per_lane, 1), c->lanes);
    cap = std::min<size_t>(cap, 16384);
    cap = std::min<size_t>(cap, ((size_t)want + c->lanes - 1) / c->lanes * c->lanes);
    cap = std::max<size_t>(cap / c->lanes * c->lanes, c->lanes);
    return (uint32_t)cap;
}

template <int KL>
int launch_phases(bfg_ctx *c, uint32_t nl, float *up_ms, float *sweep_ms, uint32_t *launches) {
    const bfg::Layout &L = c->L;
    hipStream_t s = c->stream;
    const uint32_t groups = (nl + KL - 1) / KL;
    const uint32_t n = L.n, D1 = d1_of(c);
    const uint32_t nsl = (uint32_t)L.seglvl_seg.size() - 1;
    uint32_t nlaunch = 0;
    HIP_TRY(hipEventRecord(c->ev[0], s));
    if (L.n_top) {
        const uint64_t words = (uint64_t)L.n_top * KL;
        const uint64_t total = words * groups;
        const uint32_t blocks = (uint32_t)std::min<uint64_t>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(fill_top_kernel, dim3(blocks), dim3(256), 0, s, c->arena.p, (uint64_t)n * KL, words,
                           groups);
    }
    hipLaunchKernelGGL(bottom_walk_kernel<KL>, dim3(nl), dim3(64), 0, s, c->seed_off.p, c->seed_rank.p,
                       c->seed_pos.p, c->seed_cost.p, c->seed_depth.p, c->lane_limit.p, c->parent.p,
                       c->pos_parent.p, c->pos_depth.p, c->up_off.p, c->up_tgt.p, c->up_w.p, L.n_top, c->pathbuf.p,
                       D1, c->arena.p, n);
    for (int l = (int)nsl - 1; l >= 0; --l) {
        const uint32_t s0 = L.seglvl_seg[l], s1 = L.seglvl_seg[l + 1];
        hipLaunchKernelGGL(up_tile_kernel<KL>, dim3(groups, s1 - s0), dim3(TILE_THREADS), 0, s, s0, l > 0 ? 1 : 0,
                           c->seg_pos.p, c->tb_up.p, c->tb_w.p, c->pt_blk.p, c->pt_tgt.p, c->pr_off.p, c->pr_x.p,
                           c->pr_w.p, c->arena.p, n);
        ++nlaunch;
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->ev[1], s));
    for (uint32_t l = 0; l < nsl; ++l) {
        const uint32_t s0 = L.seglvl_seg[l], s1 = L.seglvl_seg[l + 1];
        hipLaunchKernelGGL(down_tile_kernel<KL>, dim3(groups, s1 - s0), dim3(TILE_THREADS), 0, s, s0, c->seg_pos.p,
                           c->xr_off.p, c->xr_src.p, c->xr_w.p, c->tb_dn.p, c->tb_w.p, c->arena.p, n);
        ++nlaunch;
    }
    const uint32_t nch = (uint32_t)L.ch_off.size() - 1;
    if (nch) {
        hipLaunchKernelGGL((bottom_chunk_kernel<KL, SEG, NTB>), dim3(nch, groups), dim3(NTB), 0, s, c->ch_off.p,
                           c->ch_pos.p, c->cs_off.p, c->cs_src.p, c->cs_aoff.p, c->ca_v.p, c->ca_w.p, c->pos_pre.p,
                           c->pos_sub.p, c->pos_depth.p, c->seed_off.p, c->seed_pre.p, c->seed_depth.p, c->pathbuf.p,
                           D1, nl, c->arena.p, n);
        ++nlaunch;
    }
    constexpr int NSP = NT / KL;
    for (size_t b = 0; b + 1 < L.botlvl_pos.size(); ++b) {
        const uint32_t p0 = L.botlvl_pos[b], p1 = L.botlvl_pos[b + 1];
        hipLaunchKernelGGL(bottom_pull_kernel<KL>, dim3((p1 - p0 + NSP - 1) / NSP, groups), dim3(NT), 0, s, p0, p1,
                           c->in_off.p, c->in_split.p, c->in_src.p, c->in_w.p, c->arena.p, n);
        ++nlaunch;
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->ev[2], s));
    HIP_TRY(hipEventSynchronize(c->ev[2]));
    float a = 0, b = 0;
    HIP_TRY(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
    HIP_TRY(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
    *up_ms += a;
    *sweep_ms += b;
    *launches += nlaunch;
    return BFG_SUCCESS;
}

// One batch: lanes [0, nl) in sorted order, staged host seed lists; runs the
// upward and downward phases; the output step is done by the caller.
int run_batch(bfg_ctx *c, uint32_t nl, const std::vector<uint32_t> &soff, const std::vector<uint32_t> &srank,
              const std::vector<uint32_t> &scost, const std::vector<uint32_t> &limits,
              const std::vector<uint32_t> &rows, float *up_ms, float *sweep_ms, uint32_t *launches) {
    const size_t ns = srank.size();
    int rc = ensure_lane_buffers(c, nl, ns);
    if (rc) return rc;
    std::vector<uint32_t> sdepth(ns), spre(ns), spos(ns);
    for (size_t j = 0; j < ns; ++j) {
        sdepth[j] = c->L.depth[srank[j]];
        spre[j] = c->L.pre[srank[j]];
        spos[j] = c->L.pos_of_rank[srank[j]];
    }
    hipStream_t s = c->stream;
    HIP_TRY(hipMemcpyAsync(c->seed_off.p, soff.data(), (nl + 1) * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->lane_limit.p, limits.data(), nl * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->row_map.p, rows.data(), nl * 4, hipMemcpyHostToDevice, s));
    if (ns) {
        HIP_TRY(hipMemcpyAsync(c->seed_rank.p, srank.data(), ns * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(c->seed_pos.p, spos.data(), ns * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(c->seed_cost.p, scost.data(), ns * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(c->seed_depth.p, sdepth.data(), ns * 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(c->seed_pre.p, spre.data(), ns * 4, hipMemcpyHostToDevice, s));
    }
    switch (c->lanes) {
        case 16: rc = launch_phases<16>(c, nl, up_ms, sweep_ms, launches); break;
        case 64: rc = launch_phases<64>(c, nl, up_ms, sweep_ms, launches); break;
        default: rc = launch_phases<32>(c, nl, up_ms, sweep_ms, launches); break;
    }
    return rc;  // launch_phases synchronised the stream: the staged vectors may go
}

int gather_out(bfg_ctx *c, uint32_t nl, const uint32_t *d_tpos, uint32_t T, const uint32_t *d_rows, uint32_t *d_out,
               uint64_t ld, unsigned long long *d_bits, uint64_t bits_ld, float *out_ms) {
    hipStream_t s = c->stream;
    const uint32_t groups = (nl + c->lanes - 1) / c->lanes;
    dim3 grid((T + 63) / 64, groups);
    HIP_TRY(hipEventRecord(c->ev[0], s));
    switch (c->lanes) {
        case 16:
            hipLaunchKernelGGL(gather_kernel<16>, grid, dim3(NT), 0, s, c->arena.p, c->L.n, d_tpos, T,
                               c->lane_limit.p, nl, d_rows, d_out, ld, d_bits, bits_ld);
            break;
        case 64:
            hipLaunchKernelGGL(gather_kernel<64>, grid, dim3(NT), 0, s, c->arena.p, c->L.n, d_tpos, T,
                               c->lane_limit.p, nl, d_rows, d_out, ld, d_bits, bits_ld);
            break;
        default:
            hipLaunchKernelGGL(gather_kernel<32>, grid, dim3(NT), 0, s, c->arena.p, c->L.n, d_tpos, T,
                               c->lane_limit.p, nl, d_rows, d_out, ld, d_bits, bits_ld);
            break;
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->ev[1], s));
    HIP_TRY(hipEventSynchronize(c->ev[1]));
    float a = 0;
    HIP_TRY(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
    *out_ms += a;
    return BFG_SUCCESS;
}

struct Req {
    // per lane seeds and limit (host), original lane order
    std::vector<uint32_t> soff, srank, scost, limit;
};

int check_ranks(const bfg_ctx *c, const uint32_t *r, size_t k, const char *what) {
    for (size_t i = 0; i < k; ++i)
        if (r[i] >= c->L.n) return fail(BFG_INVALID_PARAMETER, std::string(what) + " rank out of range");
    return BFG_SUCCESS;
}

// Runs every lane of `req`: lanes are ordered by the pre-order of their first
// seed and cut into batches of lane_cap; after each batch,
// out(batch_first, batch_lanes, rows, &ms) writes the outputs, where rows[i]
// is the original index of sorted lane batch_first + i.
template <class F>
int run_all(bfg_ctx *c, const Req &req, F &&out) {
    const uint32_t L = (uint32_t)req.limit.size();
    if (!c->cap_fixed) c->lane_cap = auto_lane_cap(c, L);
    int rc = ensure_arena(c);
    if (rc) return rc;
    std::vector<uint32_t> order(L);
    std::iota(order.begin(), order.end(), 0);
    std::vector<uint32_t> key(L, 0);
    for (uint32_t i = 0; i < L; ++i)
        if (req.soff[i + 1] > req.soff[i]) key[i] = c->L.pre[req.srank[req.soff[i]]];
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
    auto t_start = std::chrono::steady_clock::now();
    bfg_timings tm{};
    for (uint32_t b0 = 0; b0 < L; b0 += c->lane_cap) {
        const uint32_t nl = std::min(c->lane_cap, L - b0);
        std::vector<uint32_t> soff(nl + 1), srank, scost, lim(nl), rows(nl);
        for (uint32_t i = 0; i < nl; ++i) {
            const uint32_t o = order[b0 + i];
            rows[i] = o;
            soff[i] = (uint32_t)srank.size();
            for (uint32_t k = req.soff[o]; k < req.soff[o + 1]; ++k) {
                srank.push_back(req.srank[k]);
                scost.push_back(req.scost[k]);
            }
            lim[i] = req.limit[o];
        }
        soff[nl] = (uint32_t)srank.size();
        rc = run_batch(c, nl, soff, srank, scost, lim, rows, &tm.up_ms, &tm.sweep_ms, &tm.sweep_launches);
        if (rc) return rc;
        rc = out(b0, nl, rows, &tm.output_ms);
        if (rc) return rc;
        tm.batches++;
    }
    tm.lanes_run = L;
    tm.total_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    c->last = tm;
    return BFG_SUCCESS;
}

Req single_source_req(const uint32_t *src, uint32_t k, uint32_t threshold) {
    Req r;
    r.soff.resize(k + 1);
    r.srank.assign(src, src + k);
    r.scost.assign(k, 0);
    r.limit.assign(k, threshold);
    for (uint32_t i = 0; i <= k; ++i) r.soff[i] = i;
    return r;
}

int stage_targets(bfg_ctx *c, const uint32_t *targets, uint32_t T) {
    std::vector<uint32_t> tp(T);
    for (uint32_t t = 0; t < T; ++t) tp[t] = c->L.pos_of_rank[targets[t]];
    if (T > c->tgt_cap) {
        HIP_TRY(c->tgt_pos.alloc(T));
        c->tgt_cap = T;
    }
    HIP_TRY(hipMemcpyAsync(c->tgt_pos.p, tp.data(), (size_t)T * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));  // tp dies with this frame
    return BFG_SUCCESS;
}

// Full rows for all ranks (dist and/or reach bits) into host buffers indexed
// by original lane.
int dense_rows(bfg_ctx *c, const Req &req, uint32_t *dist_host, uint64_t *bits_host) {
    const uint32_t n = c->L.n;
    const uint64_t words = (n + 63) / 64;
    DevBuf<uint32_t> d_rows, d_all_pos;
    DevBuf<unsigned long long> d_bits;
    std::vector<uint32_t> all_pos(c->L.pos_of_rank);
    HIP_TRY(d_all_pos.upload(all_pos, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<uint32_t> hrow;
    std::vector<uint64_t> hbits;
    uint32_t cap = 0;
    return run_all(c, req, [&](uint32_t b0, uint32_t nl, const std::vector<uint32_t> &rows, float *oms) -> int {
        (void)b0;
        if (cap < nl) {
            if (dist_host) HIP_TRY(d_rows.alloc((size_t)nl * n));
            if (bits_host) HIP_TRY(d_bits.alloc((size_t)nl * words));
            cap = nl;
        }
        int rc = gather_out(c, nl, d_all_pos.p, n, nullptr, dist_host ? d_rows.p : nullptr, n,
                            bits_host ? d_bits.p : nullptr, words, oms);
        if (rc) return rc;
        if (dist_host) {
            hrow.resize((size_t)nl * n);
            HIP_TRY(hipMemcpy(hrow.data(), d_rows.p, (size_t)nl * n * 4, hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < nl; ++i)
                memcpy(dist_host + (size_t)rows[i] * n, hrow.data() + (size_t)i * n, (size_t)n * 4);
        }
        if (bits_host) {
            hbits.resize((size_t)nl * words);
            HIP_TRY(hipMemcpy(hbits.data(), d_bits.p, (size_t)nl * words * 8, hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < nl; ++i)
                memcpy(bits_host + (size_t)rows[i] * words, hbits.data() + (size_t)i * words, (size_t)words * 8);
        }
        return BFG_SUCCESS;
    });
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

}  // namespace

extern "C" {

const char *bfg_last_error(void) { return g_last_error.c_str(); }

int bfg_ctx_create(const bfg_graph_desc *gd, int device, bfg_ctx **out) {
    if (!gd || !out) return fail(BFG_INVALID_PARAMETER, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(BFG_DEVICE_ERROR, "no HIP device");
    if (device < 0 || device >= ndev) return fail(BFG_INVALID_PARAMETER, "device index out of range");
    std::unique_ptr<bfg_ctx> c(new bfg_ctx());
    c->device = device;
    HIP_TRY(hipSetDevice(device));
    const bfg::CchInput in = input_of(gd);
    std::string err;
    bfg::LayoutParams lp;
    lp.segment_limit = SEG;
    if (!bfg::build_layout(in, lp, c->L, err)) return fail(BFG_INVALID_PARAMETER, err);
    // every slot, position and segment index the kernels derive rests on these
    // properties; refuse the graph rather than launch on a broken layout
    if (!bfg::check_layout(in, c->L, err)) return fail(BFG_INTERNAL_ERROR, err);
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto &e : c->ev) HIP_TRY(hipEventCreate(&e));
    hipStream_t s = c->stream;
    auto &L = c->L;
    HIP_TRY(c->parent.upload(L.parent, s));
    HIP_TRY(c->pos_parent.upload(L.pos_parent, s));
    HIP_TRY(c->pos_pre.upload(L.pos_pre, s));
    HIP_TRY(c->pos_sub.upload(L.pos_subtree, s));
    HIP_TRY(c->pos_depth.upload(L.pos_depth, s));
    HIP_TRY(c->seg_pos.upload(L.seg_pos, s));
    HIP_TRY(c->in_off.upload(L.in_off, s));
    HIP_TRY(c->in_split.upload(L.in_split, s));
    HIP_TRY(c->in_src.upload(L.in_src, s));
    HIP_TRY(c->in_w.upload(L.in_w, s));
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
    HIP_TRY(c->xr_src.upload(L.xr_src, s));
    HIP_TRY(c->xr_w.upload(L.xr_w, s));
    HIP_TRY(c->tb_dn.upload(L.tb_dn, s));
    HIP_TRY(c->tb_up.upload(L.tb_up, s));
    HIP_TRY(c->tb_w.upload(L.tb_w, s));
    HIP_TRY(c->pt_blk.upload(L.pt_blk, s));
    HIP_TRY(c->pt_tgt.upload(L.pt_tgt, s));
    HIP_TRY(c->pr_off.upload(L.pr_off, s));
    HIP_TRY(c->pr_x.upload(L.pr_x, s));
    HIP_TRY(c->pr_w.upload(L.pr_w, s));
    HIP_TRY(hipStreamSynchronize(s));
    c->graph_bytes = 0;
    for (auto *b : {&c->parent, &c->pos_parent, &c->pos_pre, &c->pos_sub, &c->pos_depth, &c->seg_pos, &c->in_off,
                    &c->in_split, &c->in_src, &c->in_w, &c->ch_off, &c->ch_pos, &c->cs_off, &c->cs_src, &c->cs_aoff,
                    &c->ca_w, &c->up_off, &c->up_tgt, &c->up_w, &c->xr_off, &c->xr_src, &c->xr_w, &c->tb_dn,
                    &c->tb_up, &c->tb_w, &c->pt_blk, &c->pt_tgt, &c->pr_off, &c->pr_x, &c->pr_w})
        c->graph_bytes += b->n * sizeof(uint32_t);
    c->graph_bytes += c->ca_v.n * sizeof(uint16_t);
    // host copies the kernels alone need
    for (auto *v : {&L.in_src, &L.in_w, &L.up_tgt, &L.up_w, &L.bs_src, &L.bs_aoff, &L.ba_w, &L.ut_tgt, &L.ut_aoff,
                    &L.uta_w, &L.sd_w, &L.su_w, &L.cs_src, &L.cs_aoff, &L.ca_w, &L.ch_pos, &L.xr_src, &L.xr_w,
                    &L.tb_dn, &L.tb_up, &L.tb_w, &L.pt_tgt, &L.pr_off, &L.pr_x, &L.pr_w})
        std::vector<uint32_t>().swap(*v);
    for (auto *v : {&L.ba_v, &L.uta_x, &L.ca_v, &L.sd_tgt, &L.su_tgt}) std::vector<uint16_t>().swap(*v);
    *out = c.release();
    return BFG_SUCCESS;
}

void bfg_ctx_destroy(bfg_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int bfg_validate_graph(const bfg_graph_desc *gd, bfg_ctx_info *info) {
    if (!gd || !info) return fail(BFG_INVALID_PARAMETER, "null argument");
    const bfg::CchInput in = input_of(gd);
    bfg::Layout L;
    std::string err;
    bfg::LayoutParams lp;
    lp.segment_limit = SEG;
    if (!bfg::build_layout(in, lp, L, err)) return fail(BFG_INVALID_PARAMETER, err);
    if (!bfg::check_layout(in, L, err)) return fail(BFG_INTERNAL_ERROR, err);
    fill_info(L, info);
    return BFG_SUCCESS;
}

int bfg_ctx_get_info(const bfg_ctx *c, bfg_ctx_info *info) {
    if (!c || !info) return fail(BFG_INVALID_PARAMETER, "null argument");
    fill_info(c->L, info);
    info->lanes = c->lanes;
    info->lane_capacity = c->lane_cap;
    info->device_bytes = c->graph_bytes + c->arena.n * 4 + c->pathbuf.n * 4;
    return BFG_SUCCESS;
}

int bfg_ctx_last_timings(const bfg_ctx *c, bfg_timings *t) {
    if (!c || !t) return fail(BFG_INVALID_PARAMETER, "null argument");
    *t = c->last;
    return BFG_SUCCESS;
}

int bfg_ctx_configure(bfg_ctx *c, uint32_t lanes, uint32_t lane_capacity) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    if (lanes != 0) {
        if (lanes != 16 && lanes != 32 && lanes != 64) return fail(BFG_INVALID_PARAMETER, "lanes must be 16, 32 or 64");
        c->lanes = lanes;
    }
    if (lane_capacity != 0) {
        c->lane_cap = (lane_capacity + c->lanes - 1) / c->lanes * c->lanes;
        c->cap_fixed = true;
        return ensure_arena(c);
    }
    c->cap_fixed = false;
    if (c->lane_cap) c->lane_cap = (c->lane_cap + c->lanes - 1) / c->lanes * c->lanes;
    return BFG_SUCCESS;
}

static int matrix_impl(bfg_ctx *c, const uint32_t *src, uint32_t S, const uint32_t *tgt, uint32_t T, uint32_t thr,
                       uint32_t *d_out, uint32_t *host_out) {
    Req req = single_source_req(src, S, thr);
    int rc = stage_targets(c, tgt, T);
    if (rc) return rc;
    DevBuf<uint32_t> tmp;
    if (!d_out) {
        HIP_TRY(tmp.alloc((size_t)S * T));
        d_out = tmp.p;
    }
    rc = run_all(c, req, [&](uint32_t, uint32_t nl, const std::vector<uint32_t> &, float *oms) -> int {
        return gather_out(c, nl, c->tgt_pos.p, T, c->row_map.p, d_out, T, nullptr, 0, oms);
    });
    if (rc) return rc;
    if (host_out) HIP_TRY(hipMemcpy(host_out, d_out, (size_t)S * T * 4, hipMemcpyDeviceToHost));
    return BFG_SUCCESS;
}

int bfg_matrix_bounded(bfg_ctx *c, const uint32_t *src, uint32_t S, const uint32_t *tgt, uint32_t T, uint32_t thr,
                       uint32_t *out) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if ((S && !src) || (T && !tgt) || (S && T && !out)) return fail(BFG_INVALID_PARAMETER, "null buffer");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    int rc = check_ranks(c, src, S, "source");
    if (!rc) rc = check_ranks(c, tgt, T, "target");
    if (rc) return rc;
    if (S == 0 || T == 0) return BFG_SUCCESS;
    return matrix_impl(c, src, S, tgt, T, thr, nullptr, out);
}

int bfg_matrix(bfg_ctx *c, const uint32_t *src, uint32_t S, const uint32_t *tgt, uint32_t T, uint32_t *out) {
    return bfg_matrix_bounded(c, src, S, tgt, T, INF, out);
}

int bfg_matrix_device(bfg_ctx *c, const uint32_t *d_src, uint32_t S, const uint32_t *d_tgt, uint32_t T, uint32_t thr,
                      uint32_t *d_out) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if ((S && !d_src) || (T && !d_tgt) || (S && T && !d_out)) return fail(BFG_INVALID_PARAMETER, "null buffer");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    if (S == 0 || T == 0) return BFG_SUCCESS;
    std::vector<uint32_t> src(S), tgt(T);
    HIP_TRY(hipMemcpy(src.data(), d_src, (size_t)S * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(tgt.data(), d_tgt, (size_t)T * 4, hipMemcpyDeviceToHost));
    int rc = check_ranks(c, src.data(), S, "source");
    if (!rc) rc = check_ranks(c, tgt.data(), T, "target");
    if (rc) return rc;
    return matrix_impl(c, src.data(), S, tgt.data(), T, thr, d_out, nullptr);
}

int bfg_query_batch(bfg_ctx *c, const uint32_t *src, uint32_t S, uint32_t thr, uint32_t *dist) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if (S && (!src || !dist)) return fail(BFG_INVALID_PARAMETER, "null buffer");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    int rc = check_ranks(c, src, S, "source");
    if (rc) return rc;
    if (S == 0) return BFG_SUCCESS;
    return dense_rows(c, single_source_req(src, S, thr), dist, nullptr);
}

int bfg_isochrone(bfg_ctx *c, const uint32_t *origins, const uint32_t *limits, uint32_t L, uint64_t *bits,
                  uint32_t *times) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if (L && (!origins || !limits || !bits)) return fail(BFG_INVALID_PARAMETER, "null buffer");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    int rc = check_ranks(c, origins, L, "origin");
    if (rc) return rc;
    if (L == 0) return BFG_SUCCESS;
    Req req = single_source_req(origins, L, INF);
    req.limit.assign(limits, limits + L);
    return dense_rows(c, req, times, bits);
}

int bfg_phast_seeded(bfg_ctx *c, const uint32_t *seed_offsets, const uint32_t *seed_ranks, const uint32_t *seed_costs,
                     uint32_t L, const uint32_t *thresholds, uint32_t *dist) {
    if (!c) return fail(BFG_INVALID_PARAMETER, "null context");
    if (L && (!seed_offsets || !thresholds || !dist)) return fail(BFG_INVALID_PARAMETER, "null buffer");
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    if (L == 0) return BFG_SUCCESS;
    if (seed_offsets[0] != 0) return fail(BFG_INVALID_PARAMETER, "seed_offsets[0] must be 0");
    Req req;
    req.soff.assign(seed_offsets, seed_offsets + L + 1);
    for (uint32_t i = 0; i < L; ++i) {
        if (req.soff[i + 1] < req.soff[i]) return fail(BFG_INVALID_PARAMETER, "seed_offsets not monotone");
        if (req.soff[i + 1] - req.soff[i] > 64) return fail(BFG_NOT_SUPPORTED, "more than 64 seeds in one lane");
    }
    const uint32_t ns = req.soff[L];
    if (ns && (!seed_ranks || !seed_costs)) return fail(BFG_INVALID_PARAMETER, "null seed buffer");
    int rc = check_ranks(c, seed_ranks, ns, "seed");
    if (rc) return rc;
    req.srank.assign(seed_ranks, seed_ranks + ns);
    req.scost.assign(seed_costs, seed_costs + ns);
    req.limit.assign(thresholds, thresholds + L);
    return dense_rows(c, req, dist, nullptr);
}

}  // extern "C"
