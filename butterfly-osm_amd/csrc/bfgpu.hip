// bfgpu.hip — MI355X (gfx950) K-lane batched PHAST: kernels + C ABI.
//
// Phases of one batch of source lanes (sources are CCH ranks):
//   K1 up_walk   — per lane, the upward search of the reference
//                  (batched_phast.rs:156-189, isochrone_handler.rs:388-411)
//                  as an elimination-tree ancestor walk: the union of the
//                  seeds' root paths is visited in ascending rank (a valid
//                  order for UP arcs), relaxing UP arcs into per-seed path
//                  slots indexed by depth.  One wavefront per lane.
//   K2 sweep     — the rank-descending DOWN scan of the reference
//                  (batched_phast.rs:205-250, :538-590) as a pull sweep over
//                  DOWN arcs keyed by head, scheduled by the etree: kernel
//                  launches walk task levels, a workgroup owns one task for
//                  KL interleaved lanes, steps inside a task are separated by
//                  workgroup barriers.  Labels live in an arena
//                  [group][position][KL] so each arc read serves KL lanes with
//                  one coalesced KL*4-byte access.
//   K4 gather    — target extraction (batched_phast.rs:693-701) through an
//                  LDS transpose into row-major S x T, with the per-lane limit
//                  mask; also serves full distance rows and isochrone bits.
// Arithmetic is u32 with saturating adds (v_add_u32 clamp) and min; the
// result is the unique shortest up-down distance, so it is bit-identical to
// the reference whatever the traversal order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
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

__device__ __forceinline__ uint32_t sat_add(uint32_t a, uint32_t b) {
    return __builtin_elementwise_add_sat(a, b);
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}

// ---------------------------------------------------------------------------
// K1: upward ancestor walk, one wavefront (64 threads) per lane.
// Lane l owns seeds j in [seed_off[l], seed_off[l+1]) (<= 64), each with a path
// slot pathbuf[j*D1 + i] holding the label of the node i levels above seed j.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void up_walk_kernel(
    const uint32_t *__restrict__ seed_off, const uint32_t *__restrict__ seed_rank,
    const uint32_t *__restrict__ seed_cost, const uint32_t *__restrict__ seed_depth,
    const uint32_t *__restrict__ lane_limit, const uint32_t *__restrict__ up_off,
    const uint32_t *__restrict__ up_tgt, const uint32_t *__restrict__ up_w,
    const uint32_t *__restrict__ parent, const uint32_t *__restrict__ depth,
    uint32_t *__restrict__ pathbuf, uint32_t D1) {
    const uint32_t l = blockIdx.x;
    const uint32_t t = threadIdx.x;
    const uint32_t j0 = seed_off[l];
    const uint32_t ns = seed_off[l + 1] - j0;
    const uint32_t lim = lane_limit[l];
    for (uint32_t j = 0; j < ns; ++j) {
        uint32_t *pb = pathbuf + (size_t)(j0 + j) * D1;
        const uint32_t dd = seed_depth[j0 + j];
        for (uint32_t i = t; i <= dd; i += 64) pb[i] = INF;
    }
    __syncthreads();
    uint32_t cur = NONE;
    if (t < ns) {
        cur = seed_rank[j0 + t];
        pathbuf[(size_t)(j0 + t) * D1] = seed_cost[j0 + t];
    }
    __syncthreads();
    for (;;) {
        const uint32_t x = wave_min(cur);
        if (x == NONE) break;
        const bool at = (cur == x);
        const uint32_t dxn = depth[x];
        uint32_t mine = INF;
        size_t my_slot = 0;
        if (at) {
            my_slot = (size_t)(j0 + t) * D1 + (seed_depth[j0 + t] - dxn);
            mine = pathbuf[my_slot];
        }
        const uint32_t dx = wave_min(mine);
        const uint64_t mask = __ballot(at);
        const uint32_t first = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
        __syncthreads();
        if (at) pathbuf[my_slot] = dx;
        if (dx != INF && dx <= lim) {
            const uint32_t js = j0 + first;
            const uint32_t dsj = seed_depth[js];
            uint32_t *pb = pathbuf + (size_t)js * D1;
            const uint32_t a1 = up_off[x + 1];
            for (uint32_t a = up_off[x] + t; a < a1; a += 64) {
                const uint32_t y = up_tgt[a];
                const uint32_t nd = sat_add(dx, up_w[a]);
                uint32_t *slot = pb + (dsj - depth[y]);
                if (nd < *slot) *slot = nd;
            }
        }
        __syncthreads();
        if (at) cur = parent[x];
    }
}

// ---------------------------------------------------------------------------
// K2: pull sweep for the tasks of one level.  grid = (tasks, lane groups).
// ---------------------------------------------------------------------------
template <int KL>
__global__ __launch_bounds__(256) void sweep_kernel(
    uint32_t task_base, const uint32_t *__restrict__ task_step, const uint32_t *__restrict__ step_pos,
    const uint32_t *__restrict__ in_off, const uint32_t *__restrict__ in_src,
    const uint32_t *__restrict__ in_w, const uint32_t *__restrict__ pos_pre,
    const uint32_t *__restrict__ pos_sub, const uint32_t *__restrict__ pos_depth,
    const uint32_t *__restrict__ seed_off, const uint32_t *__restrict__ seed_pre,
    const uint32_t *__restrict__ seed_depth, const uint32_t *__restrict__ pathbuf, uint32_t D1,
    uint32_t n_lanes, uint32_t *__restrict__ arena, uint32_t n) {
    constexpr int NT = 256, NS = NT / KL, C = 8192 / KL;
    __shared__ uint32_t acc[C * KL];
    const uint32_t task = task_base + blockIdx.x;
    const uint32_t g = blockIdx.y;
    const uint32_t lane = threadIdx.x % KL, slot = threadIdx.x / KL;
    const uint32_t gl = g * KL + lane;
    uint32_t *lab = arena + (size_t)g * n * KL;
    uint32_t sj0 = 0, sj1 = 0;
    if (gl < n_lanes) { sj0 = seed_off[gl]; sj1 = seed_off[gl + 1]; }
    const uint32_t s_end = task_step[task + 1];
    for (uint32_t s = task_step[task]; s < s_end; ++s) {
        const uint32_t p0 = step_pos[s], p1 = step_pos[s + 1];
        for (uint32_t c0 = p0; c0 < p1; c0 += C) {
            const uint32_t c1 = min(p1, c0 + (uint32_t)C);
            for (uint32_t i = slot; i < c1 - c0; i += NS) {
                const uint32_t p = c0 + i;
                uint32_t v = INF;
                if (sj0 < sj1) {
                    const uint32_t pr = pos_pre[p], ps = pos_sub[p], pd = pos_depth[p];
                    for (uint32_t j = sj0; j < sj1; ++j) {
                        const uint32_t sp = seed_pre[j];
                        if (sp >= pr && sp - pr < ps) v = min(v, pathbuf[(size_t)j * D1 + (seed_depth[j] - pd)]);
                    }
                }
                acc[i * KL + lane] = v;
            }
            __syncthreads();
            const uint32_t A0 = in_off[c0], A1 = in_off[c1];
            const uint32_t per = (A1 - A0 + NS - 1) / NS;
            uint32_t a = A0 + slot * per;
            const uint32_t ae = min(A1, a + per);
            if (a < ae) {
                uint32_t lo = c0, hi = c1;  // in_off[lo] <= a < in_off[hi]
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (in_off[mid] <= a) lo = mid; else hi = mid;
                }
                uint32_t node = lo, nend = in_off[node + 1];
                uint32_t m = INF;
                for (; a < ae; ++a) {
                    while (a >= nend) {
                        if (m != INF) atomicMin(&acc[(node - c0) * KL + lane], m);
                        m = INF;
                        ++node;
                        nend = in_off[node + 1];
                    }
                    const uint32_t u = in_src[a];
                    const uint32_t du = lab[(size_t)u * KL + lane];
                    m = min(m, sat_add(du, in_w[a]));
                }
                if (m != INF) atomicMin(&acc[(node - c0) * KL + lane], m);
            }
            __syncthreads();
            for (uint32_t i = slot; i < c1 - c0; i += NS) lab[(size_t)(c0 + i) * KL + lane] = acc[i * KL + lane];
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// K4: gather target columns of one batch into row-major rows, with the lane
// limit mask.  Tile = 64 targets x KL lanes, transposed through LDS.
// out[(row0 + lane) * ld + t] for lanes < n_lanes.  Optional reach bits:
// bits[(row0+lane) * bits_ld + t/64] (t0 multiple of 64).
// ---------------------------------------------------------------------------
template <int KL>
__global__ __launch_bounds__(256) void gather_kernel(
    const uint32_t *__restrict__ arena, uint32_t n, const uint32_t *__restrict__ tgt_pos, uint32_t T,
    const uint32_t *__restrict__ lane_limit, uint32_t n_lanes, uint32_t *__restrict__ out, uint64_t ld,
    uint64_t row0, unsigned long long *__restrict__ bits, uint64_t bits_ld) {
    __shared__ uint32_t tile[64][KL + 1];
    const uint32_t t0 = blockIdx.x * 64;
    const uint32_t g = blockIdx.y;
    const uint32_t *lab = arena + (size_t)g * n * KL;
    for (uint32_t idx = threadIdx.x; idx < 64 * KL; idx += 256) {
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
        for (uint32_t idx = threadIdx.x; idx < 64 * KL; idx += 256) {
            const uint32_t lane = idx / 64, ti = idx % 64;
            const uint32_t gl = g * KL + lane, t = t0 + ti;
            if (gl < n_lanes && t < T) out[(row0 + gl) * ld + t] = tile[ti][lane];
        }
    }
    if (bits) {
        const uint32_t w = threadIdx.x / 64, ti = threadIdx.x % 64;
        for (uint32_t lane = w; lane < KL; lane += 4) {
            const uint32_t gl = g * KL + lane;
            const bool r = (t0 + ti < T) && tile[ti][lane] != INF;
            const unsigned long long word = __ballot(r);
            if (ti == 0 && gl < n_lanes) bits[(row0 + gl) * bits_ld + (t0 / 64)] = word;
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

}  // namespace

struct bfg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    bfg::Layout L;
    uint32_t lanes = 32;        // KL
    uint32_t lane_cap = 0;      // lanes per batch (multiple of KL)
    // resident graph
    DevBuf<uint32_t> up_off, up_tgt, up_w, parent, depth;
    DevBuf<uint32_t> in_off, in_src, in_w, pos_pre, pos_sub, pos_depth;
    DevBuf<uint32_t> step_pos, task_step, pos_of_rank;
    // arenas (per batch)
    DevBuf<uint32_t> arena, pathbuf;
    size_t pathbuf_cap = 0;
    DevBuf<uint32_t> seed_off, seed_rank, seed_cost, seed_depth, seed_pre, lane_limit;
    size_t seeds_cap = 0, lanes_alloc = 0;
    DevBuf<uint32_t> tgt_pos;
    size_t tgt_cap = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bfg_timings last{};
    uint64_t graph_bytes = 0;
};

namespace {

int ensure_lane_buffers(bfg_ctx *c, size_t n_lanes, size_t n_seeds) {
    if (n_lanes > c->lanes_alloc) {
        HIP_TRY(c->seed_off.alloc(n_lanes + 1));
        HIP_TRY(c->lane_limit.alloc(n_lanes));
        c->lanes_alloc = n_lanes;
    }
    if (n_seeds > c->seeds_cap) {
        HIP_TRY(c->seed_rank.alloc(n_seeds));
        HIP_TRY(c->seed_cost.alloc(n_seeds));
        HIP_TRY(c->seed_depth.alloc(n_seeds));
        HIP_TRY(c->seed_pre.alloc(n_seeds));
        c->seeds_cap = n_seeds;
    }
    const size_t need_pb = n_seeds * (size_t)(c->L.max_depth + 1);
    if (need_pb > c->pathbuf_cap) {
        HIP_TRY(c->pathbuf.alloc(need_pb));
        c->pathbuf_cap = need_pb;
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
    size_t have = free_b + c->arena.n * sizeof(uint32_t);
    size_t per_lane = (size_t)c->L.n * 4 + (size_t)(c->L.max_depth + 1) * 4;
    size_t budget = (size_t)(have * 0.6);
    size_t cap = std::max<size_t>(budget / std::max<size_t>(per_lane, 1), c->lanes);
    cap = std::min<size_t>(cap, 8192);
    cap = std::min<size_t>(cap, ((size_t)want + c->lanes - 1) / c->lanes * c->lanes);
    cap = std::max<size_t>(cap / c->lanes * c->lanes, c->lanes);
    return (uint32_t)cap;
}

typedef void (*sweep_fn_t)(dim3, dim3, hipStream_t, bfg_ctx *, uint32_t, uint32_t);

template <int KL>
void launch_sweep(dim3 grid, dim3 block, hipStream_t s, bfg_ctx *c, uint32_t task_base, uint32_t n_lanes) {
    hipLaunchKernelGGL(sweep_kernel<KL>, grid, block, 0, s, task_base, c->task_step.p, c->step_pos.p,
                       c->in_off.p, c->in_src.p, c->in_w.p, c->pos_pre.p, c->pos_sub.p, c->pos_depth.p,
                       c->seed_off.p, c->seed_pre.p, c->seed_depth.p, c->pathbuf.p, c->L.max_depth + 1,
                       n_lanes, c->arena.p, c->L.n);
}

template <int KL>
void launch_gather(dim3 grid, hipStream_t s, bfg_ctx *c, const uint32_t *tpos, uint32_t T, uint32_t n_lanes,
                   uint32_t *out, uint64_t ld, uint64_t row0, unsigned long long *bits, uint64_t bits_ld) {
    hipLaunchKernelGGL(gather_kernel<KL>, grid, dim3(256), 0, s, c->arena.p, c->L.n, tpos, T, c->lane_limit.p,
                       n_lanes, out, ld, row0, bits, bits_ld);
}

// One batch: lanes [0, nl) with host seed lists and limits already staged in
// host vectors; runs K1 + K2; output is done by the caller.
int run_batch(bfg_ctx *c, uint32_t nl, const std::vector<uint32_t> &soff, const std::vector<uint32_t> &srank,
              const std::vector<uint32_t> &scost, const std::vector<uint32_t> &limits, float *up_ms,
              float *sweep_ms, uint32_t *launches) {
    const size_t ns = srank.size();
    int rc = ensure_lane_buffers(c, nl, ns);
    if (rc) return rc;
    std::vector<uint32_t> sdepth(ns), spre(ns);
    for (size_t j = 0; j < ns; ++j) { sdepth[j] = c->L.depth[srank[j]]; spre[j] = c->L.pre[srank[j]]; }
    hipStream_t s = c->stream;
    HIP_TRY(hipMemcpyAsync(c->seed_off.p, soff.data(), (nl + 1) * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->lane_limit.p, limits.data(), nl * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->seed_rank.p, srank.data(), ns * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->seed_cost.p, scost.data(), ns * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->seed_depth.p, sdepth.data(), ns * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->seed_pre.p, spre.data(), ns * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipEventRecord(c->ev[0], s));
    hipLaunchKernelGGL(up_walk_kernel, dim3(nl), dim3(64), 0, s, c->seed_off.p, c->seed_rank.p, c->seed_cost.p,
                       c->seed_depth.p, c->lane_limit.p, c->up_off.p, c->up_tgt.p, c->up_w.p, c->parent.p,
                       c->depth.p, c->pathbuf.p, c->L.max_depth + 1);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c->ev[1], s));
    const uint32_t groups = (nl + c->lanes - 1) / c->lanes;
    const size_t nlev = c->L.level_task.size() - 1;
    uint32_t nlaunch = 0;
    for (size_t l = 0; l < nlev; ++l) {
        const uint32_t t0 = c->L.level_task[l], t1 = c->L.level_task[l + 1];
        if (t1 == t0) continue;
        dim3 grid(t1 - t0, groups);
        switch (c->lanes) {
            case 16: launch_sweep<16>(grid, dim3(256), s, c, t0, nl); break;
            case 64: launch_sweep<64>(grid, dim3(256), s, c, t0, nl); break;
            default: launch_sweep<32>(grid, dim3(256), s, c, t0, nl); break;
        }
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

int gather_out(bfg_ctx *c, uint32_t nl, const uint32_t *d_tpos, uint32_t T, uint32_t *d_out, uint64_t ld, uint64_t row0,
               unsigned long long *d_bits, uint64_t bits_ld, float *out_ms) {
    hipStream_t s = c->stream;
    const uint32_t groups = (nl + c->lanes - 1) / c->lanes;
    dim3 grid((T + 63) / 64, groups);
    HIP_TRY(hipEventRecord(c->ev[0], s));
    switch (c->lanes) {
        case 16: launch_gather<16>(grid, s, c, d_tpos, T, nl, d_out, ld, row0, d_bits, bits_ld); break;
        case 64: launch_gather<64>(grid, s, c, d_tpos, T, nl, d_out, ld, row0, d_bits, bits_ld); break;
        default: launch_gather<32>(grid, s, c, d_tpos, T, nl, d_out, ld, row0, d_bits, bits_ld); break;
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
    // per lane seeds and limit (host)
    std::vector<uint32_t> soff, srank, scost, limit;
};

int check_ranks(const bfg_ctx *c, const uint32_t *r, size_t k, const char *what) {
    for (size_t i = 0; i < k; ++i)
        if (r[i] >= c->L.n) return fail(BFG_INVALID_PARAMETER, std::string(what) + " rank out of range");
    return BFG_SUCCESS;
}

// Runs every lane of `req` in batches of lane_cap; after each batch, the
// callback `out(batch_first_lane, batch_lanes)` performs the output step.
template <class F>
int run_all(bfg_ctx *c, const Req &req, F &&out) {
    const uint32_t L = (uint32_t)req.limit.size();
    if (c->lane_cap == 0) c->lane_cap = auto_lane_cap(c, L);
    int rc = ensure_arena(c);
    if (rc) return rc;
    auto t_start = std::chrono::steady_clock::now();
    bfg_timings tm{};
    for (uint32_t b0 = 0; b0 < L; b0 += c->lane_cap) {
        const uint32_t nl = std::min(c->lane_cap, L - b0);
        std::vector<uint32_t> soff(nl + 1), srank, scost, lim(nl);
        for (uint32_t i = 0; i < nl; ++i) {
            soff[i] = (uint32_t)srank.size();
            for (uint32_t k = req.soff[b0 + i]; k < req.soff[b0 + i + 1]; ++k) {
                srank.push_back(req.srank[k]);
                scost.push_back(req.scost[k]);
            }
            lim[i] = req.limit[b0 + i];
        }
        soff[nl] = (uint32_t)srank.size();
        rc = run_batch(c, nl, soff, srank, scost, lim, &tm.up_ms, &tm.sweep_ms, &tm.sweep_launches);
        if (rc) return rc;
        rc = out(b0, nl, &tm.output_ms);
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

// Dense rows for all ranks of a batch (dist and/or bits) to host buffers.
int dense_rows(bfg_ctx *c, const Req &req, uint32_t *dist_host, uint64_t *bits_host) {
    const uint32_t n = c->L.n;
    const uint64_t words = (n + 63) / 64;
    DevBuf<uint32_t> d_rows;
    DevBuf<unsigned long long> d_bits;
    uint32_t cap = 0;
    return run_all(c, req, [&](uint32_t b0, uint32_t nl, float *oms) -> int {
        if (cap < nl) {
            if (dist_host) HIP_TRY(d_rows.alloc((size_t)nl * n));
            if (bits_host) HIP_TRY(d_bits.alloc((size_t)nl * words));
            cap = nl;
        }
        int rc = gather_out(c, nl, c->pos_of_rank.p, n, dist_host ? d_rows.p : nullptr, n, 0,
                            bits_host ? d_bits.p : nullptr, words, oms);
        if (rc) return rc;
        if (dist_host)
            HIP_TRY(hipMemcpy(dist_host + (size_t)b0 * n, d_rows.p, (size_t)nl * n * 4, hipMemcpyDeviceToHost));
        if (bits_host)
            HIP_TRY(hipMemcpy(bits_host + (size_t)b0 * words, d_bits.p, (size_t)nl * words * 8,
                              hipMemcpyDeviceToHost));
        return BFG_SUCCESS;
    });
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
    bfg::CchInput in;
    in.n = gd->n_nodes;
    in.up_off = gd->up_offsets; in.up_tgt = gd->up_targets; in.up_w = gd->up_weights;
    in.dn_off = gd->down_offsets; in.dn_tgt = gd->down_targets; in.dn_w = gd->down_weights;
    std::string err;
    bfg::LayoutParams lp;
    if (!bfg::build_layout(in, lp, c->L, err)) return fail(BFG_INVALID_PARAMETER, err);
    // every path-slot and position index the kernels derive rests on these
    // properties; refuse the graph rather than launch on a broken layout
    if (!bfg::check_ancestry(in, c->L, err)) return fail(BFG_INTERNAL_ERROR, err);
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto &e : c->ev) HIP_TRY(hipEventCreate(&e));
    hipStream_t s = c->stream;
    auto &L = c->L;
    HIP_TRY(c->up_off.upload(L.up_off, s));
    HIP_TRY(c->up_tgt.upload(L.up_tgt, s));
    HIP_TRY(c->up_w.upload(L.up_w, s));
    HIP_TRY(c->parent.upload(L.parent, s));
    HIP_TRY(c->depth.upload(L.depth, s));
    HIP_TRY(c->in_off.upload(L.in_off, s));
    HIP_TRY(c->in_src.upload(L.in_src, s));
    HIP_TRY(c->in_w.upload(L.in_w, s));
    HIP_TRY(c->pos_pre.upload(L.pos_pre, s));
    HIP_TRY(c->pos_sub.upload(L.pos_subtree, s));
    HIP_TRY(c->pos_depth.upload(L.pos_depth, s));
    HIP_TRY(c->step_pos.upload(L.step_pos, s));
    HIP_TRY(c->task_step.upload(L.task_step, s));
    HIP_TRY(c->pos_of_rank.upload(L.pos_of_rank, s));
    HIP_TRY(hipStreamSynchronize(s));
    c->graph_bytes = 4ull * (L.up_off.size() + 2 * L.up_tgt.size() + 2ull * L.n + L.in_off.size() +
                             2 * L.in_src.size() + 4ull * L.n + L.step_pos.size() + L.task_step.size());
    // release host copies the kernels no longer need
    std::vector<uint32_t>().swap(L.in_src);
    std::vector<uint32_t>().swap(L.in_w);
    std::vector<uint32_t>().swap(L.up_tgt);
    std::vector<uint32_t>().swap(L.up_w);
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
    bfg::CchInput in;
    in.n = gd->n_nodes;
    in.up_off = gd->up_offsets; in.up_tgt = gd->up_targets; in.up_w = gd->up_weights;
    in.dn_off = gd->down_offsets; in.dn_tgt = gd->down_targets; in.dn_w = gd->down_weights;
    bfg::Layout L;
    std::string err;
    if (!bfg::build_layout(in, bfg::LayoutParams(), L, err)) return fail(BFG_INVALID_PARAMETER, err);
    if (!bfg::check_ancestry(in, L, err)) return fail(BFG_INTERNAL_ERROR, err);
    memset(info, 0, sizeof(*info));
    info->n_nodes = L.n;
    info->m_up = L.m_up;
    info->m_down = L.m_down;
    info->etree_depth = L.max_depth;
    info->n_levels = (uint32_t)L.level_task.size() - 1;
    info->n_tasks = (uint32_t)L.task_step.size() - 1;
    info->n_steps = (uint32_t)L.step_pos.size() - 1;
    return BFG_SUCCESS;
}

int bfg_ctx_get_info(const bfg_ctx *c, bfg_ctx_info *info) {
    if (!c || !info) return fail(BFG_INVALID_PARAMETER, "null argument");
    info->n_nodes = c->L.n;
    info->m_up = c->L.m_up;
    info->m_down = c->L.m_down;
    info->etree_depth = c->L.max_depth;
    info->n_levels = (uint32_t)c->L.level_task.size() - 1;
    info->n_tasks = (uint32_t)c->L.task_step.size() - 1;
    info->n_steps = (uint32_t)c->L.step_pos.size() - 1;
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
    if (lane_capacity != 0) c->lane_cap = (lane_capacity + c->lanes - 1) / c->lanes * c->lanes;
    else if (lanes != 0 && c->lane_cap) c->lane_cap = (c->lane_cap + c->lanes - 1) / c->lanes * c->lanes;
    return ensure_arena(c);
}

static int matrix_impl(bfg_ctx *c, const uint32_t *src, uint32_t S, const uint32_t *tgt, uint32_t T, uint32_t thr,
                       uint32_t *d_out, bool out_is_host, uint32_t *host_out) {
    Req req = single_source_req(src, S, thr);
    int rc = stage_targets(c, tgt, T);
    if (rc) return rc;
    DevBuf<uint32_t> tmp;
    uint32_t tmp_rows = 0;
    return run_all(c, req, [&](uint32_t b0, uint32_t nl, float *oms) -> int {
        if (!out_is_host) return gather_out(c, nl, c->tgt_pos.p, T, d_out, T, b0, nullptr, 0, oms);
        if (tmp_rows < nl) {
            HIP_TRY(tmp.alloc((size_t)nl * T));
            tmp_rows = nl;
        }
        int r2 = gather_out(c, nl, c->tgt_pos.p, T, tmp.p, T, 0, nullptr, 0, oms);
        if (r2) return r2;
        HIP_TRY(hipMemcpy(host_out + (size_t)b0 * T, tmp.p, (size_t)nl * T * 4, hipMemcpyDeviceToHost));
        return BFG_SUCCESS;
    });
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
    return matrix_impl(c, src, S, tgt, T, thr, nullptr, true, out);
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
    return matrix_impl(c, src.data(), S, tgt.data(), T, thr, d_out, false, nullptr);
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
