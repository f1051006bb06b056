// customize.hip — CCH triangle relaxation (customization) on gfx950.
//
// The reference customizes the metric-independent CCH in two steps
// (route/src/customization.rs): a bottom-up pass that gives every arc a first
// weight (original arcs their metric weight, shortcuts the sum over their
// contraction middle), then triangle relaxation to the fixpoint
//     w(x -> y) = min(w(x -> y), w(x -> m) + w(m -> y))
// over every lower triangle: apex m below both x and y, x -> m a DOWN arc of
// x, m -> y an UP arc of m (triangle_relax_parallel, :1083-1200).  The
// reference runs passes over all apexes in parallel with a 64-bit fetch_min
// on the packed (weight, middle) word until a pass changes nothing.
//
// Here the apexes are processed in dependency levels instead.  An arc at
// node v (a DOWN arc x -> v or an UP arc v -> y) is only ever relaxed through
// apexes m' that are lower neighbours of v (an UP arc m' -> v or a DOWN arc
// v -> m'), so with level(v) = 1 + max level of v's lower neighbours every
// arc an apex reads is final when its level runs.  Each triangle is then
// evaluated once with final halves, and one pass over the levels gives what
// the reference's last pass holds: per arc the least of its initial packed
// word and of (w(x -> m) + w(m -> y), m) over its apexes.  Integer mins do not
// depend on the order in which they are taken, so weights and middles equal
// the reference's bit for bit (middle ties go to the smallest apex, as the
// packed order makes them).  The (time, length) variant
// (triangle_relax_lex_parallel, :1253-1440) packs both channels instead and
// elects each shortcut's middle in a second pass over the converged words.
//
// Work layout: a row is one DOWN arc x -> m seen from its apex m (the
// reference's ReverseDownAdj, :832-905); rows are ordered by the level of m,
// then by m and x.  One wavefront takes one row and walks m's UP arcs m -> y
// 64 at a time, one per lane; the arc x -> y is in x's UP list when y > x and
// in its DOWN list otherwise.  Because a separator's nodes hold consecutive
// ranks, y and its arc index advance together along most of a chunk: from an
// anchor (one binary search for a row's first y in a list) each lane guesses
// anchor + (y - y0) and checks that guess with one coalesced load, and only
// the lanes that miss search (in the range the guess bounds).  A plain load
// of the arc's current word screens the atomic, which is issued only when
// the candidate is smaller.  Later chunks of a row start from the last index
// the previous chunk found, so most of them search nothing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bfgpu.h"
#include "errors.hpp"

namespace {

constexpr uint32_t INF = 0xFFFFFFFFu;
constexpr uint64_t NOT_FOUND = ~0ull;
constexpr uint32_t NOT_A_NODE = 0xFFFFFFFFu;
constexpr uint32_t WAVES_PER_WG = 4;

int fail(int code, const std::string &msg) {
    bfg::set_last_error(msg);
    return code;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(e_ == hipErrorOutOfMemory ? BFG_OUT_OF_MEMORY : BFG_DEVICE_ERROR,   \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                  \
    } while (0)

__device__ __forceinline__ uint32_t sat_add(uint32_t a, uint32_t b) { return __builtin_elementwise_add_sat(a, b); }

// first index in a[lo, hi) whose value is >= key
__device__ __forceinline__ uint64_t lower_bound_u32(const uint32_t *__restrict__ a, uint64_t lo, uint64_t hi,
                                                    uint32_t key) {
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// The first index in a[lo, hi) whose value is >= key, for a key and range
// the same in every lane (called by the whole wavefront): each step probes 64
// evenly spaced places at once and a ballot counts those below the key, so
// one dependent load narrows the range 64-fold (two for a list of 4096).
__device__ __forceinline__ uint64_t wave_lower_bound(const uint32_t *__restrict__ a, uint64_t lo, uint64_t hi,
                                                     uint32_t key) {
    const uint64_t lane = threadIdx.x & 63;
    while (hi - lo > 64) {
        const uint64_t step = (hi - lo + 63) / 64;
        const uint64_t p = lo + lane * step;
        const uint32_t k = (uint32_t)__popcll(__ballot(p < hi && a[p] < key));
        if (k == 0) return lo;
        // a[lo + (k-1) step] < key, and a[lo + k step] >= key unless past hi
        const uint64_t nhi = min(hi, lo + (uint64_t)k * step);
        lo = lo + (uint64_t)(k - 1) * step + 1;
        hi = nhi;
    }
    const uint64_t p = lo + lane;
    return lo + (uint64_t)__popcll(__ballot(p < hi && a[p] < key));
}

// A wavefront's place in one of x's arc lists (targets strictly ascending)
// while it walks m's UP targets in ascending order: the list's range and the
// last target it found there.
struct Cursor {
    uint64_t lo, end;   // the list: tgt[lo, end)
    uint64_t last;      // index of the last target found, NOT_FOUND before the first
    uint32_t last_y;
};

// The index of target y in the cursor's list for every lane in `want` (lanes'
// y ascending, above every target found before), NOT_FOUND where y is absent;
// *word receives the packed word at that index.  Targets are distinct
// integers, so y sits at most y - y_b places after a base index holding y_b
// (or, for a base from a search, where y_b would go): the lane guesses that
// place and checks it with one load, loading the word beside it, and searches
// only when the guess misses.  The base is the last target found (no search),
// or for a list's first chunk one wavefront-wide search for the first wanted
// y.  The cursor then moves to the last target this chunk found.
__device__ __forceinline__ uint64_t locate(const uint32_t *__restrict__ tgt, const unsigned long long *pk, Cursor &c,
                                           uint32_t y, bool want, unsigned long long *word) {
    const unsigned long long b = __ballot(want);
    if (!b) return NOT_FOUND;
    uint64_t base = c.last, lo = c.last + 1;
    uint32_t yb = c.last_y;
    if (c.last == NOT_FOUND) {
        const int f = __ffsll((long long)b) - 1;
        yb = (uint32_t)__builtin_amdgcn_readfirstlane(__shfl((int)y, f, 64));
        base = lo = wave_lower_bound(tgt, c.lo, c.end, yb);
    }
    uint64_t idx = NOT_FOUND;
    if (want) {
        const uint64_t g = base + (uint64_t)(y - yb);
        if (g < c.end) {
            const uint32_t t = tgt[g];
            const unsigned long long w = pk[g];
            if (t == y) {
                idx = g;
                *word = w;
            }
        }
        if (idx == NOT_FOUND) {
            const uint64_t p = lower_bound_u32(tgt, lo, g < c.end ? g : c.end, y);
            if (p < c.end && tgt[p] == y) {
                idx = p;
                *word = pk[p];
            }
        }
    }
    const unsigned long long fb = __ballot(idx != NOT_FOUND);
    if (fb) {
        const int l = 63 - __clzll((long long)fb);
        c.last = __shfl(idx, l, 64);
        c.last_y = (uint32_t)__shfl((int)y, l, 64);
    }
    return idx;
}

// Enumerates the triangles of row r (DOWN arc x -> m, apex m) chunk by chunk
// (64 of m's UP arcs each) and calls f(ok, p_my, idx, is_down, p_xy) once per
// chunk on every lane: ok when the lane holds a triangle with both halves
// finite, p_my the word of m -> y, idx the index of x -> y in x's DOWN
// (is_down) or UP list, or NOT_FOUND, and p_xy that arc's word as read.
template <class F>
__device__ __forceinline__ void for_row_triangles(uint32_t m, uint32_t x, const uint64_t *__restrict__ up_off,
                                                  const uint32_t *__restrict__ up_tgt,
                                                  const uint64_t *__restrict__ dn_off,
                                                  const uint32_t *__restrict__ dn_tgt,
                                                  const unsigned long long *up_pk,
                                                  const unsigned long long *dn_pk, F &&f) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t j0 = up_off[m], j1 = up_off[m + 1];
    Cursor cu{up_off[x], up_off[x + 1], NOT_FOUND, 0};
    Cursor cd{dn_off[x], dn_off[x + 1], NOT_FOUND, 0};
    for (uint64_t jb = j0; jb < j1; jb += 64) {
        const uint64_t j = jb + lane;
        const bool act = j < j1;
        const uint32_t y = act ? up_tgt[j] : INF;
        const unsigned long long pmy = act ? up_pk[j] : ~0ull;
        const bool ok = act && y != x && (uint32_t)(pmy >> 32) != INF;
        const bool isdn = y < x;
        unsigned long long wd = ~0ull, wu = ~0ull;
        const uint64_t idn = locate(dn_tgt, dn_pk, cd, y, ok && isdn, &wd);
        const uint64_t iup = locate(up_tgt, up_pk, cu, y, ok && !isdn, &wu);
        f(ok, pmy, isdn ? idn : iup, isdn, isdn ? wd : wu);
    }
}

// Rows are dealt to wavefronts round-robin over a grid of at most MAX_WGS
// workgroups (a wavefront takes rows w, w + W, .. for W wavefronts in the
// grid): consecutive wavefronts take consecutive rows, usually of one apex,
// and no workgroup is launched per handful of rows.
constexpr uint32_t MAX_WGS = 4096;

__device__ __forceinline__ uint64_t first_row() {
    return (uint64_t)blockIdx.x * WAVES_PER_WG + (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}
__device__ __forceinline__ uint64_t row_stride() { return (uint64_t)gridDim.x * WAVES_PER_WG; }

// One dependency level: rows [r0, r1), one wavefront per row at a time.
// counters[0] += triangles evaluated (both halves finite), counters[1] += atomic
// mins issued (the candidate was below the arc's word as read).
template <bool LEX>
__global__ __launch_bounds__(256) void relax_level_kernel(const uint32_t *__restrict__ row_m,
                                                          const uint32_t *__restrict__ row_x,
                                                          const uint32_t *__restrict__ row_e, uint64_t r0,
                                                          uint64_t r1, const uint64_t *__restrict__ up_off,
                                                          const uint32_t *__restrict__ up_tgt,
                                                          const uint64_t *__restrict__ dn_off,
                                                          const uint32_t *__restrict__ dn_tgt,
                                                          unsigned long long *up_pk, unsigned long long *dn_pk,
                                                          unsigned long long *__restrict__ counters) {
    uint32_t ntri = 0, nimp = 0;
    for (uint64_t r = r0 + first_row(); r < r1; r += row_stride()) {
        const uint32_t m = row_m[r], x = row_x[r];
        const unsigned long long pxm = dn_pk[row_e[r]];
        const uint32_t wxm = (uint32_t)(pxm >> 32), lxm = (uint32_t)pxm;
        if (wxm == INF) continue;
        for_row_triangles(m, x, up_off, up_tgt, dn_off, dn_tgt, up_pk, dn_pk,
                          [&](bool ok, unsigned long long pmy, uint64_t idx, bool isdn, unsigned long long pxy) {
                              bool imp = false;
                              if (ok && idx != NOT_FOUND) {
                                  const uint32_t w = sat_add(wxm, (uint32_t)(pmy >> 32));
                                  const uint32_t lo = LEX ? sat_add(lxm, (uint32_t)pmy) : m;
                                  const unsigned long long cand = ((unsigned long long)w << 32) | lo;
                                  // words only fall, so a candidate not below the word read
                                  // cannot lower it; the others go to the atomic, whose old
                                  // value is not needed
                                  if (cand < pxy) {
                                      atomicMin((isdn ? dn_pk : up_pk) + idx, cand);
                                      imp = true;
                                  }
                              }
                              ntri += __popcll(__ballot(ok));
                              nimp += __popcll(__ballot(imp));
                          });
    }
    if ((threadIdx.x & 63) == 0 && (ntri || nimp)) {
        atomicAdd(counters, (unsigned long long)ntri);
        atomicAdd(counters + 1, (unsigned long long)nimp);
    }
}

// The (time, length) variant's middle election over converged words: for a
// shortcut arc, the least apex whose halves sum to exactly its word
// (customization.rs:1362-1405).
__global__ __launch_bounds__(256) void elect_middles_kernel(
    const uint32_t *__restrict__ row_m, const uint32_t *__restrict__ row_x, const uint32_t *__restrict__ row_e,
    uint64_t n_rows, const uint64_t *__restrict__ up_off, const uint32_t *__restrict__ up_tgt,
    const uint64_t *__restrict__ dn_off, const uint32_t *__restrict__ dn_tgt, const unsigned long long *up_pk,
    const unsigned long long *dn_pk, const unsigned long long *__restrict__ up_sc,
    const unsigned long long *__restrict__ dn_sc, uint32_t *up_mid, uint32_t *dn_mid) {
    for (uint64_t r = first_row(); r < n_rows; r += row_stride()) {
        const uint32_t m = row_m[r], x = row_x[r];
        const unsigned long long pxm = dn_pk[row_e[r]];
        const uint32_t txm = (uint32_t)(pxm >> 32), lxm = (uint32_t)pxm;
        if (txm == INF) continue;
        for_row_triangles(m, x, up_off, up_tgt, dn_off, dn_tgt, up_pk, dn_pk,
                          [&](bool ok, unsigned long long pmy, uint64_t idx, bool isdn, unsigned long long pxy) {
                              if (!ok || idx == NOT_FOUND) return;
                              const unsigned long long *sc = isdn ? dn_sc : up_sc;
                              if (!((sc[idx >> 6] >> (idx & 63)) & 1ull)) return;
                              const unsigned long long cand =
                                  ((unsigned long long)sat_add(txm, (uint32_t)(pmy >> 32)) << 32) |
                                  sat_add(lxm, (uint32_t)pmy);
                              if (cand == pxy) atomicMin((isdn ? dn_mid : up_mid) + idx, m);
                          });
    }
}

// out[i] = hi[i] << 32 | (lo ? lo[i] : lo_const)
__global__ void pack_kernel(const uint32_t *__restrict__ hi, const uint32_t *__restrict__ lo, uint32_t lo_const,
                            uint64_t n, unsigned long long *__restrict__ out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = ((unsigned long long)hi[i] << 32) | (lo ? lo[i] : lo_const);
}

__global__ void unpack_kernel(const unsigned long long *__restrict__ pk, uint64_t n, uint32_t *__restrict__ hi,
                              uint32_t *__restrict__ lo) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long v = pk[i];
        hi[i] = (uint32_t)(v >> 32);
        if (lo) lo[i] = (uint32_t)v;
    }
}

__global__ void fill_u32_kernel(uint32_t *__restrict__ p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

uint32_t grid_for(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 65536)); }

// Device buffer freed on scope exit.
template <class T>
struct Buf {
    T *p = nullptr;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t count) { return hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T)); }
    hipError_t upload(const T *src, size_t count, hipStream_t s) {
        hipError_t e = alloc(count);
        if (e != hipSuccess || count == 0) return e;
        return hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s);
    }
};

struct StreamGuard {
    hipStream_t s = nullptr;
    ~StreamGuard() {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
};

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// Host side: the CSR checks find_edge_index relies on (ascending targets per
// tail, UP targets above the tail and DOWN targets below it), the dependency
// levels and the rows in level order.
struct Rows {
    std::vector<uint32_t> m, x, e;
    std::vector<uint64_t> level_off;  // rows of level L: [level_off[L], level_off[L+1])
};

// The arcs of node v break a rule find_edge_index relies on: null if they
// keep them (offsets already checked).
const char *node_error(const bfg_graph_desc *g, uint32_t v) {
    const uint32_t n = g->n_nodes;
    const uint64_t u0 = g->up_offsets[v], u1 = g->up_offsets[v + 1];
    for (uint64_t i = u0; i < u1; ++i) {
        const uint32_t t = g->up_targets[i];
        if (t <= v || t >= n || (i > u0 && t <= g->up_targets[i - 1]))
            return "must have UP targets ascending, above the node and < n_nodes";
    }
    const uint64_t d0 = g->down_offsets[v], d1 = g->down_offsets[v + 1];
    for (uint64_t i = d0; i < d1; ++i) {
        const uint32_t t = g->down_targets[i];
        if (t >= v || (i > d0 && t <= g->down_targets[i - 1])) return "must have DOWN targets ascending and below it";
    }
    return nullptr;
}

int check_topo(const bfg_graph_desc *g) {
    if (!g || !g->up_offsets || !g->down_offsets) return fail(BFG_INVALID_PARAMETER, "null topology");
    const uint32_t n = g->n_nodes;
    if (g->up_offsets[0] != 0 || g->down_offsets[0] != 0) return fail(BFG_INVALID_PARAMETER, "offsets[0] != 0");
    if ((g->up_offsets[n] && (!g->up_targets || !g->up_weights)) ||
        (g->down_offsets[n] && (!g->down_targets || !g->down_weights)))
        return fail(BFG_INVALID_PARAMETER, "null arc arrays");
    // offsets first (ascending offsets keep every node's arcs inside the arrays)
    for (uint32_t v = 0; v < n; ++v)
        if (g->up_offsets[v + 1] < g->up_offsets[v] || g->down_offsets[v + 1] < g->down_offsets[v])
            return fail(BFG_INVALID_PARAMETER, "offsets not ascending at node " + std::to_string(v));
    if (g->down_offsets[n] >= (1ull << 32)) return fail(BFG_NOT_SUPPORTED, "more than 2^32 - 1 DOWN arcs");
    // then the arcs, by threads over ranges of nodes; the first failing node is reported
    const uint64_t m = g->up_offsets[n] + g->down_offsets[n];
    const uint32_t T = m < (1u << 22) ? 1u : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<uint32_t> first(T, NOT_A_NODE);
    auto scan = [&](uint32_t t) {
        const uint32_t v0 = (uint32_t)((uint64_t)n * t / T), v1 = (uint32_t)((uint64_t)n * (t + 1) / T);
        for (uint32_t v = v0; v < v1; ++v)
            if (node_error(g, v)) {
                first[t] = v;
                return;
            }
    };
    if (T == 1) {
        scan(0);
    } else {
        std::vector<std::thread> th;
        for (uint32_t t = 0; t < T; ++t) th.emplace_back(scan, t);
        for (auto &h : th) h.join();
    }
    const uint32_t v = *std::min_element(first.begin(), first.end());
    if (v != NOT_A_NODE) return fail(BFG_INVALID_PARAMETER, "node " + std::to_string(v) + " " + node_error(g, v));
    return BFG_SUCCESS;
}

Rows build_rows(const bfg_graph_desc *g) {
    const uint32_t n = g->n_nodes;
    // level(v) = 1 + max level over v's lower neighbours: DOWN targets of v and
    // tails of UP arcs into v (ascending v: a node's level is final before its
    // UP arcs pass it on)
    std::vector<uint32_t> lev(n, 0);
    for (uint32_t v = 0; v < n; ++v) {
        uint32_t l = lev[v];
        for (uint64_t i = g->down_offsets[v]; i < g->down_offsets[v + 1]; ++i)
            l = std::max(l, lev[g->down_targets[i]] + 1);
        lev[v] = l;
        for (uint64_t i = g->up_offsets[v]; i < g->up_offsets[v + 1]; ++i) {
            uint32_t &t = lev[g->up_targets[i]];
            t = std::max(t, l + 1);
        }
    }
    const uint32_t nlev = n ? *std::max_element(lev.begin(), lev.end()) + 1 : 0;
    // rows per apex (DOWN in-degree), apexes ordered by (level, rank)
    std::vector<uint64_t> indeg(n, 0);
    const uint64_t md = g->down_offsets[n];
    for (uint64_t i = 0; i < md; ++i) indeg[g->down_targets[i]]++;
    std::vector<uint64_t> lev_rows(nlev + 1, 0);
    for (uint32_t v = 0; v < n; ++v) lev_rows[lev[v] + 1] += indeg[v];
    for (uint32_t L = 0; L < nlev; ++L) lev_rows[L + 1] += lev_rows[L];
    std::vector<uint64_t> start(n);
    {
        std::vector<uint64_t> cur(lev_rows.begin(), lev_rows.end() - 1);
        for (uint32_t v = 0; v < n; ++v) {
            start[v] = cur[lev[v]];
            cur[lev[v]] += indeg[v];
        }
    }
    Rows R;
    R.m.resize(md);
    R.x.resize(md);
    R.e.resize(md);
    // thread t writes the rows of the apexes m = t (mod T), scanning every DOWN
    // arc in tail order, so each apex's rows keep ascending x
    const uint32_t T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto fill = [&](uint32_t t) {
        for (uint32_t x = 0; x < n; ++x)
            for (uint64_t i = g->down_offsets[x]; i < g->down_offsets[x + 1]; ++i) {
                const uint32_t m = g->down_targets[i];
                if (m % T != t) continue;
                const uint64_t r = start[m]++;
                R.m[r] = m;
                R.x[r] = x;
                R.e[r] = (uint32_t)i;
            }
    };
    if (md < (1u << 22)) {
        for (uint32_t t = 0; t < T; ++t) fill(t);
    } else {
        std::vector<std::thread> th;
        for (uint32_t t = 0; t < T; ++t) th.emplace_back(fill, t);
        for (auto &h : th) h.join();
    }
    R.level_off = std::move(lev_rows);
    return R;
}

// The relaxation proper.  hi_*: the arcs' weights (times); lo_*: the low words
// (middles, or lengths when lex).  On return hi/lo hold the converged words;
// for lex, mid_* receive the elected middles (BFG_INF where none).
int relax(const bfg_graph_desc *g, int device, bool lex, const uint32_t *up_lo, const uint32_t *dn_lo,
          const uint64_t *up_sc, const uint64_t *dn_sc, uint32_t *up_hi_out, uint32_t *dn_hi_out,
          uint32_t *up_lo_out, uint32_t *dn_lo_out, uint32_t *up_mid_out, uint32_t *dn_mid_out,
          bfg_relax_stats *st) {
    const auto t0 = std::chrono::steady_clock::now();
    int rc = check_topo(g);
    if (rc) return rc;
    const uint32_t n = g->n_nodes;
    const uint64_t mu = g->up_offsets[n], md = g->down_offsets[n];
    HIP_TRY(hipSetDevice(device));
    Rows R = build_rows(g);
    const uint32_t nlev = (uint32_t)R.level_off.size() - 1;

    StreamGuard sg;
    HIP_TRY(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
    hipStream_t s = sg.s;
    Buf<uint64_t> d_uoff, d_doff;
    Buf<uint32_t> d_utgt, d_dtgt, d_rm, d_rx, d_re, d_tmp_hi, d_tmp_lo;
    Buf<unsigned long long> d_upk, d_dpk, d_cnt, d_usc, d_dsc;
    HIP_TRY(d_uoff.upload(g->up_offsets, (size_t)n + 1, s));
    HIP_TRY(d_doff.upload(g->down_offsets, (size_t)n + 1, s));
    HIP_TRY(d_utgt.upload(g->up_targets, mu, s));
    HIP_TRY(d_dtgt.upload(g->down_targets, md, s));
    HIP_TRY(d_rm.upload(R.m.data(), md, s));
    HIP_TRY(d_rx.upload(R.x.data(), md, s));
    HIP_TRY(d_re.upload(R.e.data(), md, s));
    HIP_TRY(d_upk.alloc(mu));
    HIP_TRY(d_dpk.alloc(md));
    HIP_TRY(d_cnt.alloc(2));
    HIP_TRY(hipMemsetAsync(d_cnt.p, 0, 2 * sizeof(unsigned long long), s));
    const uint64_t mmax = std::max(mu, md);
    HIP_TRY(d_tmp_hi.alloc(mmax));
    HIP_TRY(d_tmp_lo.alloc(mmax));
    // packed words: (weight, middle or length)
    auto pack = [&](const uint32_t *hi, const uint32_t *lo, uint64_t cnt, unsigned long long *out) -> int {
        if (!cnt) return BFG_SUCCESS;
        HIP_TRY(hipMemcpyAsync(d_tmp_hi.p, hi, cnt * 4, hipMemcpyHostToDevice, s));
        if (lo) HIP_TRY(hipMemcpyAsync(d_tmp_lo.p, lo, cnt * 4, hipMemcpyHostToDevice, s));
        pack_kernel<<<grid_for(cnt), 256, 0, s>>>(d_tmp_hi.p, lo ? d_tmp_lo.p : nullptr, INF, cnt, out);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(s));  // the staging buffers are reused
        return BFG_SUCCESS;
    };
    if ((rc = pack(g->up_weights, up_lo, mu, d_upk.p))) return rc;
    if ((rc = pack(g->down_weights, dn_lo, md, d_dpk.p))) return rc;

    EventPair ev;
    HIP_TRY(hipEventCreate(&ev.a));
    HIP_TRY(hipEventCreate(&ev.b));
    HIP_TRY(hipEventRecord(ev.a, s));
    uint32_t launches = 0;
    for (uint32_t L = 0; L < nlev; ++L) {
        const uint64_t r0 = R.level_off[L], r1 = R.level_off[L + 1];
        if (r1 == r0) continue;
        const uint64_t wgs = std::min<uint64_t>((r1 - r0 + WAVES_PER_WG - 1) / WAVES_PER_WG, MAX_WGS);
        if (lex)
            relax_level_kernel<true><<<(uint32_t)wgs, WAVES_PER_WG * 64, 0, s>>>(
                d_rm.p, d_rx.p, d_re.p, r0, r1, d_uoff.p, d_utgt.p, d_doff.p, d_dtgt.p, d_upk.p, d_dpk.p,
                d_cnt.p);
        else
            relax_level_kernel<false><<<(uint32_t)wgs, WAVES_PER_WG * 64, 0, s>>>(
                d_rm.p, d_rx.p, d_re.p, r0, r1, d_uoff.p, d_utgt.p, d_doff.p, d_dtgt.p, d_upk.p, d_dpk.p,
                d_cnt.p);
        ++launches;
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev.b, s));

    if (lex && md) {
        HIP_TRY(d_usc.upload(reinterpret_cast<const unsigned long long *>(up_sc), (mu + 63) / 64, s));
        HIP_TRY(d_dsc.upload(reinterpret_cast<const unsigned long long *>(dn_sc), (md + 63) / 64, s));
        // the middles reuse the staging buffers: up in d_tmp_hi, down in d_tmp_lo
        fill_u32_kernel<<<grid_for(mmax), 256, 0, s>>>(d_tmp_hi.p, mmax, INF);
        fill_u32_kernel<<<grid_for(mmax), 256, 0, s>>>(d_tmp_lo.p, mmax, INF);
        const uint64_t wgs = std::min<uint64_t>((md + WAVES_PER_WG - 1) / WAVES_PER_WG, MAX_WGS);
        elect_middles_kernel<<<(uint32_t)wgs, WAVES_PER_WG * 64, 0, s>>>(d_rm.p, d_rx.p, d_re.p, md, d_uoff.p,
                                                                         d_utgt.p, d_doff.p, d_dtgt.p, d_upk.p,
                                                                         d_dpk.p, d_usc.p, d_dsc.p, d_tmp_hi.p,
                                                                         d_tmp_lo.p);
        HIP_TRY(hipGetLastError());
        if (up_mid_out && mu) HIP_TRY(hipMemcpyAsync(up_mid_out, d_tmp_hi.p, mu * 4, hipMemcpyDeviceToHost, s));
        if (dn_mid_out) HIP_TRY(hipMemcpyAsync(dn_mid_out, d_tmp_lo.p, md * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    } else if (lex) {
        if (up_mid_out) std::fill(up_mid_out, up_mid_out + mu, INF);
    }
    auto unpack = [&](const unsigned long long *pk, uint64_t cnt, uint32_t *hi, uint32_t *lo) -> int {
        if (!cnt) return BFG_SUCCESS;
        unpack_kernel<<<grid_for(cnt), 256, 0, s>>>(pk, cnt, d_tmp_hi.p, lo ? d_tmp_lo.p : nullptr);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(hi, d_tmp_hi.p, cnt * 4, hipMemcpyDeviceToHost, s));
        if (lo) HIP_TRY(hipMemcpyAsync(lo, d_tmp_lo.p, cnt * 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        return BFG_SUCCESS;
    };
    if ((rc = unpack(d_upk.p, mu, up_hi_out, up_lo_out))) return rc;
    if ((rc = unpack(d_dpk.p, md, dn_hi_out, dn_lo_out))) return rc;
    unsigned long long cnt[2] = {0, 0};
    HIP_TRY(hipMemcpy(cnt, d_cnt.p, sizeof(cnt), hipMemcpyDeviceToHost));
    float kms = 0.f;
    HIP_TRY(hipEventElapsedTime(&kms, ev.a, ev.b));
    if (st) {
        st->triangles = cnt[0];
        st->improvements = cnt[1];
        st->levels = nlev;
        st->launches = launches;
        st->passes = 1;
        st->kernel_ms = kms;
        st->total_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return BFG_SUCCESS;
}

}  // namespace

extern "C" {

int bfg_triangle_relax(const bfg_graph_desc *topo, int device, const uint32_t *up_middle,
                       const uint32_t *down_middle, uint32_t *up_weights_out, uint32_t *down_weights_out,
                       uint32_t *up_middle_out, uint32_t *down_middle_out, bfg_relax_stats *stats) {
    if (!topo || !up_weights_out || !down_weights_out) return fail(BFG_INVALID_PARAMETER, "null argument");
    const uint32_t n = topo->n_nodes;
    if (!topo->up_offsets || !topo->down_offsets) return fail(BFG_INVALID_PARAMETER, "null topology");
    const uint64_t mu = topo->up_offsets[n], md = topo->down_offsets[n];
    // middles left out start at u32::MAX, the reference's mark of an original arc
    std::vector<uint32_t> mid_up_tmp, mid_dn_tmp;
    uint32_t *umo = up_middle_out, *dmo = down_middle_out;
    if (!umo) {
        mid_up_tmp.resize(mu);
        umo = mid_up_tmp.data();
    }
    if (!dmo) {
        mid_dn_tmp.resize(md);
        dmo = mid_dn_tmp.data();
    }
    return relax(topo, device, false, up_middle, down_middle, nullptr, nullptr, up_weights_out, down_weights_out, umo,
                 dmo, nullptr, nullptr, stats);
}

int bfg_triangle_relax_lex(const bfg_graph_desc *topo, int device, const uint32_t *up_length,
                           const uint32_t *down_length, const uint64_t *up_is_shortcut,
                           const uint64_t *down_is_shortcut, const uint32_t *up_middle, const uint32_t *down_middle,
                           uint32_t *up_times_out, uint32_t *down_times_out, uint32_t *up_middle_out,
                           uint32_t *down_middle_out, bfg_relax_stats *stats) {
    if (!topo || !up_length || !down_length || !up_is_shortcut || !down_is_shortcut || !up_times_out ||
        !down_times_out || !up_middle_out || !down_middle_out)
        return fail(BFG_INVALID_PARAMETER, "null argument");
    const uint32_t n = topo->n_nodes;
    if (!topo->up_offsets || !topo->down_offsets) return fail(BFG_INVALID_PARAMETER, "null topology");
    const uint64_t mu = topo->up_offsets[n], md = topo->down_offsets[n];
    int rc = relax(topo, device, true, up_length, down_length, up_is_shortcut, down_is_shortcut, up_times_out,
                   down_times_out, nullptr, nullptr, up_middle_out, down_middle_out, stats);
    if (rc) return rc;
    // a shortcut with no elected apex keeps its given middle (:1420-1437)
    for (uint64_t i = 0; i < mu; ++i)
        if (up_middle_out[i] == INF) up_middle_out[i] = up_middle ? up_middle[i] : INF;
    for (uint64_t i = 0; i < md; ++i)
        if (down_middle_out[i] == INF) down_middle_out[i] = down_middle ? down_middle[i] : INF;
    return BFG_SUCCESS;
}

}  // extern "C"
