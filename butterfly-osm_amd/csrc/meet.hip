// meet.hip — the join of the many-to-many (bucket) matrix on gfx950.
//
// Forward labels fw(s, x) (upward phase of the source lanes on the CCH) and
// backward labels bw(t, x) (upward phase of the target lanes on the
// transposed CCH) are finite only on the root paths of s and t, so
// d(s, t) = min over common ancestors x of fw(s, x) + bw(t, x): the join of
// the reference's bucket many-to-many (backward_join_parallel_prefix,
// route/src/matrix/bucket_ch.rs:3935-4010) with the buckets replaced by the
// tree.  A tile is 64 source lanes x 64 target lanes (one wave group of each
// side); its common meeting nodes are the top-region segments both wave
// groups reach (an AND of two reach bitsets), and the tile is a min-plus
// product over their positions.  Labels of lanes whose node is below a
// reached segment but not under a given position are INF there, so every
// position of a common segment can enter the product.
//
// meet_join_kernel computes in u32 with saturating adds (v_add_u32 clamp) and
// v_min3; meet_join16_kernel in packed u16 on capped label copies, with the
// tiles it cannot settle in 16 bits recomputed by meet_join_kernel.  Either
// way the cells equal the reference's bucket join and its PHAST matrix bit
// for bit.
#include "meet.hpp"

#include <algorithm>

namespace bfg_meet {
namespace {

constexpr uint32_t INF = 0xFFFFFFFFu;
constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint32_t XCDS = 8;  // MI355X accelerator complex dies, each with its own L2

__device__ __forceinline__ uint32_t sat_add(uint32_t a, uint32_t b) { return __builtin_elementwise_add_sat(a, b); }

// ---------------------------------------------------------------------------
// reach: workgroup (word w, wave group g) tests segments 64w .. 64w+63; lane l
// checks its seeds against the segment head's pre-order interval and a ballot
// gives the bit.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void meet_reach_kernel(const uint32_t *__restrict__ seed_off,
                                                        const uint32_t *__restrict__ seed_pre, uint32_t n_lanes,
                                                        const uint32_t *__restrict__ pos_pre,
                                                        const uint32_t *__restrict__ pos_sub,
                                                        const uint32_t *__restrict__ seg_pos, uint32_t nseg,
                                                        uint32_t words, unsigned long long *__restrict__ reach) {
    const uint32_t w = blockIdx.x, g = blockIdx.y, gl = g * 64 + threadIdx.x;
    uint32_t j0 = 0, j1 = 0;
    if (gl < n_lanes) {
        j0 = seed_off[gl];
        j1 = seed_off[gl + 1];
    }
    const uint32_t first = gl < n_lanes && j1 > j0 ? seed_pre[j0] : 0;
    unsigned long long word = 0;
    const uint32_t s1 = min(nseg, w * 64 + 64);
    for (uint32_t s = w * 64; s < s1; ++s) {
        const uint32_t p = seg_pos[s];
        const uint32_t pr = pos_pre[p], ps = pos_sub[p];
        bool r = j1 > j0 && first - pr < ps;
        for (uint32_t j = j0 + 1; j < j1; ++j) r |= seed_pre[j] - pr < ps;
        if (__ballot(r)) word |= 1ull << (s - w * 64);
    }
    if (threadIdx.x == 0) reach[(size_t)g * words + w] = word;
}

// ---------------------------------------------------------------------------
// join: one wavefront per tile (source wave group A, target wave group B).
// Thread (i, j) = (lane / 8, lane % 8) holds the 8 x 8 cells of source lanes
// 64A + 8i .. and target lanes 64B + 8j .. in registers; per position x it
// loads its 8 source labels and 8 target labels (two 16-byte loads each; the
// 8 threads of a row share an address) and relaxes 64 cells.  Two positions
// per step: acc = min3(acc, a0 + b0, a1 + b1).  With both sides' segment
// minima (per 8 lanes) a tile takes its common segments deepest first and
// skips a segment where no thread's 8 x 8 block could fall.
// Tiles go to the XCDs in contiguous ranges (block b runs on XCD b % 8), so
// the tiles of one source wave group run side by side on one XCD and share its
// labels in that XCD's L2.  With `splits` > 1 (few tiles), split sp of a tile
// takes its common segments k = sp, sp + splits, .. and merges its cells with
// atomic min into an output filled with INF beforehand.
// ---------------------------------------------------------------------------
struct JoinArgs {
    Side A, B;
    const uint32_t *seg_pos;
    uint32_t words, nseg, wga, wgb, thr, splits;
    uint32_t *out;
    uint64_t ld;
    unsigned long long *visits;
    const uint32_t *only;  // tiles to compute (null: all)
    uint32_t *redo;        // 16-bit join: tiles the u32 join recomputes
};

__device__ __forceinline__ void load8(uint32_t (&v)[8], const uint32_t *p) {
    const uint4 lo = *reinterpret_cast<const uint4 *>(p);
    const uint4 hi = *reinterpret_cast<const uint4 *>(p + 4);
    v[0] = lo.x; v[1] = lo.y; v[2] = lo.z; v[3] = lo.w;
    v[4] = hi.x; v[5] = hi.y; v[6] = hi.z; v[7] = hi.w;
}

__global__ __launch_bounds__(64) void meet_join_kernel(JoinArgs a) {
    const uint32_t tiles = a.wga * a.wgb * a.splits;
    const uint32_t per_xcd = (tiles + XCDS - 1) / XCDS;
    const uint32_t u = (blockIdx.x % XCDS) * per_xcd + blockIdx.x / XCDS;
    if (u >= tiles) return;
    const uint32_t t = u / a.splits, sp = u % a.splits;
    if (a.only && a.only[t] == 0) return;  // wave-uniform; this kernel has no barrier
    const uint32_t ta = t / a.wgb, tb = t % a.wgb;
    const uint32_t i = threadIdx.x >> 3, j = threadIdx.x & 7;
    const uint32_t la = ta * 64 + 8 * i, lb = tb * 64 + 8 * j;
    const uint32_t kla = a.A.kl, klb = a.B.kl;
    const uint32_t *fa = a.A.lab + (size_t)(la / kla) * a.A.stride * kla + la % kla;
    const uint32_t *fb = a.B.lab + (size_t)(lb / klb) * a.B.stride * klb + lb % klb;
    const unsigned long long *ra = a.A.reach + (size_t)ta * a.words;
    const unsigned long long *rb = a.B.reach + (size_t)tb * a.words;
    // cells of lanes past either side's count start at 0: they are never
    // written out, and they cannot raise the tile's largest cell
    uint32_t acc[64];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[r * 8 + c] = la + r < a.A.n_lanes && lb + c < a.B.n_lanes ? INF : 0u;
    const bool prune = a.A.segmin && a.B.segmin;
    const uint32_t *sma = prune ? a.A.segmin + (size_t)ta * a.nseg * 8 + i : nullptr;
    const uint32_t *smb = prune ? a.B.segmin + (size_t)tb * a.nseg * 8 + j : nullptr;
    uint32_t k = 0, seen = 0;  // common segments seen so far, positions visited
    // deepest segments first (a segment's index grows with its depth), so
    // the cells are tight by the time the wide segments near the root come
    for (uint32_t w = a.words; w-- > 0;) {
        unsigned long long m = ra[w] & rb[w];
        while (m) {
            const uint32_t bit = 63u - (uint32_t)__builtin_clzll(m);
            m &= ~(1ull << bit);
            const uint32_t s = w * 64 + bit;
            if (k++ % a.splits != sp) continue;
            if (prune) {
                // every sum through s for this thread's 8 x 8 cells is at least
                // its lane groups' minima: where that bound reaches the block's
                // largest cell (or passes thr) the segment lowers no kept cell;
                // the wavefront skips it when that holds for every block
                const uint32_t bound = sat_add(sma[(size_t)s * 8], smb[(size_t)s * 8]);
                uint32_t top = 0;
#pragma unroll
                for (int q = 0; q < 64; ++q) top = max(top, acc[q]);
                if (__ballot(bound < top && bound <= a.thr) == 0) continue;
            }
            const uint32_t p0 = a.seg_pos[s], p1 = a.seg_pos[s + 1];
            seen += p1 - p0;
            for (uint32_t x = p0; x < p1; x += 2) {
                const uint32_t y = min(x + 1, p1 - 1);  // an odd last position pairs with itself
                uint32_t a0[8], a1[8], b0[8], b1[8];
                load8(a0, fa + (size_t)x * kla);
                load8(b0, fb + (size_t)x * klb);
                load8(a1, fa + (size_t)y * kla);
                load8(b1, fb + (size_t)y * klb);
#pragma unroll
                for (int r = 0; r < 8; ++r)
#pragma unroll
                    for (int c = 0; c < 8; ++c)
                        acc[r * 8 + c] = min(acc[r * 8 + c], min(sat_add(a0[r], b0[c]), sat_add(a1[r], b1[c])));
            }
        }
    }
    if (a.visits && threadIdx.x == 0 && seen) atomicAdd(a.visits, (unsigned long long)seen);
    uint32_t col[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) col[c] = lb + c < a.B.n_lanes ? a.B.index[lb + c] : 0u;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        if (la + r >= a.A.n_lanes) break;
        uint32_t *o = a.out + (uint64_t)a.A.index[la + r] * a.ld;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (lb + c >= a.B.n_lanes) continue;
            const uint32_t v = acc[r * 8 + c] > a.thr ? INF : acc[r * 8 + c];
            if (a.splits == 1) o[col[c]] = v;
            else if (v != INF) atomicMin(o + col[c], v);
        }
    }
}

// ---------------------------------------------------------------------------
// join16: meet_join_kernel on the 16-bit label copies, one wavefront per tile
// (splits 1).  Thread (i, j) holds its 8 x 8 cells as 32 packed pairs:
// acc[4r + h] = cells (source lane 8i + r; target lanes 8j + 2h, 8j + 2h + 1).
// Per position one 16-byte load brings the thread's 8 source labels and one
// its 8 target labels (already paired as the cells are); a source label goes
// to both halves by an operand swizzle, so a pair of cells costs one
// v_pk_add_u16 (clamp) and one v_pk_min_u16.  A label is 0xFFFF for INF and
// min(label, CAP16) else, so the sum of two finite labels never saturates
// (at most 0xFFFE) and a sum through INF is 0xFFFF: a cell of 0xFFFF is INF,
// a cell below CAP16 is the u32 join's (a sum through a capped label is at
// least CAP16), and a tile with a kept cell in [CAP16, 0xFFFE] while thr
// admits such a value is flagged for the u32 join to redo.
// ---------------------------------------------------------------------------
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ u16x2 as_u16x2(uint32_t x) { return __builtin_bit_cast(u16x2, x); }

// the 8 16-bit labels of a thread's lanes at position x of a pair's copy
__device__ __forceinline__ uint4 ld16(const uint16_t *p, uint32_t x) {
    return *reinterpret_cast<const uint4 *>(p + (size_t)x * 64);
}

__global__ __launch_bounds__(64) void meet_join16_kernel(JoinArgs a) {
    const uint32_t tiles = a.wga * a.wgb;
    const uint32_t per_xcd = (tiles + XCDS - 1) / XCDS;
    const uint32_t t = (blockIdx.x % XCDS) * per_xcd + blockIdx.x / XCDS;
    if (t >= tiles) return;  // wave-uniform; this kernel has no barrier
    const uint32_t ta = t / a.wgb, tb = t % a.wgb;
    const uint32_t i = threadIdx.x >> 3, j = threadIdx.x & 7;
    const uint32_t la = ta * 64 + 8 * i, lb = tb * 64 + 8 * j;
    const uint32_t nA = a.A.n_lanes, nB = a.B.n_lanes;
    const unsigned long long *ra = a.A.reach + (size_t)ta * a.words;
    const unsigned long long *rb = a.B.reach + (size_t)tb * a.words;
    const uint32_t *qa_of = a.A.slot + (size_t)ta * a.nseg, *qb_of = a.B.slot + (size_t)tb * a.nseg;
    // cells of lanes past either side's count start at 0, as in meet_join_kernel
    u16x2 acc[32];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const bool ok = la + r < nA;
            acc[4 * r + h] = (u16x2){(unsigned short)(ok && lb + 2 * h < nB ? 0xFFFF : 0),
                                     (unsigned short)(ok && lb + 2 * h + 1 < nB ? 0xFFFF : 0)};
        }
    const bool prune = a.A.segmin && a.B.segmin;
    const uint32_t *sma = prune ? a.A.segmin + (size_t)ta * a.nseg * 8 + i : nullptr;
    const uint32_t *smb = prune ? a.B.segmin + (size_t)tb * a.nseg * 8 + j : nullptr;
    uint32_t seen = 0;
    for (uint32_t w = a.words; w-- > 0;) {
        unsigned long long m = ra[w] & rb[w];
        while (m) {
            const uint32_t bit = 63u - (uint32_t)__builtin_clzll(m);
            m &= ~(1ull << bit);
            const uint32_t s = w * 64 + bit;
            if (prune) {  // as meet_join_kernel, a cell of 0xFFFF being INF
                const uint32_t bound = sat_add(sma[(size_t)s * 8], smb[(size_t)s * 8]);
                u16x2 hi2 = acc[0];
#pragma unroll
                for (int q = 1; q < 32; ++q) hi2 = __builtin_elementwise_max(hi2, acc[q]);
                const uint32_t hi = max((uint32_t)hi2.x, (uint32_t)hi2.y);
                const uint32_t top = hi == 0xFFFFu ? INF : hi;
                if (__ballot(bound < top && bound <= a.thr) == 0) continue;
            }
            const uint32_t qa = qa_of[s], qb = qb_of[s];
            if (qa == NONE || qb == NONE) continue;  // not reached by one of the wave groups: INF
            if (a.A.lmin16 && a.B.lmin16) {
                // per cell: the sum of its two lanes' least labels in the
                // segment (16-bit, saturating) against the cell; the segment
                // can lower no cell where no sum is below its cell
                const uint4 ma = ld16(a.A.lmin16 + (size_t)qa * 64 + 8 * i, 0);
                const uint4 mb = ld16(a.B.lmin16 + (size_t)qb * 64 + 8 * j, 0);
                const uint32_t MA[4] = {ma.x, ma.y, ma.z, ma.w}, MB[4] = {mb.x, mb.y, mb.z, mb.w};
                bool need = false;
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const u16x2 p = as_u16x2(MA[r / 2]);
                    const u16x2 pr = (r & 1) ? p.yy : p.xx;
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        const u16x2 lo = __builtin_elementwise_min(__builtin_elementwise_add_sat(pr, as_u16x2(MB[h])),
                                                                   acc[4 * r + h]);
                        need |= __builtin_bit_cast(uint32_t, lo) != __builtin_bit_cast(uint32_t, acc[4 * r + h]);
                    }
                }
                if (__ballot(need) == 0) continue;
            }
            const uint32_t len = a.seg_pos[s + 1] - a.seg_pos[s];
            seen += len;
            const uint16_t *pa = a.A.lab16 + (size_t)qa * 4096 + 8 * i;
            const uint16_t *pb = a.B.lab16 + (size_t)qb * 4096 + 8 * j;
            // two positions a step (an odd last one pairs with itself), the
            // next step's labels loaded before this step's updates and waited
            // for after them (a read past the last position re-reads it)
            uint4 na0 = ld16(pa, 0), nb0 = ld16(pb, 0);
            uint4 na1 = ld16(pa, min(1u, len - 1)), nb1 = ld16(pb, min(1u, len - 1));
            for (uint32_t x = 0; x < len; x += 2) {
                const uint4 a0 = na0, b0 = nb0, a1 = na1, b1 = nb1;
                na0 = ld16(pa, min(x + 2, len - 1));
                nb0 = ld16(pb, min(x + 2, len - 1));
                na1 = ld16(pa, min(x + 3, len - 1));
                nb1 = ld16(pb, min(x + 3, len - 1));
                const uint32_t A0[4] = {a0.x, a0.y, a0.z, a0.w}, A1[4] = {a1.x, a1.y, a1.z, a1.w};
                const uint32_t B0[4] = {b0.x, b0.y, b0.z, b0.w}, B1[4] = {b1.x, b1.y, b1.z, b1.w};
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const u16x2 p = as_u16x2(A0[r / 2]), q = as_u16x2(A1[r / 2]);
                    const u16x2 pr = (r & 1) ? p.yy : p.xx, qr = (r & 1) ? q.yy : q.xx;
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        const u16x2 s0 = __builtin_elementwise_add_sat(pr, as_u16x2(B0[h]));
                        const u16x2 s1 = __builtin_elementwise_add_sat(qr, as_u16x2(B1[h]));
                        acc[4 * r + h] = __builtin_elementwise_min(acc[4 * r + h], __builtin_elementwise_min(s0, s1));
                    }
                }
                asm volatile("" : "+v"(na0.x), "+v"(na0.y), "+v"(na0.z), "+v"(na0.w), "+v"(nb0.x), "+v"(nb0.y),
                             "+v"(nb0.z), "+v"(nb0.w));
                asm volatile("" : "+v"(na1.x), "+v"(na1.y), "+v"(na1.z), "+v"(na1.w), "+v"(nb1.x), "+v"(nb1.y),
                             "+v"(nb1.z), "+v"(nb1.w));
            }
        }
    }
    if (a.visits && threadIdx.x == 0 && seen) atomicAdd(a.visits, (unsigned long long)seen);
    uint32_t col[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) col[c] = lb + c < nB ? a.B.index[lb + c] : 0u;
    bool redo = false;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        if (la + r >= nA) break;
        uint32_t *o = a.out + (uint64_t)a.A.index[la + r] * a.ld;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (lb + c >= nB) continue;
            const uint32_t v16 = (c & 1) ? acc[4 * r + c / 2].y : acc[4 * r + c / 2].x;
            const bool capped = v16 >= CAP16 && v16 != 0xFFFFu;
            redo |= capped && a.thr >= CAP16;
            o[col[c]] = capped || v16 == 0xFFFFu || v16 > a.thr ? INF : v16;
        }
    }
    if (__ballot(redo) != 0 && threadIdx.x == 0) {
        a.redo[t] = 1;
        if (a.visits) atomicAdd(a.visits + 1, 1ull);
    }
}

// ---------------------------------------------------------------------------
// bottom pairs: one thread per pair of a source and a target in one bottom
// tree.  Both walk up to their lowest common ancestor; every common ancestor
// from there to the tree's root (the last bottom node) is a meeting node whose
// labels are the lanes' path slots.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void meet_bottom_kernel(const uint32_t *__restrict__ pairs, uint32_t n_pairs,
                                                          PathSide A, PathSide B, uint32_t d1,
                                                          const uint32_t *__restrict__ pos_parent,
                                                          const uint32_t *__restrict__ pos_depth, uint32_t n_top,
                                                          uint32_t thr, uint32_t *__restrict__ out, uint64_t ld) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_pairs) return;
    const uint32_t la = pairs[2 * k], lb = pairs[2 * k + 1];
    const uint32_t ja = A.seed_off[la], jb = B.seed_off[lb];
    const uint32_t pa = A.seed_pos[ja], pb = B.seed_pos[jb];
    const uint32_t da = pos_depth[pa], db = pos_depth[pb];
    uint32_t x = pa, y = pb, dx = da, dy = db;
    while (dx > dy && x != NONE) { x = pos_parent[x]; --dx; }
    while (dy > dx && y != NONE) { y = pos_parent[y]; --dy; }
    while (x != y && x != NONE && y != NONE) { x = pos_parent[x]; y = pos_parent[y]; }
    uint32_t best = INF;
    const uint32_t *sa = A.path + (size_t)ja * d1, *sb = B.path + (size_t)jb * d1;
    while (x != NONE && x >= n_top && x == y) {
        const uint32_t d = pos_depth[x];
        if (da - d >= d1 || db - d >= d1) break;  // deeper than a path holds: not a bottom ancestor
        best = min(best, sat_add(sa[da - d], sb[db - d]));
        x = pos_parent[x];
        y = x;
    }
    if (best <= thr) atomicMin(out + (uint64_t)A.index[la] * ld + B.index[lb], best);
}

__global__ __launch_bounds__(256) void meet_fill_kernel(uint32_t *__restrict__ p, uint64_t count, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256) p[i] = v;
}

}  // namespace

hipError_t launch_fill(hipStream_t st, uint32_t *p, uint64_t count, uint32_t value) {
    if (!count) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>((count + 255) / 256, 16384);
    hipLaunchKernelGGL(meet_fill_kernel, dim3((uint32_t)blocks), dim3(256), 0, st, p, count, value);
    return hipGetLastError();
}

hipError_t launch_reach(hipStream_t st, const uint32_t *seed_off, const uint32_t *seed_pre, uint32_t n_lanes,
                        const uint32_t *pos_pre, const uint32_t *pos_sub, const uint32_t *seg_pos, uint32_t nseg,
                        uint32_t words, unsigned long long *reach) {
    const uint32_t wgs = (n_lanes + 63) / 64;
    if (!wgs || !words) return hipSuccess;
    hipLaunchKernelGGL(meet_reach_kernel, dim3(words, wgs), dim3(64), 0, st, seed_off, seed_pre, n_lanes, pos_pre,
                       pos_sub, seg_pos, nseg, words, reach);
    return hipGetLastError();
}

hipError_t launch_join(hipStream_t st, const Side &A, const Side &B, const uint32_t *seg_pos, uint32_t words,
                       uint32_t nseg, uint32_t thr, uint32_t splits, uint32_t *out, uint64_t ld,
                       unsigned long long *visits, const uint32_t *only) {
    if ((A.segmin || B.segmin) && (uint64_t)words * 64 < nseg) return hipErrorInvalidValue;
    JoinArgs a{A, B, seg_pos, words, nseg, (A.n_lanes + 63) / 64, (B.n_lanes + 63) / 64, thr, splits ? splits : 1,
               out, ld, visits, only, nullptr};
    const uint64_t tiles = (uint64_t)a.wga * a.wgb * a.splits;
    if (!tiles) return hipSuccess;
    const uint64_t blocks = (tiles + XCDS - 1) / XCDS * XCDS;
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(meet_join_kernel, dim3((uint32_t)blocks), dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_join16(hipStream_t st, const Side &A, const Side &B, const uint32_t *seg_pos, uint32_t words,
                         uint32_t nseg, uint32_t thr, uint32_t *out, uint64_t ld, unsigned long long *visits,
                         uint32_t *redo) {
    if (!A.lab16 || !B.lab16 || !A.slot || !B.slot || A.kl != 64 || B.kl != 64 || !redo)
        return hipErrorInvalidValue;
    if ((A.segmin || B.segmin) && (uint64_t)words * 64 < nseg) return hipErrorInvalidValue;
    JoinArgs a{A, B, seg_pos, words, nseg, (A.n_lanes + 63) / 64, (B.n_lanes + 63) / 64, thr, 1,
               out, ld, visits, nullptr, redo};
    const uint64_t tiles = (uint64_t)a.wga * a.wgb;
    if (!tiles) return hipSuccess;
    const uint64_t blocks = (tiles + XCDS - 1) / XCDS * XCDS;
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(meet_join16_kernel, dim3((uint32_t)blocks), dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_bottom_pairs(hipStream_t st, const uint32_t *pairs, uint32_t n_pairs, const PathSide &A,
                               const PathSide &B, uint32_t d1, const uint32_t *pos_parent,
                               const uint32_t *pos_depth, uint32_t n_top, uint32_t thr, uint32_t *out, uint64_t ld) {
    if (!n_pairs) return hipSuccess;
    hipLaunchKernelGGL(meet_bottom_kernel, dim3((n_pairs + 255) / 256), dim3(256), 0, st, pairs, n_pairs, A, B, d1,
                       pos_parent, pos_depth, n_top, thr, out, ld);
    return hipGetLastError();
}

}  // namespace bfg_meet
